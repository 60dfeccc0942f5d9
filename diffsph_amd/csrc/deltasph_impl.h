// Internal interface of the delta-SPH kernels shared with the batched step
// driver (batched.hip): the right-hand side's argument block and launchers
// that resolve the periodic box per particle.
#pragma once

#include "common.h"

namespace dsph {

struct RhsArgs {
  const float2* vel;
  const float* mass;
  const float* rho;
  const float2* grad_rho;
  float2* dudt;
  float* drhodt;
  float* terms;  // optional [n, 8]: a_p.xy, visc.xy, continuity, diffusion, pressure, 0
  float rho0;
  float eos_c;     // cs^2 (isothermal) or kappa (Tait)
  float delta, kscale, cs;
  float visc_c;    // viscid: 2 (d+2) nu rho0; inviscid: alpha
  float eps;
  float gx, gy;
  int visc_scheme;  // -1 none, 0 inviscid, 1 viscid
  int pi_switch;
  int diffusion;
};

// Validates the parameters and fills the argument block (pointers included).
int make_rhs_args(const dsph_deltasph_params* prm, const float* vel, const float* mass, const float* rho,
                  const float* grad_rho, float* dudt, float* drhodt, float* terms, RhsArgs* out);

// Fluid-only normalization / density gradient and right-hand side of a
// particle set whose box is doms[dom_of[i]] (P.dom is ignored), launched
// with `block` threads per workgroup (64, 128 or 256).
int launch_normalization_per_particle(const PairsDev& P, const Dom* doms, const int32_t* dom_of,
                                      const float* mass, const float* rho, const float* vol, float2* grad_rho,
                                      int block, hipStream_t st);
// vol[i] = mass[i] / rho[i]: the particle volume the normalization pass
// reads as V_j instead of forming the same quotient once per pair.
int launch_volume(int n, const float* mass, const float* rho, float* vol, hipStream_t st);
int launch_rhs_per_particle(const PairsDev& P, const RhsArgs& A, const Dom* doms, const int32_t* dom_of,
                            int block, hipStream_t st);

}  // namespace dsph
