// delta-SPH right-hand side, RK4 stage arithmetic and the time-step reduction.
//
// One right-hand side (v2/simulationSchemes/deltaPlus.py:176-422, fluid only)
// is two pair kernels over the neighbor list:
//
//   k_dsph_normalization  per particle i, two passes over its row:
//       M_i = sum_j V_j (x_j - x_i) (x) grad W_ij      (normalizationMatrices.py:18-23)
//       M_i = I if the row has fewer than 4 entries     (deltaPlus.py:245)
//       L_i = pinv(M_i)                                 (math.py:10-57)
//       grad^L rho_i = sum_j V_j (rho_i - rho_j) L_i grad W_ij   (densityDiffusion.py:7-26)
//   k_dsph_rhs            per particle i, one pass:
//       density diffusion  D_i   (densityDiffusion.py:49-73, scheme deltaSPH)
//       viscosity          (viscosity.py:123-170 inviscid, :200-218 viscid)
//       continuity         (momentumEquation.py:6-39)
//       EOS + pressure     (pressureEOS.py:28-64, pressureForce.py:10-12)
//       dudt = a_p + g + visc, drho/dt = continuity + D   (deltaPlus.py:349-350)
//
// Pair geometry (d = x_i - x_j min-image, r, xhat, q, W, grad W) is evaluated
// on the fly with the bit-reproducible helpers of common.h; V_j = m_j / rho_j
// per pair.  Every expression is written in the order of the reference's
// tensor expressions, and the CPU oracle evaluates the same expressions.
#include "common.h"
#include "deltasph_impl.h"
#include "mdbc_op.h"
#include "pmath.h"

namespace dsph {

// The wall particles seen by the fluid rows: pair list with rows = fluid
// particles (same order as the fluid list) and entries = boundary indices,
// rpos = boundary positions; the boundary's mass, mDBC density, velocity
// and its viscosity parameters (boundaryDiffusion, boundaryViscosity.py).
struct BoundaryDev {
  PairsDev P;
  const float* mass;
  const float* rho;
  const float2* vel;
  float visc_coeff, visc_eps;
  float visc_c_viscid;  // 2 (d+2) nu_b rho0 for the viscid wall term
  int viscosity;
  int pi_switch;
  int on;
};

inline BoundaryDev make_boundary(const dsph_boundary* b) {
  BoundaryDev B;
  B.on = b != nullptr;
  if (!b) {
    B.mass = B.rho = nullptr;
    B.vel = nullptr;
    B.visc_coeff = B.visc_eps = B.visc_c_viscid = 0.f;
    B.viscosity = -1;
    B.pi_switch = 0;
    return B;
  }
  B.P = make_pairs(*b->pairs);
  B.mass = b->mass;
  B.rho = b->rho;
  B.vel = reinterpret_cast<const float2*>(b->vel);
  B.visc_coeff = b->visc_coeff;
  B.visc_eps = b->visc_eps;
  B.visc_c_viscid = 0.f;
  B.viscosity = b->viscosity;
  B.pi_switch = b->pi_switch;
  return B;
}

// Gathered neighbor data of the pair loops.
struct NbrXMR {
  float2 x;
  float m, r;
};
struct NbrRhs {
  float2 x, u, g;
  float r, m;
};

// The periodic box a particle lives in: one for a single domain, or one per
// particle (its domain's) for a batch of independent domains.
struct OneDomain {
  Dom d;
  __device__ __forceinline__ Dom operator()(int) const { return d; }
};
struct DomainPerParticle {
  const Dom* doms;
  const int32_t* dom_of;
  __device__ __forceinline__ Dom operator()(int i) const { return doms[dom_of[i]]; }
};

// With a boundary: M_i = M_ff + M_bf ('fluidwBoundary', deltaPlus.py:241),
// the identity test uses the fluid row length (the fluid numNeighbors), and
// the wall term of the density gradient is the plain 'difference' gradient
// sum_b V_b (rho_i - rho_b) gradW_ib (densityGradient, deltaPlus.py:262-263).
// One pass over the row: L_i does not depend on j, so the renormalized
// gradient is L_i applied once to s_i = sum_j V_j (rho_i - rho_j) gradW_ij.
template <class DF, bool kVol>
__global__ void __launch_bounds__(kBlock) k_dsph_normalization(PairsDev P, BoundaryDev B, DF domain,
                                                               const float* __restrict__ mass,
                                                               const float* __restrict__ rho,
                                                               const float* __restrict__ vol,
                                                               float4* __restrict__ L_out,
                                                               float2* __restrict__ ev_out,
                                                               float2* __restrict__ grad_rho) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const Dom dm = domain(i);
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const KSel ks = ksel(P);
  const RowRef rr = row_ref(P, i);
  const float rho_i = rho[i];
  float m00 = 0.f, m01 = 0.f, m10 = 0.f, m11 = 0.f, sx = 0.f, sy = 0.f;
  // with kVol the record carries V_j (= m_j / rho_j, the same quotient
  // formed once per particle) in place of m_j
  auto ld = [&](int j) {
    NbrXMR d;
    d.x = P.rpos[j];
    d.m = kVol ? vol[j] : mass[j];
    d.r = rho[j];
    return d;
  };
  scan_row<4, true>(P.nbr, rr, ld, [&](int j, const NbrXMR& d) {
    const float h = pick_h(P.mode, hi, h_of(P.rh, j, P.h0));
    const Geom g = pair_geom(xi, d.x, h, ks(h), dm);
    const float Vj = kVol ? d.m : d.m / d.r;
    const float s = -(g.q * g.h);  // -(distances * supports)
    const float qx = s * g.ex, qy = s * g.ey;
    const float kx = Vj * g.gx, ky = Vj * g.gy;
    m00 += qx * kx;
    m01 += qx * ky;
    m10 += qy * kx;
    m11 += qy * ky;
    const float dr = rho_i - d.r;
    sx += dr * kx;
    sy += dr * ky;
  });
  float bx = 0.f, by = 0.f;
  if (B.on) {
    const RowRef br = row_ref(B.P, i);
    float b00 = 0.f, b01 = 0.f, b10 = 0.f, b11 = 0.f;
    for (int n = 0, k = br.base; n < br.cnt; ++n, k += br.step) {
      const int j = B.P.nbr[k];
      const float h = pick_h(B.P.mode, hi, h_of(B.P.rh, j, B.P.h0));
      const Geom g = pair_geom(xi, B.P.rpos[j], h, B.P.dom);
      const float Vj = B.mass[j] / B.rho[j];
      const float s = -(g.q * g.h);
      const float qx = s * g.ex, qy = s * g.ey;
      const float kx = Vj * g.gx, ky = Vj * g.gy;
      b00 += qx * kx;
      b01 += qx * ky;
      b10 += qy * kx;
      b11 += qy * ky;
      const float dr = rho_i - B.rho[j];
      bx += dr * kx;
      by += dr * ky;
    }
    m00 = m00 + b00;
    m01 = m01 + b01;
    m10 = m10 + b10;
    m11 = m11 + b11;
  }
  if (rr.cnt < 4) {
    m00 = 1.f;
    m01 = 0.f;
    m10 = 0.f;
    m11 = 1.f;
  }
  const dsph_pm::Pinv2 Lm = dsph_pm::pinv2x2(m00, m01, m10, m11);
  if (L_out) L_out[i] = make_float4(Lm.l00, Lm.l01, Lm.l10, Lm.l11);
  if (ev_out) ev_out[i] = make_float2(Lm.ev0, Lm.ev1);
  float ax = Lm.l00 * sx + Lm.l01 * sy;
  float ay = Lm.l10 * sx + Lm.l11 * sy;
  if (B.on) {
    ax = ax + bx;
    ay = ay + by;
  }
  grad_rho[i] = make_float2(ax, ay);
}


__device__ __forceinline__ float eos_p(float rho, float rho0, float c) { return c * (rho - rho0); }

// With a boundary (deltaPlus.py:281-329, direction 'all' / 'boundaryToFluid'
// for the fluid rows): continuity and pressure add the wall pairs as a
// second sum (ftf + btf), the viscosity adds computeViscosity_Boundary
// (boundaryViscosity.py:123-170 inviscid with V_b = m_b/(rho_i + rho_b),
// :198-218 viscid with V_b = m_b/rho_b) with its own coefficients.
template <class DF>
__global__ void __launch_bounds__(kBlock) k_dsph_rhs(PairsDev P, BoundaryDev B, RhsArgs A, DF domain) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const Dom dm = domain(i);
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const float2 ui = A.vel[i];
  const float rho_i = A.rho[i];
  const float2 gri = A.diffusion ? A.grad_rho[i] : make_float2(0.f, 0.f);
  const float p_i = eos_p(rho_i, A.rho0, A.eos_c);
  const KSel ks = ksel(P);
  float accD = 0.f, accVx = 0.f, accVy = 0.f, accC = 0.f, accPx = 0.f, accPy = 0.f;
  const RowRef rr = row_ref(P, i);
  auto ld = [&](int j) {
    NbrRhs d;
    d.x = P.rpos[j];
    d.u = A.vel[j];
    d.r = A.rho[j];
    d.m = A.mass[j];
    d.g = A.diffusion ? A.grad_rho[j] : make_float2(0.f, 0.f);
    return d;
  };
  scan_row<4, true>(P.nbr, rr, ld, [&](int j, const NbrRhs& e) {
    const float h = pick_h(P.mode, hi, h_of(P.rh, j, P.h0));
    const Geom g = pair_geom(xi, e.x, h, ks(h), dm);
    const float2 uj = e.u;
    const float rho_j = e.r;
    const float m_j = e.m;
    const float Vj = m_j / rho_j;
    const float kx = Vj * g.gx, ky = Vj * g.gy;
    const float rij = g.q * g.h;  // distances * supports
    if (A.diffusion) {
      const float2 grj = e.g;
      const float r_ij = (2.0f * (rho_j - rho_i)) / (rij + 1e-6f * g.h);
      const float psx = (-r_ij) * g.ex + (gri.x + grj.x);
      const float psy = (-r_ij) * g.ey + (gri.y + grj.y);
      accD += psx * kx + psy * ky;
    }
    if (A.visc_scheme >= 0) {
      const float vx = ui.x - uj.x, vy = ui.y - uj.y;
      const float vr = vx * g.ex + vy * g.ey;
      float pi = vr / (rij + A.eps * (g.h * g.h));
      if (A.pi_switch) pi = vr < 0.0f ? pi : 0.0f;
      const float vs = Vj * pi;
      accVx += vs * g.gx;
      accVy += vs * g.gy;
    }
    {
      const float qx = (uj.x - ui.x) * m_j, qy = (uj.y - ui.y) * m_j;
      float kq = qx * g.gx + qy * g.gy;
      kq = (kq * rho_i) / rho_j;
      accC += kq;
    }
    {
      const float ps = p_i + eos_p(rho_j, A.rho0, A.eos_c);
      accPx += ps * kx;
      accPy += ps * ky;
    }
  });
  float apx = (-accPx) / rho_i, apy = (-accPy) / rho_i;
  float vsx = 0.f, vsy = 0.f;
  if (A.visc_scheme == 1) {
    const float s = A.visc_c / rho_i;
    vsx = s * accVx;
    vsy = s * accVy;
  } else if (A.visc_scheme == 0) {
    const float s = (((A.visc_c * hi) * A.cs) * A.rho0) / rho_i;
    vsx = s * accVx;
    vsy = s * accVy;
  }
  float cont = -accC;
  if (B.on) {
    float bVx = 0.f, bVy = 0.f, bC = 0.f, bPx = 0.f, bPy = 0.f;
    const RowRef br = row_ref(B.P, i);
    for (int n = 0, k = br.base; n < br.cnt; ++n, k += br.step) {
      const int j = B.P.nbr[k];
      const float h = pick_h(B.P.mode, hi, h_of(B.P.rh, j, B.P.h0));
      const Geom g = pair_geom(xi, B.P.rpos[j], h, B.P.dom);
      const float2 uj = B.vel[j];
      const float rho_j = B.rho[j];
      const float m_j = B.mass[j];
      const float rij = g.q * g.h;
      if (B.viscosity >= 0) {
        const float vx = ui.x - uj.x, vy = ui.y - uj.y;
        const float vr = vx * g.ex + vy * g.ey;
        float pi = vr / (rij + B.visc_eps * (g.h * g.h));
        if (B.pi_switch) pi = vr < 0.0f ? pi : 0.0f;
        const float Vb = B.viscosity == 0 ? m_j / (rho_i + rho_j) : m_j / rho_j;
        const float vs = Vb * pi;
        bVx += vs * g.gx;
        bVy += vs * g.gy;
      }
      {
        const float qx = (uj.x - ui.x) * m_j, qy = (uj.y - ui.y) * m_j;
        float kq = qx * g.gx + qy * g.gy;
        kq = (kq * rho_i) / rho_j;
        bC += kq;
      }
      {
        const float Vj = m_j / rho_j;
        const float ps = p_i + eos_p(rho_j, A.rho0, A.eos_c);
        bPx += ps * (Vj * g.gx);
        bPy += ps * (Vj * g.gy);
      }
    }
    apx = apx + (-bPx) / rho_i;
    apy = apy + (-bPy) / rho_i;
    cont = cont + (-bC);
    if (B.viscosity == 1) {
      const float s = B.visc_c_viscid / rho_i;
      vsx = vsx + s * bVx;
      vsy = vsy + s * bVy;
    } else if (B.viscosity == 0) {
      const float s = (((B.visc_coeff * hi) * A.cs) * A.rho0) / rho_i;
      vsx = vsx + s * bVx;
      vsy = vsy + s * bVy;
    }
  }
  const float D = A.diffusion ? (((A.delta * hi) / A.kscale) * A.cs) * accD : 0.0f;
  A.dudt[i] = make_float2((apx + A.gx) + vsx, (apy + A.gy) + vsy);
  A.drhodt[i] = cont + D;
  if (A.terms) {
    float* t = A.terms + (int64_t)i * 8;
    t[0] = apx;
    t[1] = apy;
    t[2] = vsx;
    t[3] = vsy;
    t[4] = cont;
    t[5] = D;
    t[6] = p_i;
    t[7] = 0.0f;
  }
}

// out = base + k * c for the three state fields (any may be absent).
__global__ void __launch_bounds__(kBlock) k_rk_update(int n, const float* __restrict__ x0, const float* __restrict__ kx,
                                                      float* __restrict__ x, const float* __restrict__ u0,
                                                      const float* __restrict__ ku, float* __restrict__ u,
                                                      const float* __restrict__ r0, const float* __restrict__ kr,
                                                      float* __restrict__ r, float c,
                                                      const float* __restrict__ dt_dev, int half) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (dt_dev) {
    // c = fl32(dt * 0.5) or dt, as the host forms it from the float32 step
    const float t = dt_dev[0];
    c = half ? t * 0.5f : t;
  }
  if (x) {
    x[2 * i] = x0[2 * i] + kx[2 * i] * c;
    x[2 * i + 1] = x0[2 * i + 1] + kx[2 * i + 1] * c;
  }
  if (u) {
    u[2 * i] = u0[2 * i] + ku[2 * i] * c;
    u[2 * i + 1] = u0[2 * i + 1] + ku[2 * i + 1] * c;
  }
  if (r) r[i] = r0[i] + kr[i] * c;
}

// (k1 + 2 k2 + 2 k3 + k4) / 6, then state += that * dt (integration.py:286-307).
__device__ __forceinline__ float rk4_avg(float a, float b, float c, float d) {
  return (((a + 2.0f * b) + 2.0f * c) + d) / 6.0f;
}

struct Rk4Args {
  const float* k[3][4];  // [field][stage]: dxdt (2n), dudt (2n), drhodt (n)
  float* state[3];       // x, u, rho (in/out)
  float* avg[3];         // optional averaged derivatives
  float dt;
  const float* dt_dev;   // when set, the step is read from here instead of dt
  int n;
};

__global__ void __launch_bounds__(kBlock) k_rk4_combine(Rk4Args A) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const float dt = A.dt_dev ? A.dt_dev[0] : A.dt;
  for (int f = 0; f < 3; ++f) {
    if (!A.state[f]) continue;
    const int w = f == 2 ? 1 : 2;
    for (int c = 0; c < w; ++c) {
      const int64_t o = (int64_t)i * w + c;
      const float d = rk4_avg(A.k[f][0][o], A.k[f][1][o], A.k[f][2][o], A.k[f][3][o]);
      A.state[f][o] = A.state[f][o] + d * dt;
      if (A.avg[f]) A.avg[f][o] = d;
    }
  }
}

// max_i |v_i| (float32 norm with separately rounded squares): per-block
// maxima, then one block.  The maximum does not depend on the order.
__global__ void __launch_bounds__(kBlock) k_max_norm(const float2* __restrict__ v, int n, float* __restrict__ part) {
  __shared__ float sh[kBlock / 64];
  float m = 0.0f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float2 a = v[i];
    const float s = sqrt_rn(dist2_exact(a.x, a.y));
    m = s > m ? s : m;  // NaN entries are skipped (isnan-filtered in the reference)
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float t = __shfl_down(m, o, 64);
    m = t > m ? t : m;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = 0.0f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b = sh[w] > b ? sh[w] : b;
    part[blockIdx.x] = b;
  }
}

// One wave: lane l folds partials l, l+64, ... then a shuffle tree (the
// maximum does not depend on the order).  Launched as <<<1, 64>>>.
__global__ void __launch_bounds__(64) k_max_final(const float* part, int nb, float* out) {
  float m = 0.0f;
#pragma unroll 4
  for (int b = threadIdx.x; b < nb; b += 64) m = part[b] > m ? part[b] : m;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float t = __shfl_down(m, o, 64);
    m = t > m ? t : m;
  }
  if (threadIdx.x == 0) *out = m;
}

// sum of squares of n floats in float64 (NaN entries skipped): per-block
// partials in a fixed tree order, then one block, so repeated runs give the
// same bits.
__global__ void __launch_bounds__(kBlock) k_sumsq(const float* __restrict__ v, int64_t n, double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double a = (double)v[i];
    if (a == a) s += a * a;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b += sh[w];
    part[blockIdx.x] = b;
  }
}

// One wave, a fixed order: lane l sums partials l, l+64, ... in turn, then
// a shuffle tree.  Launched as <<<1, 64>>>.
__global__ void __launch_bounds__(64) k_sum_final(const double* part, int nb, double* out) {
  double s = 0.0;
#pragma unroll 4
  for (int b = threadIdx.x; b < nb; b += 64) s += part[b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (threadIdx.x == 0) *out = s;
}

// computeTimestep (v2/modules/timestep.py:8-39) on the device: the float32
// formula of the host version from the acceleration norm left by the
// reduction (global: the float32 rounding of sqrt of the float64 sum of
// squares; max: the largest per-particle norm); time += the step used.
__global__ void k_timestep_single(const double* __restrict__ sumsq, const float* __restrict__ maxn,
                                  dsph_timestep_params tp, float* __restrict__ dt, double* __restrict__ time) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const float used = dt[0];
  if (time) time[0] = time[0] + (double)used;
  if (!tp.adaptive) return;
  const float amax = sumsq ? (float)sqrt(sumsq[0]) : maxn[0];
  const float dt_a = (0.25f * sqrtf(tp.support / (amax + 1e-7f))) / tp.kernel_scale;
  float nw = used;
  if (tp.viscosity_constraint) nw = tp.dt_visc;
  if (tp.acceleration_constraint) nw = dt_a < nw ? dt_a : nw;
  if (tp.acoustic_constraint) nw = tp.dt_acoustic < nw ? tp.dt_acoustic : nw;
  nw = tp.max_dt < nw ? tp.max_dt : nw;
  nw = tp.min_dt > nw ? tp.min_dt : nw;
  dt[0] = nw;
}

// Ghost positions of wall particles (buildBoundaryGhostParticles,
// v2/modules/mDBC.py:9-17): x_g = x_b + (d + max(d, -h)) n with the signed
// distance d (negative inside the wall) and the normal n pointing into it.
__global__ void __launch_bounds__(kBlock) k_ghosts(const float2* __restrict__ bpos, const float* __restrict__ dist,
                                                   const float2* __restrict__ nrm, int nb, float h,
                                                   float2* __restrict__ gpos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const float d = dist[i];
  const float md = d + (d > -h ? d : -h);
  const float2 x = bpos[i], nn = nrm[i];
  gpos[i] = make_float2(x.x + md * nn.x, x.y + md * nn.y);
}

// The reference's C++ mDBC op (mDBCDensityKernel, v2/cppSrc/mDBC.h:129-244):
// x_gj = -(q h) xhat, b = sum m_j [W, gradW], Cramer's rule with
// det + 1e-7, Shepard density for <= 5 neighbors, clamped at rho0.
__device__ float mdbc_op_row(const PairsDev& G, int i, const float* __restrict__ mass,
                             const float* __restrict__ rho, const float2* __restrict__ bpos, float rho0,
                             float* __restrict__ shepard) {
  const float2 xg = G.qpos[i];
  const float hi = h_of(G.qh, i, G.h0);
  MdbcOpAcc acc;
  const RowRef rr = row_ref(G, i);
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    const int j = G.nbr[k];
    const float h = pick_h(G.mode, hi, h_of(G.rh, j, G.h0));
    const Geom g = pair_geom(xg, G.rpos[j], h, G.dom);
    acc.add(mass[j], rho[j], g.w, g.gx, g.gy, ((-g.q) * g.h) * g.ex, ((-g.q) * g.h) * g.ey);
  }
  const float2 xb = bpos[i];
  float shep;
  const float out = acc.finish(rr.cnt, xb.x - xg.x, xb.y - xg.y, rho0, &shep);
  if (shepard) shepard[i] = shep;
  return out;
}

// mDBC wall density (mDBCDensity, v2/modules/mDBC.py:50-136): per ghost g
// over its fluid neighbors j (V_j = m_j / rho_j, x_gj = x_j - x_g):
//   Shepard  rho_S = sum m_j W / (sum V_j W + 1e-7)
//   A = [[sum V W, sum V W x_gj^T], [sum V gradW, sum V x_gj (x) gradW]^T-layout as the
//        reference fills A_g], b = [sum m_j W, sum m_j gradW]
//   rho_b = rho0 (no neighbors), rho_S (1..5 neighbors), else
//   c = pinv(A) b (pinv only for > 4 neighbors) and rho_b = c0 + (x_b - x_g) . c[1:].
// variant 1 is the C++ op above.
__global__ void __launch_bounds__(kBlock) k_mdbc(PairsDev G, const float* __restrict__ mass,
                                                 const float* __restrict__ rho, const float2* __restrict__ bpos,
                                                 float rho0, float support, int variant, float* __restrict__ rho_b,
                                                 float* __restrict__ shepard) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nq) return;
  if (variant == 1) {
    rho_b[i] = mdbc_op_row(G, i, mass, rho, bpos, rho0, shepard);
    return;
  }
  const float2 xg = G.qpos[i];
  const float hi = h_of(G.qh, i, G.h0);
  float nom = 0.f, vol = 0.f, gsx = 0.f, gsy = 0.f, vgx = 0.f, vgy = 0.f, psx = 0.f, psy = 0.f;
  float p00 = 0.f, p01 = 0.f, p10 = 0.f, p11 = 0.f;
  const RowRef rr = row_ref(G, i);
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    const int j = G.nbr[k];
    const float h = pick_h(G.mode, hi, h_of(G.rh, j, G.h0));
    const Geom g = pair_geom(xg, G.rpos[j], h, G.dom);
    const float m_j = mass[j], rho_j = rho[j];
    const float Vj = m_j / rho_j;
    nom += m_j * g.w;
    const float kw = Vj * g.w;
    vol += 1.0f * kw;
    const float kx = Vj * g.gx, ky = Vj * g.gy;
    gsx += rho_j * kx;
    gsy += rho_j * ky;
    vgx += 1.0f * kx;
    vgy += 1.0f * ky;
    const float xx = ((-g.ex) * g.q) * support, xy = ((-g.ey) * g.q) * support;
    psx += xx * kw;
    psy += xy * kw;
    p00 += xx * kx;
    p01 += xx * ky;
    p10 += xy * kx;
    p11 += xy * ky;
  }
  const float shep = nom / (vol + 1e-7f);
  float out = rho0;
  if (rr.cnt > 0) out = shep;
  if (rr.cnt > 5) {
    const double A[9] = {vol, psx, psy, vgx, p00, p01, vgy, p10, p11};
    double Ai[9];
    dsph_pm::pinv3x3(A, 3.0 * 1.1920928955078125e-07, Ai);
    const double b0 = nom, b1 = gsx, b2 = gsy;
    const float c0 = (float)(Ai[0] * b0 + Ai[1] * b1 + Ai[2] * b2);
    const float c1 = (float)(Ai[3] * b0 + Ai[4] * b1 + Ai[5] * b2);
    const float c2 = (float)(Ai[6] * b0 + Ai[7] * b1 + Ai[8] * b2);
    const float2 xb = bpos[i];
    const float rx = xb.x - xg.x, ry = xb.y - xg.y;
    out = c0 + (rx * c1 + ry * c2);
  }
  rho_b[i] = out;
  if (shepard) shepard[i] = shep;
}

static int check_same_set(const dsph_pairs* pairs) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  return DSPH_OK;
}

// ---------------------------------------------------------------------------
// delta+ particle shifting (shifting.py:119-147, applied as shifting.py:
// 997-1018 and runner.py:40-42).
// ---------------------------------------------------------------------------
struct ShiftArgs {
  const float* mass;
  const float* rho;
  const float* vmax;  // device max |v| (when mach < 0)
  float2* pos;
  float2* shift_out;
  float w0, R, nexp, cfl, kscale, thr, cs, mach;
  int nint;  // integral exponent in [0, 16], else -1
};

// (W / W_0)^n: a double product chain rounded once to float for integral n.
__device__ __forceinline__ float shift_pow(float k, const ShiftArgs& A) {
  if (A.nint >= 0) {
    double p = 1.0;
    for (int t = 0; t < A.nint; ++t) p *= (double)k;
    return (float)p;
  }
  return powf(k, A.nexp);
}

__device__ __forceinline__ void shift_term(const Geom& g, float m_j, float rho_i, float rho_j, const ShiftArgs& A,
                                           float& sx, float& sy) {
  const float kn = shift_pow(g.w / A.w0, A);
  const float term = 1.0f + A.R * kn;
  const float s = term * (m_j / (rho_i + rho_j));
  sx += s * g.gx;
  sy += s * g.gy;
}

__global__ void __launch_bounds__(kBlock) k_deltaplus_shift(PairsDev P, BoundaryDev B, ShiftArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const float rho_i = A.rho[i];
  float sx = 0.f, sy = 0.f;
  const RowRef rr = row_ref(P, i);
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    const int j = P.nbr[k];
    const Geom g = pair_geom(xi, P.rpos[j], pick_h(P.mode, hi, h_of(P.rh, j, P.h0)), P.dom);
    shift_term(g, A.mass[j], rho_i, A.rho[j], A, sx, sy);
  }
  if (B.on) {
    const RowRef rb = row_ref(B.P, i);
    for (int n = 0, k = rb.base; n < rb.cnt; ++n, k += rb.step) {
      const int b = B.P.nbr[k];
      const Geom g = pair_geom(xi, B.P.rpos[b], pick_h(B.P.mode, hi, h_of(B.P.rh, b, B.P.h0)), B.P.dom);
      shift_term(g, B.mass[b], rho_i, B.rho[b], A, sx, sy);
    }
  }
  const float ma = A.mach >= 0.0f ? A.mach : *A.vmax / A.cs;
  const float hk = (hi / A.kscale) * 2.0f;
  const float sc = (-A.cfl * ma) * (hk * hk);
  float ux = -(sc * sx), uy = -(sc * sy);
  // torch.clamp(update, -thr, thr): NaN stays NaN
  ux = ux < -A.thr ? -A.thr : (ux > A.thr ? A.thr : ux);
  uy = uy < -A.thr ? -A.thr : (uy > A.thr ? A.thr : uy);
  const float2 x = A.pos[i];
  // solveShifting returns dx = (x - u) - x; the runner adds it
  const float dx = (x.x - ux) - x.x, dy = (x.y - uy) - x.y;
  A.pos[i] = make_float2(x.x + dx, x.y + dy);
  if (A.shift_out) A.shift_out[i] = make_float2(dx, dy);
}

}  // namespace dsph

using namespace dsph;

static int check_boundary(const dsph_boundary* b, const dsph_pairs* fluid) {
  if (!b) return DSPH_OK;
  int s = check_pairs(b->pairs);
  if (s != DSPH_OK) return s;
  if (b->pairs->nq != fluid->nq) return DSPH_ERR_ARG;
  if (b->pairs->nr > 0 && (!b->mass || !b->rho || !b->vel)) return DSPH_ERR_ARG;
  if (b->viscosity < -1 || b->viscosity > 1) return DSPH_ERR_ARG;
  return DSPH_OK;
}

__global__ void __launch_bounds__(kBlock) k_volume(int n, const float* __restrict__ mass,
                                                   const float* __restrict__ rho, float* __restrict__ vol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) vol[i] = mass[i] / rho[i];
}

int dsph::launch_volume(int n, const float* mass, const float* rho, float* vol, hipStream_t st) {
  if (n <= 0) return DSPH_OK;
  k_volume<<<grid_for(n), kBlock, 0, st>>>(n, mass, rho, vol);
  return launch_status();
}

extern "C" int dsph_volume(int64_t n, const float* mass, const float* rho, float* vol, void* stream) {
  if (n < 0 || n >= INT32_MAX || (n > 0 && (!mass || !rho || !vol))) return DSPH_ERR_ARG;
  return launch_volume((int)n, mass, rho, vol, as_stream(stream));
}

extern "C" int dsph_deltasph_normalization_v(const dsph_pairs* pairs, const dsph_boundary* boundary,
                                             const float* mass, const float* rho, const float* vol, float* L,
                                             float* eigvals, float* grad_rho, void* stream) {
  int s = check_same_set(pairs);
  if (s != DSPH_OK) return s;
  s = check_boundary(boundary, pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq == 0) return DSPH_OK;
  if (!mass || !rho || !grad_rho) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  BoundaryDev B = make_boundary(boundary);
  float4* Lo = reinterpret_cast<float4*>(L);
  float2* eo = reinterpret_cast<float2*>(eigvals);
  float2* go = reinterpret_cast<float2*>(grad_rho);
  if (vol)
    k_dsph_normalization<OneDomain, true><<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(
        P, B, OneDomain{P.dom}, mass, rho, vol, Lo, eo, go);
  else
    k_dsph_normalization<OneDomain, false><<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(
        P, B, OneDomain{P.dom}, mass, rho, nullptr, Lo, eo, go);
  return launch_status();
}

extern "C" int dsph_deltasph_normalization(const dsph_pairs* pairs, const dsph_boundary* boundary, const float* mass,
                                           const float* rho, float* L, float* eigvals, float* grad_rho,
                                           void* stream) {
  return dsph_deltasph_normalization_v(pairs, boundary, mass, rho, nullptr, L, eigvals, grad_rho, stream);
}

int dsph::make_rhs_args(const dsph_deltasph_params* prm, const float* vel, const float* mass, const float* rho,
                        const float* grad_rho, float* dudt, float* drhodt, float* terms, RhsArgs* out) {
  if (!prm || !out) return DSPH_ERR_ARG;
  if (prm->diffusion && !grad_rho) return DSPH_ERR_ARG;
  if (prm->viscosity < -1 || prm->viscosity > 1 || prm->eos < 0 || prm->eos > 1) return DSPH_ERR_ARG;
  if (prm->diffusion && !(prm->kernel_scale > 0.0f)) return DSPH_ERR_ARG;
  RhsArgs& A = *out;
  A.vel = reinterpret_cast<const float2*>(vel);
  A.mass = mass;
  A.rho = rho;
  A.grad_rho = reinterpret_cast<const float2*>(grad_rho);
  A.dudt = reinterpret_cast<float2*>(dudt);
  A.drhodt = drhodt;
  A.terms = terms;
  A.rho0 = prm->rho0;
  A.eos_c = prm->eos == 0 ? prm->cs * prm->cs : prm->eos_kappa;
  A.delta = prm->delta;
  A.kscale = prm->kernel_scale;
  A.cs = prm->cs;
  // viscid prefactor 2 (d+2) nu rho0 with d = 2, formed in double as the
  // reference forms it from Python scalars (viscosity.py:218)
  A.visc_c = prm->viscosity == 1 ? (float)(2.0 * 4.0 * (double)prm->visc_coeff * (double)prm->rho0) : prm->visc_coeff;
  A.eps = prm->visc_eps;
  A.gx = prm->gravity[0];
  A.gy = prm->gravity[1];
  A.visc_scheme = prm->viscosity;
  A.pi_switch = prm->pi_switch;
  A.diffusion = prm->diffusion;
  return DSPH_OK;
}

int dsph::launch_normalization_per_particle(const PairsDev& P, const Dom* doms, const int32_t* dom_of,
                                            const float* mass, const float* rho, const float* vol, float2* grad_rho,
                                            int block, hipStream_t st) {
  if (block != 64 && block != 128 && block != 256) return DSPH_ERR_ARG;
  if (P.nq == 0) return DSPH_OK;
  const DomainPerParticle dp{doms, dom_of};
  if (vol)
    k_dsph_normalization<DomainPerParticle, true><<<grid_for(P.nq, block), block, 0, st>>>(
        P, make_boundary(nullptr), dp, mass, rho, vol, nullptr, nullptr, grad_rho);
  else
    k_dsph_normalization<DomainPerParticle, false><<<grid_for(P.nq, block), block, 0, st>>>(
        P, make_boundary(nullptr), dp, mass, rho, nullptr, nullptr, nullptr, grad_rho);
  return launch_status();
}

int dsph::launch_rhs_per_particle(const PairsDev& P, const RhsArgs& A, const Dom* doms, const int32_t* dom_of,
                                  int block, hipStream_t st) {
  if (block != 64 && block != 128 && block != 256) return DSPH_ERR_ARG;
  if (P.nq == 0) return DSPH_OK;
  k_dsph_rhs<<<grid_for(P.nq, block), block, 0, st>>>(P, make_boundary(nullptr), A, DomainPerParticle{doms, dom_of});
  return launch_status();
}

extern "C" int dsph_deltasph_rhs(const dsph_pairs* pairs, const dsph_boundary* boundary,
                                 const dsph_deltasph_params* prm, const float* vel, const float* mass,
                                 const float* rho, const float* grad_rho, float* dudt, float* drhodt, float* terms,
                                 void* stream) {
  int s = check_same_set(pairs);
  if (s != DSPH_OK) return s;
  s = check_boundary(boundary, pairs);
  if (s != DSPH_OK) return s;
  if (!prm) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!vel || !mass || !rho || !dudt || !drhodt) return DSPH_ERR_ARG;
  RhsArgs A;
  s = make_rhs_args(prm, vel, mass, rho, grad_rho, dudt, drhodt, terms, &A);
  if (s != DSPH_OK) return s;
  PairsDev P = make_pairs(*pairs);
  BoundaryDev B = make_boundary(boundary);
  if (B.on && B.viscosity == 1) B.visc_c_viscid = (float)(2.0 * 4.0 * (double)B.visc_coeff * (double)prm->rho0);
  k_dsph_rhs<<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(P, B, A, OneDomain{P.dom});
  return launch_status();
}

extern "C" int dsph_rk_update(int64_t n, const float* x0, const float* dxdt, float* x, const float* u0,
                              const float* dudt, float* u, const float* rho0, const float* drhodt, float* rho,
                              float c, void* stream) {
  if (n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if ((x && (!x0 || !dxdt)) || (u && (!u0 || !dudt)) || (rho && (!rho0 || !drhodt))) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  k_rk_update<<<grid_for(n), kBlock, 0, as_stream(stream)>>>((int)n, x0, dxdt, x, u0, dudt, u, rho0, drhodt, rho,
                                                              c, nullptr, 0);
  return launch_status();
}

extern "C" int dsph_rk_update_dev(int64_t n, const float* x0, const float* dxdt, float* x, const float* u0,
                                  const float* dudt, float* u, const float* rho0, const float* drhodt, float* rho,
                                  const float* dt, int32_t half, void* stream) {
  if (n < 0 || n >= INT32_MAX || !dt) return DSPH_ERR_ARG;
  if ((x && (!x0 || !dxdt)) || (u && (!u0 || !dudt)) || (rho && (!rho0 || !drhodt))) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  k_rk_update<<<grid_for(n), kBlock, 0, as_stream(stream)>>>((int)n, x0, dxdt, x, u0, dudt, u, rho0, drhodt, rho,
                                                              0.0f, dt, half ? 1 : 0);
  return launch_status();
}

static int rk4_combine(int64_t n, const dsph_rk4_stages* k, float* x, float* u, float* rho, float dt,
                       const float* dt_dev, float* dxdt_avg, float* dudt_avg, float* drhodt_avg, void* stream) {
  if (!k || n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  Rk4Args A;
  A.dt_dev = dt_dev;
  const float* const* src[3] = {k->dxdt, k->dudt, k->drhodt};
  float* st[3] = {x, u, rho};
  float* av[3] = {dxdt_avg, dudt_avg, drhodt_avg};
  for (int f = 0; f < 3; ++f) {
    A.state[f] = st[f];
    A.avg[f] = av[f];
    for (int s = 0; s < 4; ++s) {
      A.k[f][s] = src[f][s];
      if (st[f] && !src[f][s]) return DSPH_ERR_ARG;
    }
  }
  A.dt = dt;
  A.n = (int)n;
  k_rk4_combine<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(A);
  return launch_status();
}

extern "C" int dsph_rk4_combine(int64_t n, const dsph_rk4_stages* k, float* x, float* u, float* rho, float dt,
                                float* dxdt_avg, float* dudt_avg, float* drhodt_avg, void* stream) {
  return rk4_combine(n, k, x, u, rho, dt, nullptr, dxdt_avg, dudt_avg, drhodt_avg, stream);
}

extern "C" int dsph_rk4_combine_dev(int64_t n, const dsph_rk4_stages* k, float* x, float* u, float* rho,
                                    const float* dt, float* dxdt_avg, float* dudt_avg, float* drhodt_avg,
                                    void* stream) {
  if (!dt) return DSPH_ERR_ARG;
  return rk4_combine(n, k, x, u, rho, 0.0f, dt, dxdt_avg, dudt_avg, drhodt_avg, stream);
}

extern "C" size_t dsph_max_norm_workspace(void) { return sizeof(double) * 1040; }

extern "C" int dsph_max_norm2(const float* v, int64_t n, float* out_host, void* ws, size_t ws_bytes, void* stream) {
  if (!out_host || n < 0 || n >= INT32_MAX || (n > 0 && !v)) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_max_norm_workspace()) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  const int nb = 1024;
  float* part = static_cast<float*>(ws);
  float* out = part + nb;
  k_max_norm<<<nb, kBlock, 0, st>>>(reinterpret_cast<const float2*>(v), (int)n, part);
  k_max_final<<<1, 64, 0, st>>>(part, nb, out);
  DSPH_LAUNCH_CHECK();
  DSPH_CHECK_HIP(hipMemcpyAsync(out_host, out, sizeof(float), hipMemcpyDeviceToHost, st));
  DSPH_CHECK_HIP(hipStreamSynchronize(st));
  return DSPH_OK;
}

extern "C" size_t dsph_shift_workspace(int64_t) { return dsph_max_norm_workspace(); }

extern "C" int dsph_deltaplus_shift(const dsph_pairs* pairs, const dsph_boundary* boundary,
                                    const dsph_shift_params* prm, const float* mass, const float* rho,
                                    const float* vel, float* pos, float* shift_out, void* ws, size_t ws_bytes,
                                    void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (!prm || pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  s = check_boundary(boundary, pairs);
  if (s != DSPH_OK) return s;
  if (!(prm->dx > 0.0) || !(prm->support > 0.0) || !(prm->kernel_scale > 0.f) || !(prm->cs > 0.f))
    return DSPH_ERR_ARG;
  const int64_t n = pairs->nq;
  if (n == 0) return DSPH_OK;
  if (!mass || !rho || !pos || (prm->mach < 0.f && !vel)) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_shift_workspace(n)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  // W_0 = kernel(tensor(dx / support), tensor(support)) in float32
  const float q0 = (float)(prm->dx / prm->support);
  const float h0 = (float)prm->support;
  float t = 1.0f - q0;
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  const float t4 = ((t * t) * t) * t;
  const float w0 = ((t4 * (1.0f + 4.0f * q0)) * kC2) / (h0 * h0);
  ShiftArgs A;
  A.mass = mass;
  A.rho = rho;
  A.pos = reinterpret_cast<float2*>(pos);
  A.shift_out = reinterpret_cast<float2*>(shift_out);
  A.w0 = w0;
  A.R = prm->R;
  A.nexp = prm->n;
  A.nint = (prm->n >= 0.f && prm->n <= 16.f && prm->n == floorf(prm->n)) ? (int)prm->n : -1;
  A.cfl = prm->cfl;
  A.kscale = prm->kernel_scale;
  A.thr = (float)((double)prm->threshold * prm->dx);
  A.cs = prm->cs;
  A.mach = prm->mach;
  float* vmax = reinterpret_cast<float*>(static_cast<char*>(ws) + sizeof(double) * 1032);
  A.vmax = vmax;
  if (prm->mach < 0.f) {
    const int nb = 1024;
    float* part = static_cast<float*>(ws);
    k_max_norm<<<nb, kBlock, 0, st>>>(reinterpret_cast<const float2*>(vel), (int)n, part);
    k_max_final<<<1, 64, 0, st>>>(part, nb, vmax);
    DSPH_LAUNCH_CHECK();
  }
  k_deltaplus_shift<<<grid_for(n), kBlock, 0, st>>>(make_pairs(*pairs), make_boundary(boundary), A);
  return launch_status();
}

extern "C" int dsph_sum_squares(const float* v, int64_t n, double* out_host, void* ws, size_t ws_bytes, void* stream) {
  if (!out_host || n < 0 || (n > 0 && !v)) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_max_norm_workspace()) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  const int nb = 1024;
  double* part = static_cast<double*>(ws);
  double* out = part + nb;
  k_sumsq<<<nb, kBlock, 0, st>>>(v, n, part);
  k_sum_final<<<1, 64, 0, st>>>(part, nb, out);
  DSPH_LAUNCH_CHECK();
  DSPH_CHECK_HIP(hipMemcpyAsync(out_host, out, sizeof(double), hipMemcpyDeviceToHost, st));
  DSPH_CHECK_HIP(hipStreamSynchronize(st));
  return DSPH_OK;
}

extern "C" int dsph_timestep_dev(const float* dudt_avg, int64_t n, const dsph_timestep_params* tp,
                                 int32_t accel_norm, float* dt, double* time, void* ws, size_t ws_bytes,
                                 void* stream) {
  if (!tp || !dt || n < 0 || n >= INT32_MAX || (n > 0 && !dudt_avg) || accel_norm < 0 || accel_norm > 1)
    return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_max_norm_workspace()) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  const int nb = 1024;
  if (accel_norm == 0) {
    // the reduction tree of dsph_sum_squares over the 2n components
    double* part = static_cast<double*>(ws);
    double* out = part + nb;
    k_sumsq<<<nb, kBlock, 0, st>>>(dudt_avg, 2 * n, part);
    k_sum_final<<<1, 64, 0, st>>>(part, nb, out);
    DSPH_LAUNCH_CHECK();
    k_timestep_single<<<1, 64, 0, st>>>(out, nullptr, *tp, dt, time);
  } else {
    // the reduction of dsph_max_norm2
    float* part = static_cast<float*>(ws);
    float* out = part + nb;
    k_max_norm<<<nb, kBlock, 0, st>>>(reinterpret_cast<const float2*>(dudt_avg), (int)n, part);
    k_max_final<<<1, 64, 0, st>>>(part, nb, out);
    DSPH_LAUNCH_CHECK();
    k_timestep_single<<<1, 64, 0, st>>>(nullptr, out, *tp, dt, time);
  }
  return launch_status();
}

extern "C" int dsph_boundary_ghosts(const float* bpos, const float* distances, const float* normals, int64_t nb,
                                    float support, float* ghost_pos, void* stream) {
  if (nb < 0 || nb >= INT32_MAX || (nb > 0 && (!bpos || !distances || !normals || !ghost_pos))) return DSPH_ERR_ARG;
  if (nb == 0) return DSPH_OK;
  k_ghosts<<<grid_for(nb), kBlock, 0, as_stream(stream)>>>(reinterpret_cast<const float2*>(bpos), distances,
                                                            reinterpret_cast<const float2*>(normals), (int)nb,
                                                            support, reinterpret_cast<float2*>(ghost_pos));
  return launch_status();
}

extern "C" int dsph_mdbc(const dsph_pairs* ghost_pairs, const float* fluid_mass, const float* fluid_rho,
                         const float* boundary_pos, float rho0, float support, int32_t variant, float* boundary_rho,
                         float* shepard, void* stream) {
  int s = check_pairs(ghost_pairs);
  if (s != DSPH_OK) return s;
  if (variant != 0 && variant != 1) return DSPH_ERR_ARG;
  if (ghost_pairs->nq == 0) return DSPH_OK;
  if (!boundary_pos || !boundary_rho || (ghost_pairs->nr > 0 && (!fluid_mass || !fluid_rho))) return DSPH_ERR_ARG;
  PairsDev G = make_pairs(*ghost_pairs);
  k_mdbc<<<grid_for(G.nq), kBlock, 0, as_stream(stream)>>>(G, fluid_mass, fluid_rho,
                                                            reinterpret_cast<const float2*>(boundary_pos), rho0,
                                                            support, variant, boundary_rho, shepard);
  return launch_status();
}
