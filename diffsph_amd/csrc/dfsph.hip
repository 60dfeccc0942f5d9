// DFSPH relaxed-Jacobi pressure solve (v2/modules/dfsph.py) as gather
// kernels with device-side loop control.
//
// The prepare sweep (once per step, positions fixed for the whole solve)
// evaluates the geometry of every pair and stores the scalar
// s_ij = V_j dW/dr / (r + 1e-7) in the neighbor list's layout; the
// coefficient kV_ij = V_j gradW_ij -- the product the reference forms as
// (m_j/rho_j) * gradKernels (v2/sphOps.py:50,93) -- is d_ij * s_ij, with the
// min-image offset d_ij = x_i - x_j recomputed from the positions (a gather
// that hits the caches).  Every later sweep of both solves streams s (4 B per
// entry) instead of recomputing sqrt/div per pair or streaming the 8-byte
// vector; the on-the-fly evaluation gives the same bits.
//
// Per iteration two pair sweeps:
//   A: a_i = -(1/rho_i) sum_j (p_i + p_j) kV_ij
//   B: K_i = -dtf sum_j (a_j - a_i) . kV_ij ; r_i = K_i - s_i ;
//      p_i += omega (s_i - K_i)/alpha_i (clamped at 0) -- in place, a thread
//      only touches its own p_i in B.
// B also reduces mean(clamp(r_i/rho0, min=-thr)) in fp64: every block
// writes a partial; a one-block kernel queued after B sums the partials in a
// fixed order, advances the iteration counter and sets `done` with the
// reference's loop condition (no grid-wide fence or atomic inside B).  Every
// kernel of the loop returns at once when `done` is set, so the host may
// queue several iterations ahead and synchronise only once per chunk.

#include "common.h"

namespace dsph {

struct SolveCtrl {
  int iter;
  int done;
  unsigned int ticket;
  int min_iters;
  int max_iters;
  int pad;
  double err;
  double err0;
  double thr;
};

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Deterministic block sum (fixed shuffle tree, then fixed wave order).
__device__ __forceinline__ double block_sum_d(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v = wave_sum_d(v);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
  }
  __syncthreads();
  return t;
}

// V_j gradW_ij of entry k as d_ij * s_ij (dfsph_t): s_ij stored per entry
// (kv != NULL) or evaluated on the fly.
__device__ __forceinline__ float2 coef_of(float2 xi, float2 xj, float sv, const Dom& d) {
  const float dx = mod_offset(xi.x - xj.x, d, 0);
  const float dy = mod_offset(xi.y - xj.y, d, 1);
  return make_float2(dx * sv, dy * sv);
}

__device__ __forceinline__ float2 kv_of(const PairsDev& P, const float* __restrict__ kv, int k, int j, float2 xi,
                                        float hi, const float* __restrict__ vol) {
  const float2 xj = P.rpos[j];
  if (kv) return coef_of(xi, xj, kv[k], P.dom);
  const float hj = h_of(P.rh, j, P.h0);
  Geom g = pair_geom(xi, xj, pick_h(P.mode, hi, hj), P.dom);
  const float sv = vol[j] * dfsph_t(g);
  return make_float2(g.dx * sv, g.dy * sv);
}

// ---------------------------------------------------------------------------
// Coefficients for both solves + divergence of v + viscosity, one sweep.
// ---------------------------------------------------------------------------
struct PrepArgs {
  const float2* vel;
  const float* areas;
  const float* vol;    // m/rho
  const float* rho;
  float dt, dt2, rho0;
  int visc_on;
  int visc_scheme;
  float visc_alpha, visc_cs, visc_rho0, visc_eps;
  int pi_switch;
  float visc_factor;   // 2(d+2) nu rho0 for the viscid scheme
};

// Neighbor data of one entry of the prepare pass; c2 = V_j^2 / (area_j rho0)
// of the alpha term.
struct PrepNb {
  float2 x, v;
  float h, vol, c2;
};

// One row of the prepare pass: neighbors by index idx(k) and data src(j),
// four entries' loads ahead of their sums (the sums run in entry order).
template <class Idx, class Src>
__device__ __forceinline__ void prepare_row(const PairsDev& P, const PrepArgs& A, int i, const Idx& idx,
                                            const Src& src, float* __restrict__ alpha_div,
                                            float* __restrict__ alpha_den, float* __restrict__ div_src,
                                            float2* __restrict__ visc_acc, float* __restrict__ kv_out) {
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const KSel ks = ksel(P);
  const float2 vi = A.vel[i];
  float k1x = 0.f, k1y = 0.f, k2 = 0.f, dv = 0.f, vsx = 0.f, vsy = 0.f;
  const RowRef rr = row_ref(P, i);
  int n = 0, k = rr.base;
  auto body = [&](int k, const PrepNb& d) {
    const float h = pick_h(P.mode, hi, d.h);
    Geom g = pair_geom(xi, d.x, h, ks(h), P.dom);
    const float Vj = d.vol;
    const float2 vj = d.v;
    // gradW_ij = d_ij * t_ij, V_j gradW_ij = d_ij * s_ij (dfsph_t)
    const float t = dfsph_t(g);
    const float sv = Vj * t;
    const float kx = g.dx * sv, ky = g.dy * sv;
    if (kv_out) kv_out[k] = sv;
    // alpha terms (dfsph.py:13-20)
    k1x += kx;
    k1y += ky;
    const float gx = g.dx * t, gy = g.dy * t;
    k2 += d.c2 * (gx * gx + gy * gy);
    // divergence 'difference' of v (sphOps.py:92-101)
    dv += (vj.x - vi.x) * kx + (vj.y - vi.y) * ky;
    if (A.visc_on) {
      // delta-SPH viscosity (viscosity.py:141-166 / 205-218)
      const float vx = vi.x - vj.x, vy = vi.y - vj.y;
      const float rij = g.q * g.h;
      const float vr = vx * g.ex + vy * g.ey;
      float pi = vr / (rij + A.visc_eps * g.h * g.h);
      if (A.pi_switch) pi = vr < 0.0f ? pi : 0.0f;
      const float f = Vj * pi;
      vsx += f * g.gx;
      vsy += f * g.gy;
    }
  };
  for (; n + 4 <= rr.cnt; n += 4, k += 4 * rr.step) {
    int j[4];
    PrepNb d[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = idx(k + u * rr.step);
#pragma unroll
    for (int u = 0; u < 4; ++u) d[u] = src(j[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) body(k + u * rr.step, d[u]);
  }
  for (; n < rr.cnt; ++n, k += rr.step) body(k, src(idx(k)));
  const float Vi = A.vol[i];
  const float mass_i = A.areas[i] * A.rho0;
  const float k1sq = k1x * k1x + k1y * k1y;
  {
    const float fac = -A.dt2 * Vi;
    alpha_den[i] = fac / mass_i * k1sq + fac * k2;
  }
  {
    const float fac = -A.dt * Vi;
    alpha_div[i] = fac / mass_i * k1sq + fac * k2;
  }
  div_src[i] = dv;
  if (A.visc_on && visc_acc) {
    float s;
    if (A.visc_scheme == 0)
      s = A.visc_alpha * hi * A.visc_cs * A.visc_rho0 / A.rho[i];
    else
      s = A.visc_factor / A.rho[i];
    visc_acc[i] = make_float2(s * vsx, s * vsy);
  }
}

// The neighbor data of particle j from the global arrays.
__device__ __forceinline__ PrepNb prep_nb(const PairsDev& P, const PrepArgs& A, int j) {
  PrepNb d;
  d.x = P.rpos[j];
  d.v = A.vel[j];
  d.h = h_of(P.rh, j, P.h0);
  d.vol = A.vol[j];
  d.c2 = (d.vol * d.vol) / (A.areas[j] * A.rho0);
  return d;
}

__global__ void __launch_bounds__(kBlock) k_dfsph_prepare(PairsDev P, PrepArgs A, float* __restrict__ alpha_div,
                                                          float* __restrict__ alpha_den, float* __restrict__ div_src,
                                                          float2* __restrict__ visc_acc, float* __restrict__ kv_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  prepare_row(
      P, A, i, [&](int k) { return P.nbr[k]; }, [&](int j) { return prep_nb(P, A, j); }, alpha_div, alpha_den,
      div_src, visc_acc, kv_out);
}

// The prepare pass over neighbor tiles (uniform support): (x, y, v) and
// (V, c2) of the workgroup's neighbors staged in LDS, c2 evaluated once per
// staged particle instead of once per entry (the same expression).
__global__ void __launch_bounds__(kBlock) k_dfsph_prepare_tiled(PairsDev P, PrepArgs A, float* __restrict__ alpha_div,
                                                                float* __restrict__ alpha_den,
                                                                float* __restrict__ div_src,
                                                                float2* __restrict__ visc_acc,
                                                                float* __restrict__ kv_out) {
  extern __shared__ float4 tile_xv[];
  float2* tile_vc = reinterpret_cast<float2*>(tile_xv + P.tcap);
  const int b = blockIdx.x;
  const int i = b * kBlock + threadIdx.x;
  const bool tiled = stage_tile(P, b, [&](int o, int j) {
    const PrepNb d = prep_nb(P, A, j);
    tile_xv[o] = make_float4(d.x.x, d.x.y, d.v.x, d.v.y);
    tile_vc[o] = make_float2(d.vol, d.c2);
  });
  if (i >= P.nq) return;
  if (tiled)
    prepare_row(
        P, A, i, IdxSlot{P.tslot},
        [&](int s) {
          const float4 a = tile_xv[s];
          const float2 c = tile_vc[s];
          return PrepNb{make_float2(a.x, a.y), make_float2(a.z, a.w), P.h0, c.x, c.y};
        },
        alpha_div, alpha_den, div_src, visc_acc, kv_out);
  else
    prepare_row(
        P, A, i, [&](int k) { return P.nbr[k]; }, [&](int j) { return prep_nb(P, A, j); }, alpha_div, alpha_den,
        div_src, visc_acc, kv_out);
}

// Source of a solve with advection adv (may be NULL = 0): div(v*) with
// v* = v + dt adv; density source 1 - areas_i/V_i - dt div(v*)
// (dfsph.py:57, 62-65).
// Streamed per-entry loads of the sweeps: the neighbor index and V_j gradW_ij
// are read exactly once per launch, so they bypass the caches
// (non-temporal) and leave L2 to the gathered neighbor fields.
__device__ __forceinline__ int ld_stream(const int32_t* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ld_stream(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float2 ld_stream(const float2* p) {
  return make_float2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}

// Neighbor index of entry k of row i: the int32 list, or the 16-bit modular
// delta j = (i + d) mod n of the compact copy.
struct Idx32 {
  const int32_t* a;
  __device__ __forceinline__ int operator()(int k) const { return ld_stream(a + k); }
};
struct Idx16 {
  const int16_t* a;
  int i, n;
  __device__ __forceinline__ int operator()(int k) const {
    int j = i + (int)__builtin_nontemporal_load(a + k);
    j += j < 0 ? n : 0;
    j -= j >= n ? n : 0;
    return j;
  }
};

__device__ __forceinline__ bool use_idx16(const PairsDev& P) { return P.nbr16 && *P.ovf == 0; }

// Divergence source (dfsph.py:30-33, 62-63) with the coefficients streamed:
// dv_i = sum_j (v*_j - v*_i) . V_j gradW_ij, v* = v + dt adv, entries in row
// order with kUnroll index / coefficient loads in flight.
// Neighbor position and v* = v + dt adv of entry j: separate arrays, or
// one (x, y, v*) record per particle.
struct GatherXV {
  const float2* x;
  const float2* vel;
  const float2* adv;
  float dt;
  __device__ __forceinline__ void operator()(int j, float2& xj, float2& vj) const {
    xj = x[j];
    const float2 aj = adv ? adv[j] : make_float2(0.f, 0.f);
    vj = make_float2(vel[j].x + dt * aj.x, vel[j].y + dt * aj.y);
  }
};
struct GatherXV4 {
  const float4* r;
  __device__ __forceinline__ void operator()(int j, float2& xj, float2& vj) const {
    const float4 v = r[j];
    xj = make_float2(v.x, v.y);
    vj = make_float2(v.z, v.w);
  }
};

__global__ void __launch_bounds__(kBlock) k_pack_xv(const float2* __restrict__ rpos, const float2* __restrict__ vel,
                                                    const float2* __restrict__ adv, float dt, int n,
                                                    float4* __restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 x = rpos[i];
  const float2 a = adv ? adv[i] : make_float2(0.f, 0.f);
  rec[i] = make_float4(x.x, x.y, vel[i].x + dt * a.x, vel[i].y + dt * a.y);
}

template <class Idx, class Gat>
__device__ __forceinline__ float source_kv(const PairsDev& P, float2 xi, const RowRef& rr, const Idx& idx,
                                           const float* __restrict__ kv, const Gat& gat, float2 vi) {
  float dv = 0.f;
  int n = 0, k = rr.base;
  for (; n + 4 <= rr.cnt; n += 4, k += 4 * rr.step) {
    int j[4];
    float sv[4];
    float2 c[4], vj[4], xj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      j[u] = idx(k + u * rr.step);
      sv[u] = ld_stream(kv + k + u * rr.step);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) gat(j[u], xj[u], vj[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c[u] = coef_of(xi, xj[u], sv[u], P.dom);
      dv += (vj[u].x - vi.x) * c[u].x + (vj[u].y - vi.y) * c[u].y;
    }
  }
  for (; n < rr.cnt; ++n, k += rr.step) {
    const int j = idx(k);
    float2 xj, vj;
    gat(j, xj, vj);
    const float2 c = coef_of(xi, xj, ld_stream(kv + k), P.dom);
    dv += (vj.x - vi.x) * c.x + (vj.y - vi.y) * c.y;
  }
  return dv;
}

__global__ void __launch_bounds__(kBlock) k_dfsph_source(PairsDev P, const float* __restrict__ kv,
                                                         const float2* __restrict__ vel,
                                                         const float2* __restrict__ adv,
                                                         const float* __restrict__ vol,
                                                         const float* __restrict__ areas, float dt, int density,
                                                         float* __restrict__ src, const float4* __restrict__ rec) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const float2 a_i = adv ? adv[i] : make_float2(0.f, 0.f);
  const float2 vi = make_float2(vel[i].x + dt * a_i.x, vel[i].y + dt * a_i.y);
  float dv = 0.f;
  const RowRef rr = row_ref(P, i);
  if (kv) {
    const bool i16 = use_idx16(P);
    if (rec) {
      const GatherXV4 g{rec};
      dv = i16 ? source_kv(P, xi, rr, Idx16{P.nbr16, i, P.nr}, kv, g, vi)
               : source_kv(P, xi, rr, Idx32{P.nbr}, kv, g, vi);
    } else {
      const GatherXV g{P.rpos, vel, adv, dt};
      dv = i16 ? source_kv(P, xi, rr, Idx16{P.nbr16, i, P.nr}, kv, g, vi)
               : source_kv(P, xi, rr, Idx32{P.nbr}, kv, g, vi);
    }
  } else {
    for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
      const int j = P.nbr[k];
      const float2 c = kv_of(P, kv, k, j, xi, hi, vol);
      const float2 aj = adv ? adv[j] : make_float2(0.f, 0.f);
      const float2 vj = make_float2(vel[j].x + dt * aj.x, vel[j].y + dt * aj.y);
      dv += (vj.x - vi.x) * c.x + (vj.y - vi.y) * c.y;
    }
  }
  src[i] = density ? (1.0f - areas[i] / vol[i]) - dt * dv : dv;
}

// Entries per unrolled group: the group's index and coefficient loads are
// issued together, then its gathers, then the additions in entry order (the
// sum is accumulated exactly as the one-at-a-time loop does).
#ifndef DSPH_SWEEP_UNROLL
#define DSPH_SWEEP_UNROLL 4
#endif
constexpr int kUnroll = DSPH_SWEEP_UNROLL;

// Sweep A (dfsph.py:35-37): a_i = -(1/rho_i) sum_j (p_i + p_j) V_j gradW_ij;
// with `diff` the ISPH form (incompressible.py:276-285, gradient
// 'difference'): a_i = -(1/rho_i) sum_j (p_i - p_j) V_j gradW_ij.
// `ctrl` NULL -> unconditional.  With `kv` the coefficients are streamed,
// else evaluated on the fly.
// Neighbor fields gathered by the sweeps: separate arrays, or one 16-byte
// record per particle (x, y, p) / (x, y, a) so that a pair costs one gather.
struct GatherXP {   // sweep A: position and pressure
  const float2* x;
  const float* p;
  __device__ __forceinline__ void operator()(int j, float2& xj, float& pj) const {
    xj = x[j];
    pj = p[j];
  }
};
struct GatherXP4 {
  const float4* r;
  __device__ __forceinline__ void operator()(int j, float2& xj, float& pj) const {
    const float4 v = r[j];
    xj = make_float2(v.x, v.y);
    pj = v.z;
  }
};
struct GatherXA {   // sweep B: position and pressure acceleration
  const float2* x;
  const float2* a;
  __device__ __forceinline__ void operator()(int j, float2& xj, float2& aj) const {
    xj = x[j];
    aj = a[j];
  }
};
struct GatherXA4 {
  const float4* r;
  __device__ __forceinline__ void operator()(int j, float2& xj, float2& aj) const {
    const float4 v = r[j];
    xj = make_float2(v.x, v.y);
    aj = make_float2(v.z, v.w);
  }
};

template <class Idx, class Gat>
__device__ __forceinline__ void sweep_a_kv(const PairsDev& P, float2 xi, const RowRef& rr, const Idx& idx,
                                           const float* __restrict__ kv, const Gat& gat, float pi, float sj,
                                           float& ax, float& ay) {
  int n = 0, k = rr.base;
  for (; n + kUnroll <= rr.cnt; n += kUnroll, k += kUnroll * rr.step) {
    int j[kUnroll];
    float sv[kUnroll];
    float2 xj[kUnroll];
    float pj[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      j[u] = idx(k + u * rr.step);
      sv[u] = ld_stream(kv + k + u * rr.step);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      float q;
      gat(j[u], xj[u], q);
      pj[u] = sj * q;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const float2 c = coef_of(xi, xj[u], sv[u], P.dom);
      const float s = pi + pj[u];
      ax += s * c.x;
      ay += s * c.y;
    }
  }
  for (; n < rr.cnt; ++n, k += rr.step) {
    const int j = idx(k);
    float2 xj;
    float q;
    gat(j, xj, q);
    const float2 c = coef_of(xi, xj, ld_stream(kv + k), P.dom);
    const float s = pi + sj * q;
    ax += s * c.x;
    ay += s * c.y;
  }
}

// Sweep A (dfsph.py:35-37): a_i = -(1/rho_i) sum_j (p_i + p_j) V_j gradW_ij;
// with `diff` the ISPH form (incompressible.py:276-285, gradient
// 'difference'): a_i = -(1/rho_i) sum_j (p_i - p_j) V_j gradW_ij.
// `ctrl` NULL -> unconditional.  With `kv` the coefficients are streamed,
// else evaluated on the fly.
struct SweepAArgs {
  const float* kv;
  const float* vol;
  const float* rho;
  const float* p;
  float2* acc;
  int diff;
  const float4* xp;
  float4* xa;
};

// One row of sweep A.  xp (with kv): the (x, y, p) records replace p and
// the positions in the gathers; xa: the result is written as the (x, y, a)
// record instead of acc.
__device__ __forceinline__ void sweep_a_row(const PairsDev& P, const SweepAArgs& A, int i) {
  const float* kv = A.kv;
  const float* p = A.p;
  const float4* xp = A.xp;
  const float4 own = xp ? xp[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float pi = xp ? own.z : p[i];
  // -p_j turns the sum into the difference form; negation is exact
  const float sj = A.diff ? -1.0f : 1.0f;
  float ax = 0.f, ay = 0.f;
  const RowRef rr = row_ref(P, i);
  const float2 xi = P.qpos[i];
  if (kv) {
    const bool i16 = use_idx16(P);
    if (xp) {
      const GatherXP4 g{xp};
      if (i16)
        sweep_a_kv(P, xi, rr, Idx16{P.nbr16, i, P.nr}, kv, g, pi, sj, ax, ay);
      else
        sweep_a_kv(P, xi, rr, Idx32{P.nbr}, kv, g, pi, sj, ax, ay);
    } else {
      const GatherXP g{P.rpos, p};
      if (i16)
        sweep_a_kv(P, xi, rr, Idx16{P.nbr16, i, P.nr}, kv, g, pi, sj, ax, ay);
      else
        sweep_a_kv(P, xi, rr, Idx32{P.nbr}, kv, g, pi, sj, ax, ay);
    }
  } else {
    const float hi = h_of(P.qh, i, P.h0);
    for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
      const int j = P.nbr[k];
      const float2 c = kv_of(P, nullptr, k, j, xi, hi, A.vol);
      const float s = pi + sj * p[j];
      ax += s * c.x;
      ay += s * c.y;
    }
  }
  const float r = A.rho[i];
  const float2 a = make_float2(-ax / r, -ay / r);
  if (A.xa) {
    const float2 xr = xp ? make_float2(own.x, own.y) : P.rpos[i];
    A.xa[i] = make_float4(xr.x, xr.y, a.x, a.y);
  } else {
    A.acc[i] = a;
  }
}

// Sweep A (dfsph.py:35-37): a_i = -(1/rho_i) sum_j (p_i + p_j) V_j gradW_ij;
// with `diff` the ISPH form (incompressible.py:276-285, gradient
// 'difference'): a_i = -(1/rho_i) sum_j (p_i - p_j) V_j gradW_ij.
// `ctrl` NULL -> unconditional.  With `kv` the coefficients are streamed,
// else evaluated on the fly.
__global__ void __launch_bounds__(kBlock) k_dfsph_sweep_a(PairsDev P, const float* __restrict__ kv,
                                                          const float* __restrict__ vol,
                                                          const float* __restrict__ rho, const float* __restrict__ p,
                                                          float2* __restrict__ acc, const SolveCtrl* ctrl, int diff,
                                                          const float4* __restrict__ xp, float4* __restrict__ xa) {
  if (ctrl && ctrl->done) return;
  int i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const SweepAArgs A{kv, vol, rho, p, acc, diff, xp, xa};
  sweep_a_row(P, A, i);
}

struct SweepBArgs {
  const float* kv;
  const float* vol;
  const float* src;
  const float* alpha;
  float* p;
  float* residual;
  const float2* acc;
  double* partials;
  SolveCtrl* ctrl;
  float dtf;
  float omega;
  float rho0;
  float neg_thr;
  int clamp;
  int n;
  int isph;  // error metric |r| (incompressible.py:357) instead of max(r/rho0, -thr)
  const float4* xa;  // (x, y, a) records in place of acc (with kv)
  float4* xp;        // the updated (x, y, p) records are written here too
};

// The loop condition of dfsph.py:88 from the per-block partials, summed in a
// fixed order by one block of kBlock threads.  (A last-block-ticket fusion
// into sweep B needs device-scope fences, which on this part write back the
// per-XCD L2s: 1.4-1.7x slower solves, so the reduction stays a kernel.)
__device__ __forceinline__ void reduce_partials(const double* partials, int nb, int n, SolveCtrl* ctrl, double* sh) {
  // Thread t adds partials t, t + blockDim, ... in that order; eight loads
  // are issued before their adds so that their latencies overlap.
  constexpr int kR = 8;
  double v = 0.0;
  int k = threadIdx.x;
  for (; k + (kR - 1) * (int)blockDim.x < nb; k += kR * (int)blockDim.x) {
    double t[kR];
#pragma unroll
    for (int u = 0; u < kR; ++u) t[u] = partials[k + u * (int)blockDim.x];
#pragma unroll
    for (int u = 0; u < kR; ++u) v += t[u];
  }
  for (; k < nb; k += blockDim.x) v += partials[k];
  double tot = block_sum_d(v, sh);
  if (threadIdx.x == 0) {
    double mean = tot / (double)n;
    ctrl->iter += 1;
    ctrl->err = mean;
    if (ctrl->iter == 1) ctrl->err0 = mean;
    bool go = ctrl->iter < ctrl->max_iters && (ctrl->iter < ctrl->min_iters || mean > ctrl->thr);
    ctrl->done = go ? 0 : 1;
  }
}

template <class Idx, class Gat>
__device__ __forceinline__ float sweep_b_kv(const PairsDev& P, float2 xi, const RowRef& rr, const Idx& idx,
                                            const float* __restrict__ kv, const Gat& gat, float2 ai) {
  float sum = 0.f;
  int n = 0, k = rr.base;
  for (; n + kUnroll <= rr.cnt; n += kUnroll, k += kUnroll * rr.step) {
    int j[kUnroll];
    float sv[kUnroll];
    float2 xj[kUnroll];
    float2 aj[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      j[u] = idx(k + u * rr.step);
      sv[u] = ld_stream(kv + k + u * rr.step);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) gat(j[u], xj[u], aj[u]);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const float2 c = coef_of(xi, xj[u], sv[u], P.dom);
      sum += (aj[u].x - ai.x) * c.x + (aj[u].y - ai.y) * c.y;
    }
  }
  for (; n < rr.cnt; ++n, k += rr.step) {
    const int j = idx(k);
    float2 xj, aj;
    gat(j, xj, aj);
    const float2 c = coef_of(xi, xj, ld_stream(kv + k), P.dom);
    sum += (aj.x - ai.x) * c.x + (aj.y - ai.y) * c.y;
  }
  return sum;
}

// The pressure update of row i from its neighbor sum; returns the row's
// term of the error metric.
__device__ __forceinline__ double sweep_b_finish(const SweepBArgs& B, int i, float sum, float2 xr) {
  const float K = -B.dtf * sum;
  const float s = B.src[i];
  const float res = K - s;
  float pn = B.p[i] + B.omega * (s - K) / B.alpha[i];
  if (B.clamp) pn = pn < 0.0f ? 0.0f : pn;  // torch.max(p, 0)
  B.p[i] = pn;
  if (B.xp) B.xp[i] = make_float4(xr.x, xr.y, pn, 0.f);
  if (B.residual) B.residual[i] = res;
  if (B.isph) return (double)fabsf(res);
  const float rr0 = res / B.rho0;
  return (double)(rr0 < B.neg_thr ? B.neg_thr : rr0);
}

// One row of sweep B; returns the row's term of the error metric.
__device__ __forceinline__ double sweep_b_row(const PairsDev& P, const SweepBArgs& B, int i) {
  float2 ai, xr;
  if (B.xa) {
    const float4 own = B.xa[i];
    ai = make_float2(own.z, own.w);
    xr = make_float2(own.x, own.y);
  } else {
    ai = B.acc[i];
    xr = B.xp ? P.rpos[i] : make_float2(0.f, 0.f);
  }
  float sum = 0.f;
  const RowRef rr = row_ref(P, i);
  if (B.kv) {
    const bool i16 = use_idx16(P);
    const float2 xi = P.qpos[i];
    if (B.xa) {
      const GatherXA4 g{B.xa};
      sum = i16 ? sweep_b_kv(P, xi, rr, Idx16{P.nbr16, i, P.nr}, B.kv, g, ai)
                : sweep_b_kv(P, xi, rr, Idx32{P.nbr}, B.kv, g, ai);
    } else {
      const GatherXA g{P.rpos, B.acc};
      sum = i16 ? sweep_b_kv(P, xi, rr, Idx16{P.nbr16, i, P.nr}, B.kv, g, ai)
                : sweep_b_kv(P, xi, rr, Idx32{P.nbr}, B.kv, g, ai);
    }
  } else {
    const float2 xi = P.qpos[i];
    const float hi = h_of(P.qh, i, P.h0);
    for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
      const int j = P.nbr[k];
      const float2 c = kv_of(P, nullptr, k, j, xi, hi, B.vol);
      const float2 aj = B.acc[j];
      sum += (aj.x - ai.x) * c.x + (aj.y - ai.y) * c.y;
    }
  }
  return sweep_b_finish(B, i, sum, xr);
}

// Sweep B (dfsph.py:39-50): K_i = -dtf sum_j (a_j - a_i) . V_j gradW_ij,
// p_i += omega (s_i - K_i) / alpha_i (clamped at 0, in place: a thread
// touches only its own p_i), residual K_i - s_i, per-block float64 partial
// of max(r_i/rho0, -thr).
__global__ void __launch_bounds__(kBlock) k_dfsph_sweep_b(PairsDev P, SweepBArgs B) {
  __shared__ double sh[kBlock / 64];
  if (B.ctrl && B.ctrl->done) return;
  const int lb = logical_block();
  int i = lb * blockDim.x + threadIdx.x;
  double err = 0.0;
  if (i < P.nq) err = sweep_b_row(P, B, i);
  double bs = block_sum_d(err, sh);
  if (threadIdx.x == 0) B.partials[lb] = bs;
}

// ---------------------------------------------------------------------------
// Neighbor tiles.  The 256 rows of a workgroup are consecutive particles in
// cell order; the search finds their neighbors in the 3 x 3 cells around
// each row's cell, which in cell order are a few runs of consecutive
// particles (three for a workgroup within one row of cells, ~900 particles
// at 16 per cell).  The sweeps load those runs' 16-byte records into LDS
// once per workgroup (contiguous loads) and read them there through a
// 16-bit slot per list entry, instead of one global gather per entry.
// Arithmetic and summation order are those of the gathering sweeps, so the
// results are bit-identical.
static_assert(kBlock == DSPH_TILE_ROWS, "one tile per workgroup of sweep rows");
// The tiled sweeps keep 8 waves per SIMD (their kernel arguments and two
// row paths would take 100+ SGPRs, i.e. 7 waves; the excess goes to VGPR
// lanes): 78.7 -> 72.7 us (A) and 82.1 -> 75.1 us (B) at 1M particles.
#define DSPH_TILED_ATTR __attribute__((amdgpu_waves_per_eu(8)))
constexpr int kTileMaxSegs = 4;  // rows of cells one workgroup may span (9 ranges each)


// The cell holding sorted particle i, the last c with cell_start[c] <= i,
// by a 64-way search of the wave (3 rounds for 62k cells).
__device__ __forceinline__ int wave_cell_of(const int32_t* __restrict__ cs, int ncell, int i) {
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = ncell - 1;  // cs[lo] <= i; the answer is in [lo, hi]
  while (hi > lo) {
    const int step = (hi - lo + 64) / 64;
    const int c = lo + lane * step;
    const bool ok = c <= hi && cs[c] <= i;
    const unsigned long long m = __ballot(ok);
    lo += (63 - __clzll(m)) * step;
    hi = lo + step - 1 < hi ? lo + step - 1 : hi;
  }
  return lo;
}

// Tile runs, one wave per workgroup of rows: the cells of its rows span
// rows y0..y1 of the grid (segments); each segment's 3 x 3 stencil is three
// rows of cells by one interval (plus the pieces a periodic seam wraps),
// each a range of consecutive particles.  Lane l < 36 forms range l, the
// wave sorts them by start and merges overlapping or touching ones into
// disjoint runs.  tile_n = the runs' particle count, or -1 when the rows
// span more than kTileMaxSegs rows of cells or the runs number more than
// kTileMaxRuns.
__global__ void __launch_bounds__(kBlock) k_tile_runs(int nq, int ncx, int ncy, int px, int py,
                                                      const int32_t* __restrict__ cs, int ntiles,
                                                      int32_t* __restrict__ tile_runs, int32_t* __restrict__ tile_n) {
  __shared__ int srt_s[kBlock], srt_e[kBlock];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int b = blockIdx.x * (kBlock / 64) + wv;
  if (b >= ntiles) return;  // uniform over the wave; no block barrier follows
  int* ss = srt_s + wv * 64;
  int* se = srt_e + wv * 64;
  const int ncell = ncx * ncy;
  const int i0 = b * kBlock;
  const int i1 = (i0 + kBlock < nq ? i0 + kBlock : nq) - 1;
  const int c0 = wave_cell_of(cs, ncell, i0), c1 = wave_cell_of(cs, ncell, i1);
  const int y0 = c0 / ncx, y1 = c1 / ncx;
  if (y1 - y0 + 1 > kTileMaxSegs) {
    if (lane == 0) tile_n[b] = -1;
    return;
  }
  constexpr int kEmpty = 0x7fffffff;
  int rs = kEmpty, re = kEmpty;
  if (lane < kTileMaxSegs * 9) {
    const int y = y0 + lane / 9, dy = (lane / 3) % 3 - 1, part = lane % 3;
    int r = y + dy;
    bool ok = y <= y1;
    if (r < 0 || r >= ncy) {
      ok = ok && py;
      r += r < 0 ? ncy : -ncy;
    }
    const int a = y == y0 ? c0 - y0 * ncx : 0;
    const int e = y == y1 ? c1 - y1 * ncx : ncx - 1;
    int lo = a - 1, hi = e + 1, xa = 0, xb = -1;
    if (hi - lo + 1 >= ncx) {
      if (part == 0) xa = 0, xb = ncx - 1;
    } else if (part == 0) {
      xa = lo < 0 ? 0 : lo;
      xb = hi >= ncx ? ncx - 1 : hi;
    } else if (part == 1) {
      if (lo < 0 && px) xa = ncx + lo, xb = ncx - 1;
    } else {
      if (hi >= ncx && px) xa = 0, xb = hi - ncx;
    }
    if (ok && xb >= xa) {
      const int s0 = cs[r * ncx + xa], s1 = cs[r * ncx + xb + 1];
      if (s1 > s0) rs = s0, re = s1;
    }
  }
  // rank sort by (start, lane)
  int rank = 0;
  for (int u = 0; u < 64; ++u) {
    const int su = __shfl(rs, u, 64);
    rank += (su < rs || (su == rs && u < lane)) ? 1 : 0;
  }
  ss[rank] = rs;
  se[rank] = re;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  rs = ss[lane];
  re = se[lane];
  const bool valid = rs != kEmpty;
  // merge: a range opens a run when it starts past every earlier end
  const int pm_in = wave_incl_max(valid ? re : -1);
  int pm_ex = __shfl_up(pm_in, 1, 64);
  if (lane == 0) pm_ex = -1;
  const bool opens = valid && rs > pm_ex;
  const int run = wave_incl_sum(opens ? 1 : 0) - 1;
  const int R = __shfl(run + 1, 63, 64);
  const bool next_opens = __shfl_down(opens ? 1 : 0, 1, 64) != 0 || lane == 63;
  const bool next_valid = lane < 63 ? __shfl_down(valid ? 1 : 0, 1, 64) != 0 : false;
  const bool closes = valid && (next_opens || !next_valid);
  // run r: start from its opening lane, end (the prefix max) from its closing lane
  __builtin_amdgcn_wave_barrier();
  if (opens && run < kTileMaxRuns) ss[run] = rs;
  if (closes && run < kTileMaxRuns) se[run] = pm_in;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (R > kTileMaxRuns) {
    if (lane == 0) tile_n[b] = -1;
    return;
  }
  const int len = lane < R ? se[lane] - ss[lane] : 0;
  const int tot = wave_incl_sum(len);
  int32_t* tr = tile_runs + (size_t)b * kTileRunStride;
  if (lane == 0) tr[0] = R;
  if (lane < R) {
    tr[1 + 2 * lane] = ss[lane];
    tr[2 + 2 * lane] = len;
  }
  if (lane == 63) tile_n[b] = tot;
}

// Slots of one row's entries, kSlotUnroll at a time: their index loads in
// flight together, then every run (a uniform loop, its bounds read once per
// group) claims the entries it holds.  False when a neighbor lies in no
// run.
template <int U>
__device__ __forceinline__ bool tile_claim(const int* jj, int* sl, const int* run_s, const int* run_e,
                                           const int* run_d, int R) {
#pragma unroll
  for (int u = 0; u < U; ++u) sl[u] = -1;
  for (int r = 0; r < R; ++r) {
    const int s0 = run_s[r], e0 = run_e[r], d0 = run_d[r];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (jj[u] >= s0 && jj[u] < e0) sl[u] = jj[u] + d0;
  }
  bool ok = true;
#pragma unroll
  for (int u = 0; u < U; ++u) ok = ok && sl[u] >= 0;
  return ok;
}

template <class Idx>
__device__ __forceinline__ bool tile_slots_row(const RowRef& rr, const Idx& idx, const int* run_s, const int* run_e,
                                               const int* run_d, int R, uint16_t* __restrict__ slot) {
  constexpr int kSlotUnroll = 8;
  bool ok = true;
  int c = 0, k = rr.base;
  for (; c + kSlotUnroll <= rr.cnt; c += kSlotUnroll, k += kSlotUnroll * rr.step) {
    int jj[kSlotUnroll], sl[kSlotUnroll];
#pragma unroll
    for (int u = 0; u < kSlotUnroll; ++u) jj[u] = idx(k + u * rr.step);
    ok = tile_claim<kSlotUnroll>(jj, sl, run_s, run_e, run_d, R) && ok;
#pragma unroll
    for (int u = 0; u < kSlotUnroll; ++u) slot[k + u * rr.step] = (uint16_t)sl[u];
  }
  for (; c < rr.cnt; ++c, k += rr.step) {
    int j = idx(k), sl;
    ok = tile_claim<1>(&j, &sl, run_s, run_e, run_d, R) && ok;
    slot[k] = (uint16_t)sl;
  }
  return ok;
}

// Entry slots of tile b: the position of each entry's neighbor in the runs
// laid end to end.  The run of the previous entry is tried first (a row's
// entries come a run at a time).  An entry outside every run marks the
// tile unusable (tile_n = -1).
__global__ void __launch_bounds__(kBlock) k_tile_slots(PairsDev P, int cap, uint16_t* __restrict__ slot,
                                                       int64_t slot_cap, const int32_t* __restrict__ tile_runs,
                                                       int32_t* __restrict__ tile_n) {
  __shared__ int run_s[kTileMaxRuns], run_e[kTileMaxRuns], run_d[kTileMaxRuns];
  const int b = blockIdx.x, t = threadIdx.x;
  const int S = tile_n[b];
  if (S < 0 || S > cap) return;  // uniform
  const int R = load_run_table(tile_runs, b, run_s, run_e, run_d);
  __syncthreads();
  const int i = b * kBlock + t;
  if (i >= P.nq) return;
  const RowRef rr = row_ref(P, i);
  if (rr.cnt > 0 && (rr.base < 0 || (int64_t)rr.base + (int64_t)(rr.cnt - 1) * rr.step >= slot_cap)) {
    tile_n[b] = -1;
    return;
  }
  const bool bad = use_idx16(P) ? !tile_slots_row(rr, Idx16{P.nbr16, i, P.nr}, run_s, run_e, run_d, R, slot)
                                 : !tile_slots_row(rr, Idx32{P.nbr}, run_s, run_e, run_d, R, slot);
  if (bad) tile_n[b] = -1;
}

// Loads tile lb's 16-byte records into LDS.
__device__ __forceinline__ bool load_tile(const PairsDev& P, int lb, const float4* __restrict__ rec, float4* tile) {
  return load_tile_by(P, lb, [&](int j) { return rec[j]; }, tile);
}

// Sweep A with the (x, y, p) records of the workgroup's neighbors in LDS.
__global__ void __launch_bounds__(kBlock) DSPH_TILED_ATTR k_dfsph_sweep_a_tiled(PairsDev P, SweepAArgs A,
                                                                                const SolveCtrl* ctrl) {
  extern __shared__ float4 tile[];
  if (ctrl && ctrl->done) return;
  const int lb = logical_block();
  const int i = lb * kBlock + threadIdx.x;
  if (!load_tile(P, lb, A.xp, tile)) {
    if (i < P.nq) sweep_a_row(P, A, i);
    return;
  }
  if (i >= P.nq) return;
  const float4 own = A.xp[i];
  const float sj = A.diff ? -1.0f : 1.0f;
  float ax = 0.f, ay = 0.f;
  sweep_a_kv(P, P.qpos[i], row_ref(P, i), IdxSlot{P.tslot}, A.kv, GatherXP4{tile}, own.z, sj, ax, ay);
  const float r = A.rho[i];
  const float2 a = make_float2(-ax / r, -ay / r);
  if (A.xa)
    A.xa[i] = make_float4(own.x, own.y, a.x, a.y);
  else
    A.acc[i] = a;
}

// Sweep B with the (x, y, a) records of the workgroup's neighbors in LDS.
__global__ void __launch_bounds__(kBlock) DSPH_TILED_ATTR k_dfsph_sweep_b_tiled(PairsDev P, SweepBArgs B) {
  extern __shared__ float4 tile[];
  __shared__ double sh[kBlock / 64];
  if (B.ctrl && B.ctrl->done) return;
  const int lb = logical_block();
  const int i = lb * kBlock + threadIdx.x;
  double err = 0.0;
  if (load_tile(P, lb, B.xa, tile)) {
    if (i < P.nq) {
      const float4 own = B.xa[i];
      const float2 ai = make_float2(own.z, own.w);
      const float sum = sweep_b_kv(P, P.qpos[i], row_ref(P, i), IdxSlot{P.tslot}, B.kv, GatherXA4{tile}, ai);
      err = sweep_b_finish(B, i, sum, make_float2(own.x, own.y));
    }
  } else if (i < P.nq) {
    err = sweep_b_row(P, B, i);
  }
  double bs = block_sum_d(err, sh);
  if (threadIdx.x == 0) B.partials[lb] = bs;
}

// Divergence source (as k_dfsph_source) with the (x, y, v*) records of the
// workgroup's neighbors in LDS.
__global__ void __launch_bounds__(kBlock) DSPH_TILED_ATTR k_dfsph_source_tiled(
    PairsDev P, const float* __restrict__ kv, const float2* __restrict__ vel, const float2* __restrict__ adv,
    const float* __restrict__ vol, const float* __restrict__ areas, float dt, int density, float* __restrict__ src,
    const float4* __restrict__ rec) {
  extern __shared__ float4 tile[];
  const int b = blockIdx.x;
  const int i = b * kBlock + threadIdx.x;
  const bool tiled = load_tile(P, b, rec, tile);
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float2 a_i = adv ? adv[i] : make_float2(0.f, 0.f);
  const float2 vi = make_float2(vel[i].x + dt * a_i.x, vel[i].y + dt * a_i.y);
  const RowRef rr = row_ref(P, i);
  float dv;
  if (tiled)
    dv = source_kv(P, xi, rr, IdxSlot{P.tslot}, kv, GatherXV4{tile}, vi);
  else if (use_idx16(P))
    dv = source_kv(P, xi, rr, Idx16{P.nbr16, i, P.nr}, kv, GatherXV4{rec}, vi);
  else
    dv = source_kv(P, xi, rr, Idx32{P.nbr}, kv, GatherXV4{rec}, vi);
  src[i] = density ? (1.0f - areas[i] / vol[i]) - dt * dv : dv;
}

// The warm-start feature sums (as k_ws_features with W_ij streamed and the
// (V, p_div_prev, s, p_den_prev) records): four entries' loads in flight,
// then their sums in entry order.
template <class Idx>
__device__ __forceinline__ void ws_feature_sums(const RowRef& rr, const Idx& idx, const float4* rec,
                                                const float* __restrict__ pair_w, float& s1, float& s3, float& s5) {
  int n = 0, k = rr.base;
  for (; n + 4 <= rr.cnt; n += 4, k += 4 * rr.step) {
    int j[4];
    float w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      j[u] = idx(k + u * rr.step);
      w[u] = __builtin_nontemporal_load(pair_w + k + u * rr.step);
    }
    float4 r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) r[u] = rec[j[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float vw = r[u].x * w[u];
      s1 += vw * r[u].y;
      s3 += vw * r[u].z;
      s5 += vw * r[u].w;
    }
  }
  for (; n < rr.cnt; ++n, k += rr.step) {
    const float4 r = rec[idx(k)];
    const float vw = r.x * __builtin_nontemporal_load(pair_w + k);
    s1 += vw * r.y;
    s3 += vw * r.z;
    s5 += vw * r.w;
  }
}

__global__ void __launch_bounds__(kBlock) DSPH_TILED_ATTR k_ws_features_tiled(
    PairsDev P, const float* __restrict__ rho, const float* __restrict__ pdp, const float* __restrict__ pnp,
    const float* __restrict__ src, const float* __restrict__ alpha, float rho0, const float* __restrict__ pair_w,
    const float* __restrict__ mean, const float* __restrict__ istd, float* __restrict__ X,
    const float4* __restrict__ rec, float4* __restrict__ ring) {
  extern __shared__ float4 tile[];
  const int b = blockIdx.x;
  const int i = b * kBlock + threadIdx.x;
  const bool tiled = load_tile(P, b, rec, tile);
  if (i >= P.nq) return;
  float s1 = 0.f, s3 = 0.f, s5 = 0.f;
  const RowRef rr = row_ref(P, i);
  if (tiled)
    ws_feature_sums(rr, IdxSlot{P.tslot}, tile, pair_w, s1, s3, s5);
  else if (use_idx16(P))
    ws_feature_sums(rr, Idx16{P.nbr16, i, P.nr}, rec, pair_w, s1, s3, s5);
  else
    ws_feature_sums(rr, Idx32{P.nbr}, rec, pair_w, s1, s3, s5);
  float f[8] = {pdp[i], s1, src[i], s3, pnp[i], s5, rho[i] / rho0 - 1.0f, alpha[i]};
  float* o = X + (int64_t)i * 16;
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = (f[c] - mean[c]) * istd[c];
#pragma unroll
  for (int c = 8; c < 16; ++c) o[c] = 0.0f;
  if (ring) ring[i] = make_float4(rec[i].x, s1, s3, s5);
}

// Second-ring warm-start features: the first-ring pass wrote (V, raw sums
// of p_div_prev, s, p_den_prev) per particle into `ring`; k_ws_ring sums
// those over the row exactly as the feature pass sums its records
// (ws_feature_sums, W_ij streamed) and writes them, normalised, to columns
// 8-10 of the row ((f - mean16[c]) * istd16[c] for c < nfeat, else 0).
template <bool kTiled>
__global__ void __launch_bounds__(kBlock) k_ws_ring(PairsDev P, const float* __restrict__ pair_w,
                                                    const float4* __restrict__ ring, const float* __restrict__ mean,
                                                    const float* __restrict__ istd, int nfeat, float* __restrict__ X) {
  extern __shared__ float4 tile[];
  const int b = blockIdx.x;
  const int i = b * kBlock + threadIdx.x;
  bool tiled = false;
  if (kTiled) tiled = load_tile(P, b, ring, tile);
  if (i >= P.nq) return;
  float t[3] = {0.f, 0.f, 0.f};
  const RowRef rr = row_ref(P, i);
  if (tiled)
    ws_feature_sums(rr, IdxSlot{P.tslot}, tile, pair_w, t[0], t[1], t[2]);
  else if (use_idx16(P))
    ws_feature_sums(rr, Idx16{P.nbr16, i, P.nr}, ring, pair_w, t[0], t[1], t[2]);
  else
    ws_feature_sums(rr, Idx32{P.nbr}, ring, pair_w, t[0], t[1], t[2]);
  float* o = X + (int64_t)i * 16;
#pragma unroll
  for (int c = 8; c < 11; ++c) o[c] = c < nfeat ? (t[c - 8] - mean[c]) * istd[c] : 0.0f;
}

// Sweep launches: over neighbor tiles when the list carries them and the
// sweep gathers records.
static void launch_sweep_a(const PairsDev& P, const SweepAArgs& A, const SolveCtrl* ctrl, int nb, hipStream_t st) {
  if (P.tslot && A.kv && A.xp)
    k_dfsph_sweep_a_tiled<<<nb, kBlock, (size_t)P.tcap * sizeof(float4), st>>>(P, A, ctrl);
  else
    k_dfsph_sweep_a<<<nb, kBlock, 0, st>>>(P, A.kv, A.vol, A.rho, A.p, A.acc, ctrl, A.diff, A.xp, A.xa);
}

static void launch_sweep_b(const PairsDev& P, const SweepBArgs& B, int nb, hipStream_t st) {
  if (P.tslot && B.kv && B.xa)
    k_dfsph_sweep_b_tiled<<<nb, kBlock, (size_t)P.tcap * sizeof(float4), st>>>(P, B);
  else
    k_dfsph_sweep_b<<<nb, kBlock, 0, st>>>(P, B);
}

// One block: fixed-order sum of the sweep-B partials, the reference's loop
// condition (dfsph.py:88), `done` for the kernels queued after it.
__global__ void __launch_bounds__(kBlock) k_dfsph_reduce(const double* __restrict__ partials, int nb, int n,
                                                         SolveCtrl* ctrl) {
  __shared__ double sh[kBlock / 64];
  if (ctrl->done) return;
  reduce_partials(partials, nb, n, ctrl, sh);
}

__global__ void k_ctrl_init(SolveCtrl* c, int min_iters, int max_iters, double thr) {
  c->iter = 0;
  c->ticket = 0;
  c->min_iters = min_iters;
  c->max_iters = max_iters;
  c->err = 0.0;
  c->err0 = 0.0;
  c->thr = thr;
  bool go = 0 < max_iters && (0 < min_iters || 0.0 > thr);
  c->done = go ? 0 : 1;
}

// ---------------------------------------------------------------------------
// Input features of the learned warm start: per particle i
//   0 p_div_prev_i     1 sum_j V_j W_ij p_div_prev_j
//   2 s_i (div source) 3 sum_j V_j W_ij s_j
//   4 p_den_prev_i     5 sum_j V_j W_ij p_den_prev_j
//   6 rho_i/rho0 - 1   7 alpha_div_i
// normalised as (f - mean) * istd; columns 8..15 are zero.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_ws_features(PairsDev P, const float* __restrict__ vol,
                                                        const float* __restrict__ rho, const float* __restrict__ pdp,
                                                        const float* __restrict__ pnp,
                                                        const float* __restrict__ src,
                                                        const float* __restrict__ alpha, float rho0,
                                                        const float* __restrict__ pair_w,
                                                        const float* __restrict__ mean,
                                                        const float* __restrict__ istd, float* __restrict__ X,
                                                        const float4* __restrict__ rec, float4* __restrict__ ring) {
  // rec (with pair_w): (V, p_div_prev, s, p_den_prev) per particle, so that
  // an entry costs one gather instead of four; ring (may be NULL): receives
  // (V, raw sums of p_div_prev, s, p_den_prev), the second ring's records
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const KSel ks = ksel(P);
  float s1 = 0.f, s3 = 0.f, s5 = 0.f;
  const RowRef rr = row_ref(P, i);
  if (pair_w) {
    // W_ij streamed from the density pass; same sums, same order, four
    // entries' gathers in flight
    auto body = [&](float w, int j) {
      const float vw = vol[j] * w;
      s1 += vw * pdp[j];
      s3 += vw * src[j];
      s5 += vw * pnp[j];
    };
    const bool i16 = use_idx16(P);
    int n = 0, k = rr.base;
    for (; n + 4 <= rr.cnt; n += 4, k += 4 * rr.step) {
      int j[4];
      float w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kk = k + u * rr.step;
        j[u] = i16 ? Idx16{P.nbr16, i, P.nr}(kk) : Idx32{P.nbr}(kk);
        w[u] = __builtin_nontemporal_load(pair_w + kk);
      }
      float v[4], a[4], b[4], c[4];
      if (rec) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float4 r = rec[j[u]];
          v[u] = r.x;
          a[u] = r.y;
          b[u] = r.z;
          c[u] = r.w;
        }
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          v[u] = vol[j[u]];
          a[u] = pdp[j[u]];
          b[u] = src[j[u]];
          c[u] = pnp[j[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float vw = v[u] * w[u];
        s1 += vw * a[u];
        s3 += vw * b[u];
        s5 += vw * c[u];
      }
    }
    for (; n < rr.cnt; ++n, k += rr.step)
      body(__builtin_nontemporal_load(pair_w + k), i16 ? Idx16{P.nbr16, i, P.nr}(k) : Idx32{P.nbr}(k));
  } else
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    const int j = P.nbr[k];
    const float hj = h_of(P.rh, j, P.h0);
    const float h = pick_h(P.mode, hi, hj);
    const float2 xj = P.rpos[j];
    const float dx = mod_offset(xi.x - xj.x, P.dom, 0), dy = mod_offset(xi.y - xj.y, P.dom, 1);
    const KConst kk = ks(h);
    const float q = sqrt_rn(dist2_exact(dx, dy)) * kk.ih;
    const float vw = vol[j] * wendland_w(q, kk);
    s1 += vw * pdp[j];
    s3 += vw * src[j];
    s5 += vw * pnp[j];
  }
  float f[8] = {pdp[i], s1, src[i], s3, pnp[i], s5, rho[i] / rho0 - 1.0f, alpha[i]};
  float* o = X + (int64_t)i * 16;
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = (f[c] - mean[c]) * istd[c];
#pragma unroll
  for (int c = 8; c < 16; ++c) o[c] = 0.0f;
  if (ring) ring[i] = make_float4(vol[i], s1, s3, s5);
}

// ---------------------------------------------------------------------------
// Density error (fp64): mean and max |rho/rho0 - 1|.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_density_error(const float* __restrict__ rho, int n, float rho0,
                                                          double* __restrict__ partial_sum,
                                                          double* __restrict__ partial_max) {
  __shared__ double sh[kBlock / 64];
  __shared__ double shm[kBlock / 64];
  double s = 0.0, m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double e = fabs((double)rho[i] / (double)rho0 - 1.0);
    s += e;
    m = fmax(m, e);
  }
  double bs = block_sum_d(s, sh);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mm = fmax(mm, shm[w]);
    partial_sum[blockIdx.x] = bs;
    partial_max[blockIdx.x] = mm;
  }
}

__global__ void k_density_error_final(const double* ps, const double* pm, int nb, int n, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0, m = 0.0;
  for (int k = 0; k < nb; ++k) {
    s += ps[k];
    m = fmax(m, pm[k]);
  }
  out[0] = n > 0 ? s / (double)n : 0.0;
  out[1] = m;
}

// ---------------------------------------------------------------------------
// Semi-implicit Euler (integration.py:111-138).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_integrate_sie(float2* __restrict__ pos, float2* __restrict__ vel,
                                                          const float2* __restrict__ a0, const float2* __restrict__ a1,
                                                          const float2* __restrict__ a2, const float2* __restrict__ a3,
                                                          float gx, float gy, int has_g, int n, float dt,
                                                          float2* __restrict__ dudt_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // dudt = a_p + g + visc + a_div, summed left to right (dfsph.py:209)
  float2 d = a0 ? a0[i] : make_float2(0.f, 0.f);
  if (has_g) {
    d.x = d.x + gx;
    d.y = d.y + gy;
  }
  if (a1) {
    d.x = d.x + a1[i].x;
    d.y = d.y + a1[i].y;
  }
  if (a2) {
    d.x = d.x + a2[i].x;
    d.y = d.y + a2[i].y;
  }
  if (a3) {
    d.x = d.x + a3[i].x;
    d.y = d.y + a3[i].y;
  }
  float2 v = vel[i];
  v.x = v.x + d.x * dt;
  v.y = v.y + d.y * dt;
  float2 x = pos[i];
  x.x = x.x + v.x * dt;
  x.y = x.y + v.y * dt;
  vel[i] = v;
  pos[i] = x;
  if (dudt_out) dudt_out[i] = d;
}

inline size_t al256(size_t b) { return (b + 255) / 256 * 256; }

// (x, y, p) records of the solve loop: x from the reference positions.
__global__ void __launch_bounds__(kBlock) k_pack_xp(const float2* __restrict__ rpos, const float* __restrict__ p, int n,
                                                    float4* __restrict__ xp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 x = rpos[i];
  xp[i] = make_float4(x.x, x.y, p[i], 0.f);
}

// Solve workspace: control block, block partials, source, then the two
// record arrays.
struct SolveWs {
  SolveCtrl* ctrl;
  double* partials;
  float* src;
  float4* xp;
  float4* xa;
};
SolveWs solve_ws(void* ws, int64_t n) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const size_t nb = (size_t)grid_for(n > 0 ? n : 1);
  char* w = static_cast<char*>(ws);
  SolveWs r;
  r.ctrl = reinterpret_cast<SolveCtrl*>(w);
  w += al256(sizeof(SolveCtrl));
  r.partials = reinterpret_cast<double*>(w);
  w += al256(nb * sizeof(double));
  r.src = reinterpret_cast<float*>(w);
  w += al256(sizeof(float) * 2 * nn);
  r.xp = reinterpret_cast<float4*>(w);
  w += al256(sizeof(float4) * nn);
  r.xa = reinterpret_cast<float4*>(w);
  return r;
}

SweepBArgs make_sweep_b(const dsph_dfsph_params* prm, const float* kv, const float* vol, const float* src,
                        const float* alpha, float* p, float* residual, const float* accel, double* partials,
                        SolveCtrl* ctrl, int64_t n) {
  SweepBArgs B;
  B.kv = kv;
  B.vol = vol;
  B.src = src;
  B.alpha = alpha;
  B.p = p;
  B.residual = residual;
  B.acc = reinterpret_cast<const float2*>(accel);
  B.partials = partials;
  B.ctrl = ctrl;
  B.dtf = prm->source_density ? (float)((double)prm->dt * (double)prm->dt) : prm->dt;
  B.omega = prm->omega;
  B.rho0 = prm->rho0;
  B.neg_thr = -prm->error_threshold;
  B.clamp = prm->clamp_pressure;
  B.n = (int)n;
  B.xa = nullptr;
  B.xp = nullptr;
  B.isph = prm->scheme == 1;
  return B;
}

}  // namespace dsph

using namespace dsph;

// 16-bit modular neighbor deltas (dsph_neighbor_compact16): one thread per
// row, entries in the list's own layout.
__global__ void k_ovf_clear(int32_t* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

__global__ void __launch_bounds__(kBlock) k_compact16(PairsDev P, int limit, int16_t* __restrict__ out,
                                                      int64_t cap, int32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const RowRef rr = row_ref(P, i);
  const int n = P.nr, half = P.nr / 2;
  bool bad = false;
  for (int t = 0, k = rr.base; t < rr.cnt; ++t, k += rr.step) {
    if (k < 0 || k >= cap) {
      bad = true;
      continue;
    }
    int d = P.nbr[k] - i;
    if (d > half) d -= n;
    if (d < -half) d += n;
    if (d > limit || d < -limit) {
      bad = true;
      d = 0;
    }
    out[k] = (int16_t)d;
  }
  if (bad) *flag = 1;
}

extern "C" int dsph_neighbor_compact16(const dsph_pairs* pairs, int32_t limit, int16_t* out, int64_t capacity,
                                       int32_t* overflow, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (limit < 0 || limit > 32767 || capacity < 0) return DSPH_ERR_ARG;
  if (!overflow) return DSPH_ERR_ARG;
  hipStream_t st = as_stream(stream);
  k_ovf_clear<<<1, 64, 0, st>>>(overflow);
  DSPH_LAUNCH_CHECK();
  if (pairs->nq == 0) return DSPH_OK;
  if (!out || !pairs->nbr) return DSPH_ERR_ARG;
  k_compact16<<<grid_for(pairs->nq), kBlock, 0, st>>>(make_pairs(*pairs), limit, out, capacity, overflow);
  return launch_status();
}

namespace dsph {
int launch_tile_runs(int64_t nq, const dsph_grid* grid, const dsph_domain* domain, const int32_t* cell_start,
                     int32_t* tile_runs, int32_t* tile_n, hipStream_t st) {
  if (grid->ncell[0] < 1 || grid->ncell[1] < 1 || (int64_t)grid->ncell[0] * grid->ncell[1] >= INT32_MAX)
    return DSPH_ERR_ARG;
  if (nq == 0) return DSPH_OK;
  const int nt = grid_for(nq);
  k_tile_runs<<<(nt + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, st>>>(
      (int)nq, grid->ncell[0], grid->ncell[1], domain->periodic[0] != 0, domain->periodic[1] != 0, cell_start, nt,
      tile_runs, tile_n);
  DSPH_LAUNCH_CHECK();
  return DSPH_OK;
}
}  // namespace dsph

extern "C" int dsph_neighbor_tiles(const dsph_pairs* pairs, const dsph_grid* grid, const int32_t* cell_start,
                                   int32_t cap, uint16_t* slot, int64_t slot_capacity, int32_t* tile_runs,
                                   int32_t* tile_n, void* stream) {
  if (!pairs || !grid) return DSPH_ERR_ARG;
  dsph_pairs q = *pairs;
  q.tile_slot = nullptr;  // the tiles being built are not read
  int s = check_pairs(&q);
  if (s != DSPH_OK) return s;
  if (q.nq != q.nr || cap < 1 || cap > DSPH_TILE_MAX_CAP || slot_capacity < 0) return DSPH_ERR_ARG;
  if (grid->ncell[0] < 1 || grid->ncell[1] < 1 || (int64_t)grid->ncell[0] * grid->ncell[1] >= INT32_MAX)
    return DSPH_ERR_ARG;
  if (q.nq == 0) return DSPH_OK;
  if (!slot || !tile_runs || !tile_n || !cell_start) return DSPH_ERR_ARG;
  hipStream_t st = as_stream(stream);
  const int nt = grid_for(q.nq);
  s = launch_tile_runs(q.nq, grid, &q.domain, cell_start, tile_runs, tile_n, st);
  if (s != DSPH_OK) return s;
  k_tile_slots<<<nt, kBlock, 0, st>>>(make_pairs(q), cap, slot, slot_capacity, tile_runs, tile_n);
  return launch_status();
}

extern "C" int dsph_dfsph_prepare(const dsph_pairs* pairs, const float* vel, const float* mass, const float* areas,
                                  const float* vol, const float* rho, float dt, float rho0,
                                  const dsph_visc_params* visc, float* alpha_div, float* alpha_den, float* div_src,
                                  float* visc_acc, float* kgrad, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!vel || !areas || !vol || !rho || !alpha_div || !alpha_den || !div_src) return DSPH_ERR_ARG;
  if (visc && !visc_acc) return DSPH_ERR_ARG;
  (void)mass;
  PairsDev P = make_pairs(*pairs);
  PrepArgs A;
  A.vel = reinterpret_cast<const float2*>(vel);
  A.areas = areas;
  A.vol = vol;
  A.rho = rho;
  A.dt = dt;
  A.dt2 = (float)((double)dt * (double)dt);
  A.rho0 = rho0;
  A.visc_on = visc != nullptr;
  A.visc_scheme = visc ? visc->scheme : 0;
  A.visc_alpha = visc ? visc->alpha : 0.f;
  A.visc_cs = visc ? visc->cs : 0.f;
  A.visc_rho0 = visc ? visc->rho0 : 0.f;
  A.visc_eps = visc ? visc->eps : 0.f;
  A.pi_switch = visc ? visc->pi_switch : 0;
  A.visc_factor = visc ? (float)(2.0 * (2 + 2) * (double)visc->alpha * (double)visc->rho0) : 0.f;
  if (P.tslot && !P.qh && !P.rh)
    k_dfsph_prepare_tiled<<<grid_for(P.nq), kBlock, (size_t)P.tcap * (sizeof(float4) + sizeof(float2)),
                            as_stream(stream)>>>(P, A, alpha_div, alpha_den, div_src,
                                                 reinterpret_cast<float2*>(visc_acc), kgrad);
  else
    k_dfsph_prepare<<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(
        P, A, alpha_div, alpha_den, div_src, reinterpret_cast<float2*>(visc_acc), kgrad);
  return launch_status();
}

extern "C" size_t dsph_dfsph_workspace(int64_t n) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  size_t nb = (size_t)grid_for(n > 0 ? n : 1);
  return al256(sizeof(SolveCtrl)) + al256(nb * sizeof(double)) + al256(sizeof(float) * 2 * nn) +
         2 * al256(sizeof(float4) * nn);
}

extern "C" int dsph_dfsph_solve(const dsph_pairs* pairs, const dsph_dfsph_params* prm, const float* kgrad,
                                const float* vel, const float* adv, const float* rho, const float* vol,
                                const float* areas, const float* alpha, const float* src0, float* p, float* accel,
                                float* residual, dsph_solve_stats* stats, void* ws, size_t ws_bytes, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (!prm || pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  const int64_t n = pairs->nq;
  if (n == 0) {
    if (stats) {
      stats->iterations = 0;
      stats->converged = 1;
      stats->final_error = 0.0;
      stats->initial_error = 0.0;
    }
    return DSPH_OK;
  }
  if (!vel || !rho || !vol || !areas || !alpha || !p || !accel) return DSPH_ERR_ARG;
  if (!prm->source_density && !src0 && !adv) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_dfsph_workspace(n)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  PairsDev P = make_pairs(*pairs);
  const int nb = grid_for(n);
  const SolveWs W = solve_ws(ws, n);
  SolveCtrl* ctrl = W.ctrl;
  double* partials = W.partials;
  float* src = W.src;
  const float* kv = kgrad;
  // with stored coefficients the loop gathers 16-byte records
  float4* xp = kv ? W.xp : nullptr;
  float4* xa = kv ? W.xa : nullptr;

  // Source term.
  if (prm->source_density || adv) {
    // with stored coefficients the (x, y, v*) records go in the xa slot,
    // which the loop fills only later
    float4* xv = kv ? W.xa : nullptr;
    if (xv) {
      k_pack_xv<<<nb, kBlock, 0, st>>>(reinterpret_cast<const float2*>(P.rpos), reinterpret_cast<const float2*>(vel),
                                        reinterpret_cast<const float2*>(adv), prm->dt, (int)n, xv);
      DSPH_LAUNCH_CHECK();
    }
    if (P.tslot && kv && xv)
      k_dfsph_source_tiled<<<nb, kBlock, (size_t)P.tcap * sizeof(float4), st>>>(
          P, kv, reinterpret_cast<const float2*>(vel), reinterpret_cast<const float2*>(adv), vol, areas, prm->dt,
          prm->source_density, src, xv);
    else
      k_dfsph_source<<<nb, kBlock, 0, st>>>(P, kv, reinterpret_cast<const float2*>(vel),
                                            reinterpret_cast<const float2*>(adv), vol, areas, prm->dt,
                                            prm->source_density, src, xv);
  } else {
    DSPH_CHECK_HIP(hipMemcpyAsync(src, src0, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  }
  DSPH_LAUNCH_CHECK();
  k_ctrl_init<<<1, 1, 0, st>>>(ctrl, prm->min_iters, prm->max_iters, (double)prm->error_threshold);
  DSPH_LAUNCH_CHECK();
  SweepBArgs B = make_sweep_b(prm, kgrad, vol, src, alpha, p, residual, accel, partials, ctrl, n);
  B.xa = xa;
  B.xp = xp;
  if (xp) {
    k_pack_xp<<<nb, kBlock, 0, st>>>(reinterpret_cast<const float2*>(P.rpos), p, (int)n, xp);
    DSPH_LAUNCH_CHECK();
  }

  const int chunk = prm->check_every > 0 ? prm->check_every : 8;
  SolveCtrl host;
  int launched = 0;
  bool done = prm->max_iters <= 0;
  while (!done && launched < prm->max_iters) {
    int k = chunk;
    if (launched + k > prm->max_iters) k = prm->max_iters - launched;
    for (int t = 0; t < k; ++t) {
      launch_sweep_a(P, SweepAArgs{kv, vol, rho, p, reinterpret_cast<float2*>(accel), B.isph, xp, xa}, ctrl, nb,
                     st);
      launch_sweep_b(P, B, nb, st);
      k_dfsph_reduce<<<1, kBlock, 0, st>>>(partials, nb, (int)n, ctrl);
    }
    DSPH_LAUNCH_CHECK();
    launched += k;
    DSPH_CHECK_HIP(hipMemcpyAsync(&host, ctrl, sizeof(SolveCtrl), hipMemcpyDeviceToHost, st));
    DSPH_CHECK_HIP(hipStreamSynchronize(st));
    done = host.done != 0;
  }
  if (prm->max_iters <= 0) {
    DSPH_CHECK_HIP(hipMemcpyAsync(&host, ctrl, sizeof(SolveCtrl), hipMemcpyDeviceToHost, st));
    DSPH_CHECK_HIP(hipStreamSynchronize(st));
  }
  // Final pressure acceleration (dfsph.py:105).
  launch_sweep_a(P, SweepAArgs{kv, vol, rho, p, reinterpret_cast<float2*>(accel), B.isph, xp, nullptr}, nullptr, nb,
                 st);
  DSPH_LAUNCH_CHECK();
  if (stats) {
    stats->iterations = host.iter;
    stats->final_error = host.err;
    stats->initial_error = host.err0;
    stats->converged = host.iter > 0 ? (host.err <= (double)prm->error_threshold ? 1 : 0) : 1;
  }
  return DSPH_OK;
}

extern "C" int dsph_pressure_accel(const dsph_pairs* pairs, const float* kgrad, const float* vol, const float* rho,
                                   const float* p, float* accel, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!vol || !rho || !p || !accel) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  launch_sweep_a(P, SweepAArgs{kgrad, vol, rho, p, reinterpret_cast<float2*>(accel), 0, nullptr, nullptr}, nullptr,
                 grid_for(P.nq), as_stream(stream));
  return launch_status();
}

extern "C" int dsph_pressure_update(const dsph_pairs* pairs, const dsph_dfsph_params* prm, const float* kgrad,
                                    const float* vol, const float* src, const float* alpha, const float* accel,
                                    float* p, float* residual, double* block_partials, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (!prm || pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  const int64_t n = pairs->nq;
  if (n == 0) return DSPH_OK;
  if (!src || !alpha || !accel || !p || !block_partials || (!kgrad && !vol)) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  SweepBArgs B = make_sweep_b(prm, kgrad, vol, src, alpha, p, residual, accel, block_partials, nullptr, n);
  launch_sweep_b(P, B, grid_for(n), as_stream(stream));
  return launch_status();
}

extern "C" int dsph_pack_pressure(const dsph_pairs* pairs, const float* p, float* xp, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!p || !xp) return DSPH_ERR_ARG;
  const int n = (int)pairs->nq;
  k_pack_xp<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(reinterpret_cast<const float2*>(pairs->rpos), p, n,
                                                           reinterpret_cast<float4*>(xp));
  return launch_status();
}

extern "C" int dsph_pressure_accel_rec(const dsph_pairs* pairs, const float* kgrad, const float* rho, const float* xp,
                                       float* xa, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!kgrad || !rho || !xp || !xa) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  launch_sweep_a(P,
                 SweepAArgs{kgrad, nullptr, rho, nullptr, nullptr, 0, reinterpret_cast<const float4*>(xp),
                            reinterpret_cast<float4*>(xa)},
                 nullptr, grid_for(P.nq), as_stream(stream));
  return launch_status();
}

extern "C" int dsph_pressure_update_rec(const dsph_pairs* pairs, const dsph_dfsph_params* prm, const float* kgrad,
                                        const float* src, const float* alpha, const float* xa, float* p, float* xp,
                                        float* residual, double* block_partials, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (!prm || pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  const int64_t n = pairs->nq;
  if (n == 0) return DSPH_OK;
  if (!kgrad || !src || !alpha || !xa || !p || !xp || !block_partials) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  SweepBArgs B = make_sweep_b(prm, kgrad, nullptr, src, alpha, p, residual, nullptr, block_partials, nullptr, n);
  B.xa = reinterpret_cast<const float4*>(xa);
  B.xp = reinterpret_cast<float4*>(xp);
  launch_sweep_b(P, B, grid_for(n), as_stream(stream));
  return launch_status();
}

extern "C" int dsph_dfsph_iterate(const dsph_pairs* pairs, const dsph_dfsph_params* prm, const float* kgrad,
                                  const float* rho, const float* vol, const float* alpha, const float* src, float* p,
                                  float* accel, float* residual, int32_t iterations, double* final_error, void* ws,
                                  size_t ws_bytes, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (!prm || pairs->nq != pairs->nr || iterations < 0) return DSPH_ERR_ARG;
  const int64_t n = pairs->nq;
  if (n == 0 || iterations == 0) return DSPH_OK;
  if (!rho || !vol || !alpha || !src || !p || !accel) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_dfsph_workspace(n)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  PairsDev P = make_pairs(*pairs);
  const int nb = grid_for(n);
  const SolveWs W = solve_ws(ws, n);
  SolveCtrl* ctrl = W.ctrl;
  double* partials = W.partials;
  const float* kv = kgrad;
  float4* xp = kv ? W.xp : nullptr;
  float4* xa = kv ? W.xa : nullptr;
  // Fixed count: the loop condition can never stop it early.
  k_ctrl_init<<<1, 1, 0, st>>>(ctrl, iterations, iterations, 0.0);
  SweepBArgs B = make_sweep_b(prm, kgrad, vol, src, alpha, p, residual, accel, partials, ctrl, n);
  B.xa = xa;
  B.xp = xp;
  if (xp) k_pack_xp<<<nb, kBlock, 0, st>>>(reinterpret_cast<const float2*>(P.rpos), p, (int)n, xp);
  for (int t = 0; t < iterations; ++t) {
    launch_sweep_a(P, SweepAArgs{kv, vol, rho, p, reinterpret_cast<float2*>(accel), B.isph, xp, xa}, ctrl, nb, st);
    launch_sweep_b(P, B, nb, st);
    k_dfsph_reduce<<<1, kBlock, 0, st>>>(partials, nb, (int)n, ctrl);
  }
  DSPH_LAUNCH_CHECK();
  if (final_error) {
    SolveCtrl host;
    DSPH_CHECK_HIP(hipMemcpyAsync(&host, ctrl, sizeof(SolveCtrl), hipMemcpyDeviceToHost, st));
    DSPH_CHECK_HIP(hipStreamSynchronize(st));
    *final_error = host.err;
  }
  return DSPH_OK;
}

extern "C" int dsph_density_error(const float* rho, int64_t n, float rho0, double* out2, void* ws, size_t ws_bytes,
                                  void* stream) {
  if (!out2 || n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if (n > 0 && !rho) return DSPH_ERR_ARG;
  const int nb = 1024;
  size_t need = al256(sizeof(double) * nb) * 2 + al256(sizeof(double) * 2);
  if (!ws || ws_bytes < need) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  char* w = static_cast<char*>(ws);
  double* ps = reinterpret_cast<double*>(w);
  double* pm = reinterpret_cast<double*>(w + al256(sizeof(double) * nb));
  double* out = reinterpret_cast<double*>(w + 2 * al256(sizeof(double) * nb));
  k_density_error<<<nb, kBlock, 0, st>>>(rho, (int)n, rho0, ps, pm);
  k_density_error_final<<<1, 1, 0, st>>>(ps, pm, nb, (int)n, out);
  DSPH_LAUNCH_CHECK();
  DSPH_CHECK_HIP(hipMemcpyAsync(out2, out, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
  DSPH_CHECK_HIP(hipStreamSynchronize(st));
  return DSPH_OK;
}

extern "C" int dsph_integrate_semi_implicit(float* pos, float* vel, const float* a0, const float* a1, const float* a2,
                                            const float* a3, const float* gravity2, int64_t n, float dt,
                                            float* dudt_out, void* stream) {
  if (n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  if (!pos || !vel) return DSPH_ERR_ARG;
  float gx = 0.f, gy = 0.f;
  int has_g = 0;
  if (gravity2) {
    gx = gravity2[0];
    gy = gravity2[1];
    has_g = 1;
  }
  k_integrate_sie<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(
      reinterpret_cast<float2*>(pos), reinterpret_cast<float2*>(vel), reinterpret_cast<const float2*>(a0),
      reinterpret_cast<const float2*>(a1), reinterpret_cast<const float2*>(a2), reinterpret_cast<const float2*>(a3),
      gx, gy, has_g, (int)n, dt, reinterpret_cast<float2*>(dudt_out));
  return launch_status();
}

__global__ void __launch_bounds__(kBlock) k_pack_features(int n, const float* __restrict__ vol,
                                                          const float* __restrict__ pdp, const float* __restrict__ src,
                                                          const float* __restrict__ pnp, float4* __restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rec[i] = make_float4(vol[i], pdp[i], src[i], pnp[i]);
}

namespace dsph {
static int launch_ws_features(const PairsDev& P, const float* vol, const float* rho, const float* pdp,
                              const float* pnp, const float* src, const float* alpha, float rho0, const float* pair_w,
                              const float* mean16, const float* istd16, float* X, const float4* rec, float4* ring,
                              hipStream_t st) {
  if (P.tslot && rec)
    k_ws_features_tiled<<<grid_for(P.nq), kBlock, (size_t)P.tcap * sizeof(float4), st>>>(
        P, rho, pdp, pnp, src, alpha, rho0, pair_w, mean16, istd16, X, rec, ring);
  else
    k_ws_features<<<grid_for(P.nq), kBlock, 0, st>>>(P, vol, rho, pdp, pnp, src, alpha, rho0, pair_w, mean16,
                                                     istd16, X, rec, ring);
  return launch_status();
}
}  // namespace dsph

extern "C" int dsph_ws_features(const dsph_pairs* pairs, const float* vol, const float* rho, const float* p_div_prev,
                                const float* p_den_prev, const float* div_src, const float* alpha_div, float rho0,
                                const float* pair_w, const float* mean16, const float* istd16, float* X,
                                float* records, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!vol || !rho || !p_div_prev || !p_den_prev || !div_src || !alpha_div || !mean16 || !istd16 || !X)
    return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  hipStream_t st = as_stream(stream);
  float4* rec = (records && pair_w) ? reinterpret_cast<float4*>(records) : nullptr;
  if (rec) {
    k_pack_features<<<grid_for(P.nq), kBlock, 0, st>>>(P.nq, vol, p_div_prev, div_src, p_den_prev, rec);
    DSPH_LAUNCH_CHECK();
  }
  return launch_ws_features(P, vol, rho, p_div_prev, p_den_prev, div_src, alpha_div, rho0, pair_w, mean16, istd16,
                            X, rec, nullptr, st);
}

extern "C" int dsph_ws_features_ring(const dsph_pairs* pairs, const float* vol, const float* rho,
                                     const float* p_div_prev, const float* p_den_prev, const float* div_src,
                                     const float* alpha_div, float rho0, const float* pair_w, const float* mean16,
                                     const float* istd16, int32_t nfeat, float* X, float* records,
                                     float* ring_records, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq != pairs->nr || nfeat < 11 || nfeat > 16) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!vol || !rho || !p_div_prev || !p_den_prev || !div_src || !alpha_div || !mean16 || !istd16 || !X ||
      !pair_w || !records || !ring_records)
    return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  hipStream_t st = as_stream(stream);
  float4* rec = reinterpret_cast<float4*>(records);
  float4* ring = reinterpret_cast<float4*>(ring_records);
  k_pack_features<<<grid_for(P.nq), kBlock, 0, st>>>(P.nq, vol, p_div_prev, div_src, p_den_prev, rec);
  DSPH_LAUNCH_CHECK();
  s = launch_ws_features(P, vol, rho, p_div_prev, p_den_prev, div_src, alpha_div, rho0, pair_w, mean16, istd16, X,
                         rec, ring, st);
  if (s != DSPH_OK) return s;
  const int nb = grid_for(P.nq);
  if (P.tslot)
    k_ws_ring<true><<<nb, kBlock, (size_t)P.tcap * sizeof(float4), st>>>(P, pair_w, ring, mean16, istd16, nfeat, X);
  else
    k_ws_ring<false><<<nb, kBlock, 0, st>>>(P, pair_w, ring, mean16, istd16, nfeat, X);
  return launch_status();
}
