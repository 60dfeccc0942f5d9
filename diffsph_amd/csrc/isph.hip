// ISPH pieces (gfx950): the divergence-free pressure system of
// projectVelocities (v2/modules/incompressible.py:102-133, 148-240) solved
// by BiCGStab (v2/sparse.py:141-230), and the step's closing passes
// (incompressible.py:398-460).
//
// The Laplacian is stored per list entry in the layout of the neighbor
// list itself (sliced ELL or CSR): the SpMV reads entry k's index and value
// from the same position, one 256-byte line per wavefront step in the
// sliced layout.  BiCGStab runs with its loop control on the device: every
// kernel of an iteration returns at once when the control block says the
// solve has ended, and the host reads the control block only every
// `check_every` iterations.  Dot products are float64, summed in one fixed
// tree (wave shuffle-down, four waves in order, then a strided pass over
// the block partials), so results repeat bit for bit.
#include "common.h"

namespace dsph {
namespace {

struct BiCtrl {
  int iter;     // completed iterations
  int done;
  int status;   // 0 |r| < atol, 1 |s| < atol, -10 rho breakdown, -11 omega / rv breakdown, 2 iteration limit
  int sexit;
  int maxiter;
  int zero_b;
  double rtol, atol_in, atol, rhotol;
  double rho, rho_prev, alpha, omega, beta;
  double rnorm, bnorm;
};

__device__ __forceinline__ double wave_tree(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum over a 256-thread block: each wave by the shuffle tree, the four wave
// sums in order; valid in thread 0.
__device__ __forceinline__ double block_tree(double v, double* sh) {
  v = wave_tree(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) t += sh[w];
  __syncthreads();
  return t;
}

// Final sum of `nb` block partials by one block: thread t adds partials
// t, t+256, ... in order, then the block tree.
__device__ __forceinline__ double final_sum(const double* __restrict__ part, int nb, double* sh) {
  double v = 0.0;
  for (int k = threadIdx.x; k < nb; k += kBlock) v += part[k];
  return block_tree(v, sh);
}

// computeLaplacian (incompressible.py:102-133): per entry
//   t_ij = ((2 dt) (m_j / rho_j)) (|gradW_ij| / (q_ij h + 1e-7 h)),
// a_ii = sum_j t_ij, the row's self entry holds -a_ii; the Jacobi
// preconditioner of projectVelocities is -1 / a_ii (:178).
__global__ void __launch_bounds__(kBlock) k_isph_laplacian(PairsDev P, const float* __restrict__ mass,
                                                           const float* __restrict__ rho, float c2dt,
                                                           float* __restrict__ val, float* __restrict__ aii,
                                                           float* __restrict__ precond) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const KSel ks = ksel(P);
  const RowRef rr = row_ref(P, i);
  float a = 0.0f;
  int self_k = -1;
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    const int j = P.nbr[k];
    const float hh = pick_h(P.mode, hi, h_of(P.rh, j, P.h0));
    const Geom g = pair_geom(xi, P.rpos[j], hh, ks(hh), P.dom);
    const float gn = sqrt_rn(dist2_exact(g.gx, g.gy));
    const float quot = g.q * g.h + 1e-7f * g.h;
    const float t = (c2dt * (mass[j] / rho[j])) * (gn / quot);
    val[k] = t;
    a += t;
    if (j == i) self_k = k;
  }
  if (self_k >= 0) val[self_k] = -a;
  if (aii) aii[i] = a;
  if (precond) precond[i] = -1.0f / a;
}

// y_i = sum over row i of val_k x_{j_k} (matvec_sparse_coo, sparse.py:8-10),
// in row order with four entries in flight; V is the vector type (the
// stored float values are promoted to it).
template <class V>
__device__ __forceinline__ V spmv_row(const PairsDev& P, const float* __restrict__ val, const V* __restrict__ x,
                                      int i) {
  const RowRef rr = row_ref(P, i);
  V acc = V(0);
  int n = 0, k = rr.base;
  for (; n + 4 <= rr.cnt; n += 4, k += 4 * rr.step) {
    int j[4];
    float v[4];
    V xj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      j[u] = __builtin_nontemporal_load(P.nbr + k + u * rr.step);
      v[u] = __builtin_nontemporal_load(val + k + u * rr.step);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) xj[u] = x[j[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += (V)v[u] * xj[u];
  }
  for (; n < rr.cnt; ++n, k += rr.step) acc += (V)val[k] * x[P.nbr[k]];
  return acc;
}

__global__ void __launch_bounds__(kBlock) k_spmv(PairsDev P, const float* __restrict__ val,
                                                 const float* __restrict__ x, float* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  y[i] = spmv_row<float>(P, val, x, i);
}

// ---------------------------------------------------------------------------
// BiCGStab (sparse.py:141-230), preconditioner M = diag(precond), vectors of
// type V (float, or double as projectVelocities' solverPrecision=float64,
// incompressible.py:173-200); the float inputs are promoted on entry and x
// is rounded back at the end.
// ---------------------------------------------------------------------------

template <class V>
__global__ void __launch_bounds__(kBlock) k_bicg_load(int n, const float* __restrict__ b, const float* __restrict__ x,
                                                      const float* __restrict__ pre, float xscale,
                                                      V* __restrict__ bv, V* __restrict__ xv, V* __restrict__ mv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bv[i] = (V)b[i];
  // x0 = x0_scale * x in float32 (the caller's warm start), then promoted
  xv[i] = (V)(x[i] * xscale);
  mv[i] = (V)pre[i];
}

template <class V>
__global__ void __launch_bounds__(kBlock) k_bicg_store(int n, const V* __restrict__ xv, float* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (float)xv[i];
}

// r = b - A x; rtilde = r; partials of b.b, r.r.
template <class V>
__global__ void __launch_bounds__(kBlock) k_bicg_start(PairsDev P, const float* __restrict__ val,
                                                       const V* __restrict__ b, const V* __restrict__ x,
                                                       V* __restrict__ r, V* __restrict__ rt,
                                                       double* __restrict__ part_bb, double* __restrict__ part_rr) {
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double bb = 0.0, rr = 0.0;
  if (i < P.nq) {
    const V bi = b[i];
    const V ri = bi - spmv_row<V>(P, val, x, i);
    r[i] = ri;
    rt[i] = ri;
    bb = (double)bi * (double)bi;
    rr = (double)ri * (double)ri;
  }
  const double sb = block_tree(bb, sh);
  const double sr = block_tree(rr, sh);
  if (threadIdx.x == 0) {
    part_bb[blockIdx.x] = sb;
    part_rr[blockIdx.x] = sr;
  }
}

// Loop entry (sparse.py:146-175 and the first test of the loop):
// |b| == 0 -> x = b; atol = max(atol, rtol |b|); |r| < atol -> done;
// rho = rtilde . r (= r . r here); |rho| < rhotol -> breakdown.
__global__ void __launch_bounds__(kBlock) k_bicg_entry(const double* __restrict__ part_bb,
                                                       const double* __restrict__ part_rr, int nb, BiCtrl* c) {
  __shared__ double sh[kBlock / 64];
  const double bb = final_sum(part_bb, nb, sh);
  const double rr = final_sum(part_rr, nb, sh);
  if (threadIdx.x != 0) return;
  c->iter = 0;
  c->sexit = 0;
  c->zero_b = 0;
  c->bnorm = sqrt(bb);
  c->atol = fmax(c->atol_in, c->rtol * c->bnorm);
  c->rnorm = sqrt(rr);
  c->omega = 0.0;
  c->alpha = 0.0;
  c->beta = 0.0;
  if (c->bnorm == 0.0) {
    c->zero_b = 1;
    c->done = 1;
    c->status = 0;
    return;
  }
  c->done = 0;
  if (c->maxiter <= 0) {
    c->done = 1;
    c->status = 2;
    return;
  }
  if (c->rnorm < c->atol) {
    c->done = 1;
    c->status = 0;
    return;
  }
  c->rho = rr;
  if (fabs(c->rho) < c->rhotol) {
    c->done = 1;
    c->status = -10;
  }
}

template <class V>
__global__ void __launch_bounds__(kBlock) k_copy_if_zero_b(const BiCtrl* c, const V* __restrict__ b,
                                                           V* __restrict__ x, int n) {
  if (!c->zero_b) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = b[i];
}

// p = r (first pass) or ((p - omega v) beta) + r; phat = M p.
template <class V>
__global__ void __launch_bounds__(kBlock) k_bicg_p(const BiCtrl* c, int n, const V* __restrict__ r,
                                                   const V* __restrict__ v, const V* __restrict__ M,
                                                   V* __restrict__ p, V* __restrict__ phat) {
  if (c->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V pi;
  if (c->iter == 0) {
    pi = r[i];
  } else {
    const V om = (V)c->omega, be = (V)c->beta;
    pi = ((p[i] - om * v[i]) * be) + r[i];
  }
  p[i] = pi;
  phat[i] = M[i] * pi;
}

// y = A x and the block partial of w . y.
template <class V>
__global__ void __launch_bounds__(kBlock) k_bicg_spmv_dot(PairsDev P, const BiCtrl* c, const float* __restrict__ val,
                                                          const V* __restrict__ x, V* __restrict__ y,
                                                          const V* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  if (c->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < P.nq) {
    const V yi = spmv_row<V>(P, val, x, i);
    y[i] = yi;
    d = (double)w[i] * (double)yi;
  }
  const double s = block_tree(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// t = A shat with the partials of t.s and t.t.
template <class V>
__global__ void __launch_bounds__(kBlock) k_bicg_spmv_t(PairsDev P, const BiCtrl* c, const float* __restrict__ val,
                                                        const V* __restrict__ shat, const V* __restrict__ s,
                                                        V* __restrict__ t, double* __restrict__ part_ts,
                                                        double* __restrict__ part_tt) {
  __shared__ double sh[kBlock / 64];
  if (c->done || c->sexit) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ts = 0.0, tt = 0.0;
  if (i < P.nq) {
    const V ti = spmv_row<V>(P, val, shat, i);
    t[i] = ti;
    ts = (double)ti * (double)s[i];
    tt = (double)ti * (double)ti;
  }
  const double a = block_tree(ts, sh);
  const double b = block_tree(tt, sh);
  if (threadIdx.x == 0) {
    part_ts[blockIdx.x] = a;
    part_tt[blockIdx.x] = b;
  }
}

// rv = rtilde . v; rv == 0 -> breakdown; alpha = rho / rv.
__global__ void __launch_bounds__(kBlock) k_bicg_alpha(const double* __restrict__ part, int nb, BiCtrl* c) {
  __shared__ double sh[kBlock / 64];
  if (c->done) return;
  const double rv = final_sum(part, nb, sh);
  if (threadIdx.x != 0) return;
  if (rv == 0.0) {
    c->done = 1;
    c->status = -11;
    return;
  }
  c->alpha = c->rho / rv;
}

// s = r - alpha v (in r); shat = M s; partial s.s.
template <class V>
__global__ void __launch_bounds__(kBlock) k_bicg_s(const BiCtrl* c, int n, V* __restrict__ r,
                                                   const V* __restrict__ v, const V* __restrict__ M,
                                                   V* __restrict__ shat, double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  if (c->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < n) {
    const V al = (V)c->alpha;
    const V si = r[i] - al * v[i];
    r[i] = si;
    shat[i] = M[i] * si;
    d = (double)si * (double)si;
  }
  const double s = block_tree(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_bicg_sexit(const double* __restrict__ part, int nb, BiCtrl* c) {
  __shared__ double sh[kBlock / 64];
  if (c->done) return;
  const double ss = final_sum(part, nb, sh);
  if (threadIdx.x != 0) return;
  if (sqrt(ss) < c->atol) {
    c->sexit = 1;
    c->rnorm = sqrt(ss);  // the residual the solve ends with
  }
}

__global__ void __launch_bounds__(kBlock) k_bicg_omega(const double* __restrict__ part_ts,
                                                       const double* __restrict__ part_tt, int nb, BiCtrl* c) {
  __shared__ double sh[kBlock / 64];
  if (c->done || c->sexit) return;
  const double ts = final_sum(part_ts, nb, sh);
  const double tt = final_sum(part_tt, nb, sh);
  if (threadIdx.x == 0) c->omega = ts / tt;
}

// x += alpha phat (+ omega shat); r = s - omega t; partials r.r, rtilde.r.
template <class V>
__global__ void __launch_bounds__(kBlock) k_bicg_x(const BiCtrl* c, int n, V* __restrict__ x,
                                                   const V* __restrict__ phat, const V* __restrict__ shat,
                                                   V* __restrict__ r, const V* __restrict__ t,
                                                   const V* __restrict__ rt, double* __restrict__ part_rr,
                                                   double* __restrict__ part_rtr) {
  __shared__ double sh[kBlock / 64];
  if (c->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const V al = (V)c->alpha;
  if (c->sexit) {
    if (i < n) x[i] = x[i] + al * phat[i];
    return;
  }
  double rr = 0.0, rtr = 0.0;
  if (i < n) {
    const V om = (V)c->omega;
    V xi = x[i] + al * phat[i];
    xi = xi + om * shat[i];
    x[i] = xi;
    const V ri = r[i] - om * t[i];
    r[i] = ri;
    rr = (double)ri * (double)ri;
    rtr = (double)rt[i] * (double)ri;
  }
  const double a = block_tree(rr, sh);
  const double b = block_tree(rtr, sh);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = a;
    part_rtr[blockIdx.x] = b;
  }
}

// End of an iteration and the tests that open the next one
// (sparse.py:181-193): |r| < atol, rho breakdown, omega breakdown, beta.
__global__ void __launch_bounds__(kBlock) k_bicg_next(const double* __restrict__ part_rr,
                                                      const double* __restrict__ part_rtr, int nb, BiCtrl* c) {
  __shared__ double sh[kBlock / 64];
  if (c->done) return;
  const double rr = final_sum(part_rr, nb, sh);
  const double rtr = final_sum(part_rtr, nb, sh);
  if (threadIdx.x != 0) return;
  if (c->sexit) {
    c->done = 1;
    c->status = 1;
    return;
  }
  c->rho_prev = c->rho;
  c->iter += 1;
  if (c->iter >= c->maxiter) {
    c->done = 1;
    c->status = 2;
    return;
  }
  c->rnorm = sqrt(rr);
  if (c->rnorm < c->atol) {
    c->done = 1;
    c->status = 0;
    return;
  }
  c->rho = rtr;
  if (fabs(c->rho) < c->rhotol) {
    c->done = 1;
    c->status = -10;
    return;
  }
  if (fabs(c->omega) < c->rhotol) {
    c->done = 1;
    c->status = -11;
    return;
  }
  c->beta = (c->rho / c->rho_prev) * (c->alpha / c->omega);
}

__global__ void k_bicg_ctrl_init(BiCtrl* c, int maxiter, double rtol, double atol) {
  c->maxiter = maxiter;
  c->rtol = rtol;
  c->atol_in = atol;
  // float32 eps squared (sparse.py:166-168, the tolerance of a float32 x)
  c->rhotol = 1.1920928955078125e-07 * 1.1920928955078125e-07;
  c->done = 0;
  c->status = 0;
}

inline size_t a256(size_t b) { return (b + 255) / 256 * 256; }

// Workspace: control block, three partial arrays, ten vectors of up to 8 bytes.
constexpr int kBiVecs = 10;

template <class V>
struct BiWs {
  BiCtrl* ctrl;
  double* part[3];
  V *b, *x, *M, *r, *rt, *p, *phat, *v, *shat, *t;
};

template <class V>
BiWs<V> carve(void* ws, int64_t n) {
  const int nb = grid_for(n);
  char* w = static_cast<char*>(ws);
  BiWs<V> o;
  o.ctrl = reinterpret_cast<BiCtrl*>(w);
  w += a256(sizeof(BiCtrl));
  for (int k = 0; k < 3; ++k) {
    o.part[k] = reinterpret_cast<double*>(w);
    w += a256(sizeof(double) * nb);
  }
  V** vecs[kBiVecs] = {&o.b, &o.x, &o.M, &o.r, &o.rt, &o.p, &o.phat, &o.v, &o.shat, &o.t};
  for (auto* f : vecs) {
    *f = reinterpret_cast<V*>(w);
    w += a256(sizeof(double) * n);
  }
  return o;
}

template <class V>
int bicgstab_run(const PairsDev& P, const float* values, const float* precond, const float* b, float* x,
                 float x0_scale, double rtol, double atol, int maxiter, int check_every, dsph_bicgstab_stats* stats,
                 void* ws, hipStream_t st) {
  const int64_t n = P.nq;
  const int nb = grid_for(n);
  const int ni = (int)n;
  BiWs<V> W = carve<V>(ws, n);
  k_bicg_ctrl_init<<<1, 1, 0, st>>>(W.ctrl, maxiter, rtol, atol);
  k_bicg_load<V><<<nb, kBlock, 0, st>>>(ni, b, x, precond, x0_scale, W.b, W.x, W.M);
  k_bicg_start<V><<<nb, kBlock, 0, st>>>(P, values, W.b, W.x, W.r, W.rt, W.part[0], W.part[1]);
  k_bicg_entry<<<1, kBlock, 0, st>>>(W.part[0], W.part[1], nb, W.ctrl);
  k_copy_if_zero_b<V><<<nb, kBlock, 0, st>>>(W.ctrl, W.b, W.x, ni);
  DSPH_LAUNCH_CHECK();
  const int chunk = check_every > 0 ? check_every : 8;
  BiCtrl host;
  int launched = 0;
  for (;;) {
    DSPH_CHECK_HIP(hipMemcpyAsync(&host, W.ctrl, sizeof(BiCtrl), hipMemcpyDeviceToHost, st));
    DSPH_CHECK_HIP(hipStreamSynchronize(st));
    if (host.done || launched >= maxiter) break;
    int k = chunk;
    if (launched + k > maxiter) k = maxiter - launched;
    for (int t = 0; t < k; ++t) {
      k_bicg_p<V><<<nb, kBlock, 0, st>>>(W.ctrl, ni, W.r, W.v, W.M, W.p, W.phat);
      k_bicg_spmv_dot<V><<<nb, kBlock, 0, st>>>(P, W.ctrl, values, W.phat, W.v, W.rt, W.part[0]);
      k_bicg_alpha<<<1, kBlock, 0, st>>>(W.part[0], nb, W.ctrl);
      k_bicg_s<V><<<nb, kBlock, 0, st>>>(W.ctrl, ni, W.r, W.v, W.M, W.shat, W.part[0]);
      k_bicg_sexit<<<1, kBlock, 0, st>>>(W.part[0], nb, W.ctrl);
      k_bicg_spmv_t<V><<<nb, kBlock, 0, st>>>(P, W.ctrl, values, W.shat, W.r, W.t, W.part[1], W.part[2]);
      k_bicg_omega<<<1, kBlock, 0, st>>>(W.part[1], W.part[2], nb, W.ctrl);
      k_bicg_x<V><<<nb, kBlock, 0, st>>>(W.ctrl, ni, W.x, W.phat, W.shat, W.r, W.t, W.rt, W.part[0], W.part[1]);
      k_bicg_next<<<1, kBlock, 0, st>>>(W.part[0], W.part[1], nb, W.ctrl);
    }
    DSPH_LAUNCH_CHECK();
    launched += k;
  }
  k_bicg_store<V><<<nb, kBlock, 0, st>>>(ni, W.x, x);
  DSPH_LAUNCH_CHECK();
  if (stats) {
    stats->iterations = host.iter;
    stats->status = host.status;
    stats->residual_norm = host.rnorm;
    stats->rhs_norm = host.bnorm;
  }
  return DSPH_OK;
}

// ISPH closing passes (incompressible.py:398-460), first pair pass:
//   a_i = -(sum_j (p*_i + p*_j) V_j gradW_ij) / rho0      ('summation' gradient)
//   shift_i = (-dt^2) (rho_i sum_j m_j (p_i/rho_i^2 + p_j/rho_j^2) gradW_ij) / rho0
//   v'_i = v*_i + a_i dt
struct CloseArgs {
  const float2* vel;
  const float2* adv;
  const float* mass;
  const float* rho;
  const float* pstar;
  const float* pss;
  float2* vprime;
  float2* shift;
  float dt, rho0, neg_dt2;
};

__global__ void __launch_bounds__(kBlock) k_isph_close1(PairsDev P, CloseArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const KSel ks = ksel(P);
  const float psi = A.pstar[i], pqi = A.pss[i], rhoi = A.rho[i];
  const float iri = 1.0f / (rhoi * rhoi);
  float gsx = 0.f, gsy = 0.f, gyx = 0.f, gyy = 0.f;
  const RowRef rr = row_ref(P, i);
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    const int j = P.nbr[k];
    const float hh = pick_h(P.mode, hi, h_of(P.rh, j, P.h0));
    const Geom g = pair_geom(xi, P.rpos[j], hh, ks(hh), P.dom);
    const float mj = A.mass[j], rhoj = A.rho[j];
    const float V = mj / rhoj;
    const float kx = V * g.gx, ky = V * g.gy;
    const float s = psi + A.pstar[j];
    gsx += s * kx;
    gsy += s * ky;
    const float irj = 1.0f / (rhoj * rhoj);
    const float a = pqi * iri + A.pss[j] * irj;
    gyx += a * (mj * g.gx);
    gyy += a * (mj * g.gy);
  }
  const float2 v = A.vel[i], ad = A.adv ? A.adv[i] : make_float2(0.f, 0.f);
  const float vsx = v.x + ad.x * A.dt, vsy = v.y + ad.y * A.dt;
  const float ax = (-gsx) / A.rho0, ay = (-gsy) / A.rho0;
  A.vprime[i] = make_float2(vsx + ax * A.dt, vsy + ay * A.dt);
  A.shift[i] = make_float2((A.neg_dt2 * (rhoi * gyx)) / A.rho0, (A.neg_dt2 * (rhoi * gyy)) / A.rho0);
}

// Second pair pass: gradV = sum_j V_j (v'_i - v'_j) (x) gradW_ij ('difference'),
// u = v' + gradV shift, x = (x + v' dt) + shift.
__global__ void __launch_bounds__(kBlock) k_isph_close2(PairsDev P, const float* __restrict__ mass,
                                                        const float* __restrict__ rho,
                                                        const float2* __restrict__ vprime,
                                                        const float2* __restrict__ shift, float dt,
                                                        float2* __restrict__ pos, float2* __restrict__ vel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const KSel ks = ksel(P);
  const float2 vi = vprime[i];
  float g00 = 0.f, g01 = 0.f, g10 = 0.f, g11 = 0.f;
  const RowRef rr = row_ref(P, i);
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    const int j = P.nbr[k];
    const float hh = pick_h(P.mode, hi, h_of(P.rh, j, P.h0));
    const Geom g = pair_geom(xi, P.rpos[j], hh, ks(hh), P.dom);
    const float V = mass[j] / rho[j];
    const float kx = V * g.gx, ky = V * g.gy;
    const float2 vj = vprime[j];
    const float dx = vi.x - vj.x, dy = vi.y - vj.y;
    g00 += dx * kx;
    g01 += dx * ky;
    g10 += dy * kx;
    g11 += dy * ky;
  }
  const float2 s = shift[i];
  vel[i] = make_float2(vi.x + (g00 * s.x + g01 * s.y), vi.y + (g10 * s.x + g11 * s.y));
  const float2 x = pos[i];
  pos[i] = make_float2((x.x + vi.x * dt) + s.x, (x.y + vi.y * dt) + s.y);
}

// ISPH source of projectVelocities (computeDivergenceSourceTerm,
// incompressible.py:148-165): B_i = rho0 sum_j (m_j / rho_i) (v*_j - v*_i)
// . gradW_ij with v* = v + adv dt ('divergenceConsistent', 'difference'),
// then B -= mean(B) (float64 mean).
__global__ void __launch_bounds__(kBlock) k_isph_source(PairsDev P, const float2* __restrict__ vel,
                                                        const float2* __restrict__ adv,
                                                        const float* __restrict__ mass, const float* __restrict__ rho,
                                                        float dt, float rho0, float* __restrict__ B,
                                                        double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < P.nq) {
    const float2 xi = P.qpos[i];
    const float hi = h_of(P.qh, i, P.h0);
  const KSel ks = ksel(P);
    const float rhoi = rho[i];
    const float2 ai = adv ? adv[i] : make_float2(0.f, 0.f);
    const float vix = vel[i].x + ai.x * dt, viy = vel[i].y + ai.y * dt;
    float acc = 0.f;
    const RowRef rr = row_ref(P, i);
    for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
      const int j = P.nbr[k];
      const float hh = pick_h(P.mode, hi, h_of(P.rh, j, P.h0));
    const Geom g = pair_geom(xi, P.rpos[j], hh, ks(hh), P.dom);
      const float2 aj = adv ? adv[j] : make_float2(0.f, 0.f);
      const float vjx = vel[j].x + aj.x * dt, vjy = vel[j].y + aj.y * dt;
      const float V = mass[j] / rhoi;
      const float kx = V * g.gx, ky = V * g.gy;
      acc += (vjx - vix) * kx + (vjy - viy) * ky;
    }
    const float b = rho0 * acc;
    B[i] = b;
    d = (double)b;
  }
  const double s = block_tree(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_isph_mean(const double* __restrict__ part, int nb, int n,
                                                      float* __restrict__ meanf) {
  __shared__ double sh[kBlock / 64];
  const double tot = final_sum(part, nb, sh);
  if (threadIdx.x == 0) *meanf = (float)(tot / (double)n);
}

__global__ void __launch_bounds__(kBlock) k_isph_center(float* __restrict__ B, int n, const float* __restrict__ meanf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) B[i] = B[i] - *meanf;
}

// alpha / dt of the ISPH Jacobi (incompressible.py:269) and the warm start
// 0.5 p_prev (:335-337).
__global__ void __launch_bounds__(kBlock) k_isph_jacobi_init(int n, const float* __restrict__ alpha_den, float dt,
                                                             float* __restrict__ alpha, const float* __restrict__ pprev,
                                                             float* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  alpha[i] = alpha_den[i] / dt;
  p[i] = pprev ? 0.5f * pprev[i] : 0.0f;
}

}  // namespace
}  // namespace dsph

using namespace dsph;

static int check_square(const dsph_pairs* pairs) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  return DSPH_OK;
}

extern "C" int dsph_isph_laplacian(const dsph_pairs* pairs, const float* mass, const float* rho, float dt,
                                   float* values, float* aii, float* precond, void* stream) {
  int s = check_square(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq == 0) return DSPH_OK;
  if (!mass || !rho || !values) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  const float c2dt = (float)(2.0 * (double)dt);
  k_isph_laplacian<<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(P, mass, rho, c2dt, values, aii, precond);
  return launch_status();
}

extern "C" int dsph_spmv(const dsph_pairs* pairs, const float* values, const float* x, float* y, void* stream) {
  int s = check_square(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq == 0) return DSPH_OK;
  if (!values || !x || !y) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  k_spmv<<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(P, values, x, y);
  return launch_status();
}

extern "C" size_t dsph_bicgstab_workspace(int64_t n) {
  const int64_t m = n > 0 ? n : 1;
  return a256(sizeof(BiCtrl)) + 3 * a256(sizeof(double) * grid_for(m)) + kBiVecs * a256(sizeof(double) * m);
}

extern "C" int dsph_bicgstab(const dsph_pairs* pairs, const float* values, const float* precond, const float* b,
                             float* x, float x0_scale, double rtol, double atol, int32_t maxiter,
                             int32_t check_every, int32_t fp64, dsph_bicgstab_stats* stats, void* ws,
                             size_t ws_bytes, void* stream) {
  int s = check_square(pairs);
  if (s != DSPH_OK) return s;
  if (!(rtol >= 0.0) || !(atol >= 0.0) || maxiter < 0) return DSPH_ERR_ARG;
  const int64_t n = pairs->nq;
  if (n == 0) {
    if (stats) {
      stats->iterations = 0;
      stats->status = 0;
      stats->residual_norm = 0.0;
      stats->rhs_norm = 0.0;
    }
    return DSPH_OK;
  }
  if (!values || !precond || !b || !x) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_bicgstab_workspace(n)) return DSPH_ERR_CAPACITY;
  const PairsDev P = make_pairs(*pairs);
  if (fp64)
    return bicgstab_run<double>(P, values, precond, b, x, x0_scale, rtol, atol, maxiter, check_every, stats, ws,
                                as_stream(stream));
  return bicgstab_run<float>(P, values, precond, b, x, x0_scale, rtol, atol, maxiter, check_every, stats, ws,
                             as_stream(stream));
}

extern "C" size_t dsph_isph_source_workspace(int64_t n) {
  return a256(sizeof(double) * grid_for(n > 0 ? n : 1)) + 256;
}

extern "C" int dsph_isph_source(const dsph_pairs* pairs, const float* vel, const float* adv, const float* mass,
                                const float* rho, float dt, float rho0, float* B, void* ws, size_t ws_bytes,
                                void* stream) {
  int s = check_square(pairs);
  if (s != DSPH_OK) return s;
  const int64_t n = pairs->nq;
  if (n == 0) return DSPH_OK;
  if (!vel || !mass || !rho || !B) return DSPH_ERR_ARG;
  const int nb = grid_for(n);
  if (!ws || ws_bytes < dsph_isph_source_workspace(n)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  PairsDev P = make_pairs(*pairs);
  double* part = static_cast<double*>(ws);
  float* meanf = reinterpret_cast<float*>(static_cast<char*>(ws) + a256(sizeof(double) * nb));
  k_isph_source<<<nb, kBlock, 0, st>>>(P, reinterpret_cast<const float2*>(vel), reinterpret_cast<const float2*>(adv),
                                       mass, rho, dt, rho0, B, part);
  k_isph_mean<<<1, kBlock, 0, st>>>(part, nb, (int)n, meanf);
  k_isph_center<<<nb, kBlock, 0, st>>>(B, (int)n, meanf);
  return launch_status();
}

extern "C" int dsph_isph_jacobi_init(int64_t n, const float* alpha_den, float dt, const float* p_prev, float* alpha,
                                     float* p, void* stream) {
  if (n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  if (!alpha_den || !alpha || !p || !(dt > 0.0f)) return DSPH_ERR_ARG;
  k_isph_jacobi_init<<<grid_for(n), kBlock, 0, as_stream(stream)>>>((int)n, alpha_den, dt, alpha, p_prev, p);
  return launch_status();
}

extern "C" int dsph_isph_close(const dsph_pairs* pairs, const float* mass, const float* rho, const float* vel_in,
                               const float* adv, const float* p_star, const float* p_ss, float dt, float rho0,
                               float* vprime, float* shift, float* pos, float* vel, void* stream) {
  int s = check_square(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq == 0) return DSPH_OK;
  if (!mass || !rho || !vel_in || !p_star || !p_ss || !vprime || !shift || !pos || !vel) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  hipStream_t st = as_stream(stream);
  CloseArgs A;
  A.vel = reinterpret_cast<const float2*>(vel_in);
  A.adv = reinterpret_cast<const float2*>(adv);
  A.mass = mass;
  A.rho = rho;
  A.pstar = p_star;
  A.pss = p_ss;
  A.vprime = reinterpret_cast<float2*>(vprime);
  A.shift = reinterpret_cast<float2*>(shift);
  A.dt = dt;
  A.rho0 = rho0;
  A.neg_dt2 = (float)(-((double)dt * (double)dt));
  k_isph_close1<<<grid_for(P.nq), kBlock, 0, st>>>(P, A);
  k_isph_close2<<<grid_for(P.nq), kBlock, 0, st>>>(P, mass, rho, reinterpret_cast<const float2*>(vprime),
                                                   reinterpret_cast<const float2*>(shift), dt,
                                                   reinterpret_cast<float2*>(pos), reinterpret_cast<float2*>(vel));
  return launch_status();
}
