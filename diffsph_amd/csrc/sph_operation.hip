// The reference's generic SPH operation on per-pair arrays
// (sphOperation, v2/sphOps.py:253-282, with the operators of :6-250) and the
// DFSPH module's per-pair-array pieces (computeAlpha, updatePressure and the
// convergence metric, v2/modules/dfsph.py:7-50,96).
//
// Inputs are the arrays of a reference neighborDict: row-sorted pair list
// (CSR offsets over the rows, int64 j), precomputed kernel values W_ij,
// gradients gradW_ij [P,2], normalised distances r_ij/h_ij, unit vectors
// x_ij [P,2] and supports h_ij.  One thread per row sums its segment in
// list order.  A quantity is either per particle (the reference's (q_i, q_j)
// tuple: qa indexed by i, qb by j) or per pair (a single tensor).
#include "common.h"

namespace dsph {

enum PpOp : int {
  PP_DENSITY = 0,
  PP_INTERPOLATE = 1,
  PP_GRADIENT = 2,
  PP_DIVERGENCE = 3,
  PP_DIVERGENCE_CONSISTENT = 4,
  PP_CURL = 5,
  PP_LAPLACIAN = 6,
};

enum PpMode : int { PM_NAIVE = 0, PM_DIFFERENCE = 1, PM_SUMMATION = 2, PM_SYMMETRIC = 3, PM_CONSERVING = 4,
                    PM_DIVFREE = 5, PM_DOT = 6 };

struct PpArgs {
  int n;          // rows (particles of set A)
  int64_t nb;     // particles of set B
  int64_t np;     // pairs
  const int32_t* off;   // [n+1]
  const int64_t* j;     // [np]
  const float* W;       // [np]
  const float* gW;      // [np,2]
  const float* rij;     // [np]   |x_ij| / h_ij
  const float* xij;     // [np,2] unit vectors
  const float* hij;     // [np]
  const float* lap;     // [np]   kernel Laplacians (naive Laplacian only)
  const float* ma;      // [n]
  const float* mb;      // [nb]
  const float* ra;      // [n]    densities of A
  const float* rb;      // [nb]
  const float* qa;      // [n, qw]   (per-particle quantity) or NULL
  const float* qb;      // [nb, qw]
  const float* qp;      // [np, qw]  (per-pair quantity) or NULL
  int qw;               // components of the quantity: 1, 2 or 4 (a 2x2 tensor, row-major)
  int op, mode, dot;    // dot: divergenceMode 'dot' (contract the first axis)
  float* out;
};

// Quantity difference / sum / neighbour value per the mode (q_i - q_j for a
// gradient difference, q_j - q_i for a divergence difference).
__device__ __forceinline__ void pp_q(const PpArgs& a, int i, int64_t j, int64_t k, int c, bool div_sign, float* q) {
  for (int t = 0; t < a.qw; ++t) {
    if (a.qp) {
      q[t] = a.qp[k * a.qw + t];
      continue;
    }
    const float qi = a.qa ? a.qa[(int64_t)i * a.qw + t] : 0.0f;
    const float qj = a.qb[j * a.qw + t];
    if (a.mode == PM_DIFFERENCE)
      q[t] = div_sign ? qj - qi : qi - qj;
    else if (a.mode == PM_SUMMATION)
      q[t] = div_sign ? qj + qi : qi + qj;
    else
      q[t] = qj;
  }
  (void)c;
}

__global__ void __launch_bounds__(kBlock) k_pp_operation(PpArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = 0.0f;
  int64_t b = a.off[i], e = a.off[i + 1];
  if (b < 0) b = 0;
  if (e > a.np) e = a.np;
  const float rho_i = a.ra ? a.ra[i] : 1.0f;
  const bool sym = a.mode == PM_SYMMETRIC;
  const float inv_ri2 = sym ? 1.0f / (rho_i * rho_i) : 0.0f;
  for (int64_t k = b; k < e; ++k) {
    const int64_t j = a.j[k];
    if (j < 0 || j >= a.nb) continue;
    float q[4] = {0.f, 0.f, 0.f, 0.f};
    switch (a.op) {
      case PP_DENSITY: {
        acc[0] += a.mb[j] * a.W[k];
        break;
      }
      case PP_INTERPOLATE: {
        const float kk = a.mb[j] / a.rb[j] * a.W[k];
        for (int t = 0; t < a.qw; ++t) q[t] = a.qp ? a.qp[k * a.qw + t] : a.qb[j * a.qw + t];
        for (int t = 0; t < a.qw; ++t) acc[t] += q[t] * kk;
        break;
      }
      case PP_GRADIENT:
      case PP_DIVERGENCE:
      case PP_DIVERGENCE_CONSISTENT:
      case PP_CURL: {
        const float gx = a.gW[2 * k], gy = a.gW[2 * k + 1];
        float kx, ky;
        if (sym) {
          // k = m_j gradW; q = q_i / rho_i^2 + q_j / rho_j^2; result scaled by rho_i
          kx = a.mb[j] * gx;
          ky = a.mb[j] * gy;
          const float inv_rj2 = 1.0f / (a.rb[j] * a.rb[j]);
          for (int t = 0; t < a.qw; ++t)
            q[t] = a.qa[(int64_t)i * a.qw + t] * inv_ri2 + a.qb[j * a.qw + t] * inv_rj2;
        } else {
          const float V = a.op == PP_DIVERGENCE_CONSISTENT ? a.mb[j] / rho_i : a.mb[j] / a.rb[j];
          kx = V * gx;
          ky = V * gy;
          pp_q(a, i, j, k, 0, a.op != PP_GRADIENT, q);
        }
        if (a.op == PP_GRADIENT) {
          for (int t = 0; t < a.qw; ++t) {
            acc[2 * t] += q[t] * kx;
            acc[2 * t + 1] += q[t] * ky;
          }
        } else if (a.op == PP_CURL) {
          acc[0] += q[1] * kx - q[0] * ky;
        } else if (a.qw == 2) {
          acc[0] += q[0] * kx + q[1] * ky;
        } else {
          // a 2x2 tensor field: 'div' contracts the last axis, 'dot' the first
          if (!a.dot) {
            acc[0] += q[0] * kx + q[1] * ky;
            acc[1] += q[2] * kx + q[3] * ky;
          } else {
            acc[0] += q[0] * kx + q[2] * ky;
            acc[1] += q[1] * kx + q[3] * ky;
          }
        }
        break;
      }
      case PP_LAPLACIAN: {
        const float V = a.mb[j] / a.rb[j];
        float A[4];
        for (int t = 0; t < a.qw; ++t)
          A[t] = a.qp ? a.qp[k * a.qw + t] : a.qa[(int64_t)i * a.qw + t] - a.qb[j * a.qw + t];
        if (a.mode == PM_NAIVE) {
          const float lk = -V * a.lap[k];
          for (int t = 0; t < a.qw; ++t) acc[t] += lk * A[t];
          break;
        }
        const float h = a.hij[k];
        const float quot = a.rij[k] * h + 1e-7f * h;
        const float ex = a.xij[2 * k], ey = a.xij[2 * k + 1];
        const float gx = a.gW[2 * k], gy = a.gW[2 * k + 1];
        const float kA = (gx * -ex + gy * -ey) / quot;
        if (a.qw == 1) {
          acc[0] += -A[0] * V * 2.0f * kA;
        } else if (a.mode == PM_CONSERVING) {
          const float qq = V * kA * (A[0] * ex + A[1] * ey);
          acc[0] += -qq * ex;
          acc[1] += -qq * ey;
        } else if (a.mode == PM_DIVFREE) {
          const float dd = (A[0] * ex + A[1] * ey) / quot;
          const float qq = 2.0f * 4.0f * V * dd;  // 2 (d + 2), d = 2
          acc[0] += qq * gx;
          acc[1] += qq * gy;
        } else if (a.mode == PM_DOT) {
          const float s = -4.0f * (A[0] * ex + A[1] * ey);  // -(d + 2) (A . x)
          const float f = V * kA;
          acc[0] += (s * ex - A[0]) * f;
          acc[1] += (s * ey - A[1]) * f;
        } else {
          const float qq = -2.0f * V * kA;
          acc[0] += A[0] * qq;
          acc[1] += A[1] * qq;
        }
        break;
      }
      default:
        break;
    }
  }
  int ow = 1;
  if (a.op == PP_INTERPOLATE) ow = a.qw;
  else if (a.op == PP_GRADIENT) ow = 2 * a.qw;
  else if (a.op == PP_DIVERGENCE || a.op == PP_DIVERGENCE_CONSISTENT) ow = a.qw / 2;
  else if (a.op == PP_LAPLACIAN) ow = a.qw;
  const float scale = (sym && a.op != PP_DENSITY && a.op != PP_INTERPOLATE && a.op != PP_LAPLACIAN) ? rho_i : 1.0f;
  for (int t = 0; t < ow; ++t) a.out[(int64_t)i * ow + t] = sym ? scale * acc[t] : acc[t];
}

// Row offsets of a row-sorted pair list: off[r] = first pair of row r;
// *unsorted is set when a row index decreases or lies outside [0, n).
__global__ void __launch_bounds__(kBlock) k_pp_offsets(const int64_t* __restrict__ rows, int64_t np, int n,
                                                       int32_t* __restrict__ off, int32_t* __restrict__ unsorted) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > np) return;
  // rows strictly between the previous pair's row and this one start here
  const int64_t cur = k < np ? rows[k] : (int64_t)n;
  const int64_t prev = k > 0 ? rows[k - 1] : -1;
  if (cur < prev || cur < 0 || cur > n || (k < np && cur == n)) {
    atomicOr(unsorted, 1);
    return;
  }
  for (int64_t r = prev + 1; r <= cur; ++r) off[r] = (int32_t)k;
}

__global__ void k_pp_flag_clear(int32_t* f) { f[0] = 0; }

// computeAlpha (v2/modules/dfsph.py:7-27) on per-pair gradients:
// alpha_i = fac/m_i |sum_j A_j gradW|^2 + fac sum_j A_j^2/(areas_j rho0) |gradW|^2,
// fac = -dt_pow * A_i, A = m / rho (actualArea), m_i = areas_i rho0.
__global__ void __launch_bounds__(kBlock) k_pp_alpha(int n, int64_t nb, int64_t np, const int32_t* __restrict__ off,
                                                     const int64_t* __restrict__ jj, const float* __restrict__ gW,
                                                     const float* __restrict__ act_a, const float* __restrict__ act_b,
                                                     const float* __restrict__ areas_a,
                                                     const float* __restrict__ areas_b, float dt_pow, float rho0,
                                                     float* __restrict__ alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t b = off[i], e = off[i + 1];
  if (b < 0) b = 0;
  if (e > np) e = np;
  float s1x = 0.f, s1y = 0.f, s2 = 0.f;
  for (int64_t k = b; k < e; ++k) {
    const int64_t j = jj[k];
    if (j < 0 || j >= nb) continue;
    const float gx = gW[2 * k], gy = gW[2 * k + 1];
    const float g2 = gx * gx + gy * gy;
    const float Aj = act_b[j];
    s1x += Aj * gx;
    s1y += Aj * gy;
    s2 += (Aj * Aj / (areas_b[j] * rho0)) * g2;
  }
  const float fac = -dt_pow * act_a[i];
  const float mass = areas_a[i] * rho0;
  alpha[i] = fac / mass * (s1x * s1x + s1y * s1y) + fac * s2;
}

// updatePressure (v2/modules/dfsph.py:39-50) with the convergence metric
// (:96): K_i = -dt_pow * div_difference(a)_i; p_new = p + omega (s - K) /
// alpha (clamped at 0); residual = K - s; per-block float64 partial sums of
// max(residual / rho0, -thr) in a fixed order.
__global__ void __launch_bounds__(kBlock) k_pp_update(int n, int64_t nb, int64_t np, const int32_t* __restrict__ off,
                                                      const int64_t* __restrict__ jj, const float* __restrict__ gW,
                                                      const float* __restrict__ mb, const float* __restrict__ rb,
                                                      const float* __restrict__ acc_a,
                                                      const float* __restrict__ acc_b, const float* __restrict__ src,
                                                      const float* __restrict__ alpha, const float* __restrict__ p_in,
                                                      float dt_pow, float omega, int clampp, float rho0, float thr,
                                                      float* __restrict__ p_out, float* __restrict__ residual,
                                                      double* __restrict__ partials) {
  __shared__ double red[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (i < n) {
    int64_t b = off[i], e = off[i + 1];
    if (b < 0) b = 0;
    if (e > np) e = np;
    const float ax = acc_a[2 * (int64_t)i], ay = acc_a[2 * (int64_t)i + 1];
    float sum = 0.f;
    for (int64_t k = b; k < e; ++k) {
      const int64_t j = jj[k];
      if (j < 0 || j >= nb) continue;
      const float V = mb[j] / rb[j];
      sum += (acc_b[2 * j] - ax) * (V * gW[2 * k]) + (acc_b[2 * j + 1] - ay) * (V * gW[2 * k + 1]);
    }
    const float K = -dt_pow * sum;
    const float s = src[i];
    const float r = K - s;
    float p = p_in[i] + omega * (s - K) / alpha[i];
    if (clampp) p = fmaxf(p, 0.0f);
    p_out[i] = p;
    residual[i] = r;
    const float rc = r / rho0;
    contrib = (double)(rc < -thr ? -thr : rc);
  }
  red[threadIdx.x] = contrib;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kBlock) k_pp_mean(const double* __restrict__ partials, int nblk, int n,
                                                    double* __restrict__ out) {
  __shared__ double red[kBlock];
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kBlock) s += partials[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = n > 0 ? red[0] / (double)n : 0.0;
}

}  // namespace dsph

using namespace dsph;

extern "C" int dsph_pp_offsets(const int64_t* rows, int64_t num_pairs, int32_t num_rows, int32_t* offsets,
                               int32_t* unsorted, void* stream) {
  if (num_pairs < 0 || num_pairs >= INT32_MAX || num_rows < 0 || !offsets || !unsorted) return DSPH_ERR_ARG;
  if (num_pairs > 0 && !rows) return DSPH_ERR_ARG;
  hipStream_t st = as_stream(stream);
  k_pp_flag_clear<<<1, 1, 0, st>>>(unsorted);
  DSPH_LAUNCH_CHECK();
  k_pp_offsets<<<grid_for(num_pairs + 1), kBlock, 0, st>>>(rows, num_pairs, num_rows, offsets, unsorted);
  return launch_status();
}

extern "C" int dsph_pp_operation(const dsph_pair_arrays* pa, int32_t op, int32_t mode, int32_t dot_mode,
                                 const float* masses_a, const float* masses_b, const float* densities_a,
                                 const float* densities_b, const float* q_a, const float* q_b, const float* q_pairs,
                                 int32_t q_width, float* out, void* stream) {
  if (!pa || pa->num_rows < 0 || pa->num_ref < 0 || pa->num_pairs < 0 || pa->num_rows >= INT32_MAX)
    return DSPH_ERR_ARG;
  if (op < PP_DENSITY || op > PP_LAPLACIAN || mode < PM_NAIVE || mode > PM_DOT) return DSPH_ERR_ARG;
  if (q_width != 1 && q_width != 2 && q_width != 4) return DSPH_ERR_DIM;
  if (pa->num_rows == 0) return DSPH_OK;
  if (!pa->offsets || !out || !masses_b) return DSPH_ERR_ARG;
  if (pa->num_pairs > 0 && !pa->j) return DSPH_ERR_ARG;
  const bool needs_q = op != PP_DENSITY;
  if (needs_q && !q_pairs && !q_b) return DSPH_ERR_ARG;
  if (op == PP_DENSITY && !pa->kernels && pa->num_pairs > 0) return DSPH_ERR_ARG;
  if (op == PP_INTERPOLATE && (!pa->kernels || !densities_b)) return DSPH_ERR_ARG;
  const bool grad_like = op == PP_GRADIENT || op == PP_DIVERGENCE || op == PP_DIVERGENCE_CONSISTENT || op == PP_CURL;
  if ((grad_like || op == PP_LAPLACIAN) && !pa->gradients && pa->num_pairs > 0) return DSPH_ERR_ARG;
  if (grad_like && mode > PM_SYMMETRIC) return DSPH_ERR_ARG;
  if (grad_like && mode == PM_SYMMETRIC && (!q_a || !q_b || !densities_a || !densities_b)) return DSPH_ERR_ARG;
  if (grad_like && mode != PM_SYMMETRIC && !densities_b) return DSPH_ERR_ARG;
  if (grad_like && !q_pairs && (mode == PM_DIFFERENCE || mode == PM_SUMMATION) && !q_a) return DSPH_ERR_ARG;
  if (op == PP_DIVERGENCE_CONSISTENT && mode != PM_SYMMETRIC && !densities_a) return DSPH_ERR_ARG;
  if ((op == PP_DIVERGENCE || op == PP_DIVERGENCE_CONSISTENT) && q_width < 2) return DSPH_ERR_DIM;
  if (op == PP_CURL && q_width != 2) return DSPH_ERR_DIM;
  if (op == PP_GRADIENT && q_width > 4) return DSPH_ERR_DIM;
  if (op == PP_LAPLACIAN) {
    if (q_width > 2 || !densities_b || (!q_pairs && (!q_a || !q_b))) return DSPH_ERR_ARG;
    if (mode == PM_NAIVE && !pa->laplacians && pa->num_pairs > 0) return DSPH_ERR_ARG;
    if (mode != PM_NAIVE && pa->num_pairs > 0 && (!pa->distances || !pa->vectors || !pa->supports))
      return DSPH_ERR_ARG;
  }
  PpArgs a;
  a.n = (int)pa->num_rows;
  a.nb = pa->num_ref;
  a.np = pa->num_pairs;
  a.off = pa->offsets;
  a.j = pa->j;
  a.W = pa->kernels;
  a.gW = pa->gradients;
  a.rij = pa->distances;
  a.xij = pa->vectors;
  a.hij = pa->supports;
  a.lap = pa->laplacians;
  a.ma = masses_a;
  a.mb = masses_b;
  a.ra = densities_a;
  a.rb = densities_b;
  a.qa = q_a;
  a.qb = q_b;
  a.qp = q_pairs;
  a.qw = q_width;
  a.op = op;
  a.mode = mode;
  a.dot = dot_mode != 0;
  a.out = out;
  k_pp_operation<<<grid_for(a.n), kBlock, 0, as_stream(stream)>>>(a);
  return launch_status();
}

extern "C" int dsph_pp_dfsph_alpha(const dsph_pair_arrays* pa, const float* actual_area_a,
                                   const float* actual_area_b, const float* areas_a, const float* areas_b,
                                   float dt_pow, float rho0, float* alpha, void* stream) {
  if (!pa || pa->num_rows < 0 || pa->num_rows >= INT32_MAX || pa->num_pairs < 0) return DSPH_ERR_ARG;
  if (pa->num_rows == 0) return DSPH_OK;
  if (!pa->offsets || !actual_area_a || !actual_area_b || !areas_a || !areas_b || !alpha) return DSPH_ERR_ARG;
  if (pa->num_pairs > 0 && (!pa->j || !pa->gradients)) return DSPH_ERR_ARG;
  k_pp_alpha<<<grid_for(pa->num_rows), kBlock, 0, as_stream(stream)>>>(
      (int)pa->num_rows, pa->num_ref, pa->num_pairs, pa->offsets, pa->j, pa->gradients, actual_area_a,
      actual_area_b, areas_a, areas_b, dt_pow, rho0, alpha);
  return launch_status();
}

extern "C" size_t dsph_pp_dfsph_update_workspace(int64_t num_rows) {
  const int64_t nblk = (num_rows + kBlock - 1) / kBlock;
  return (size_t)(((nblk > 0 ? nblk : 1) + 1) * sizeof(double) + 255) / 256 * 256;
}

extern "C" int dsph_pp_dfsph_update(const dsph_pair_arrays* pa, const float* masses_b, const float* densities_b,
                                    const float* accel_a, const float* accel_b, const float* source,
                                    const float* alpha, const float* pressure_in, float dt_pow, float omega,
                                    int32_t clamp_pressure, float rho0, float threshold, float* pressure_out,
                                    float* residual, double* error_device, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  if (!pa || pa->num_rows < 0 || pa->num_rows >= INT32_MAX || pa->num_pairs < 0) return DSPH_ERR_ARG;
  if (!error_device) return DSPH_ERR_ARG;
  if (!workspace || workspace_bytes < dsph_pp_dfsph_update_workspace(pa->num_rows)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  const int n = (int)pa->num_rows;
  const int nblk = (n + kBlock - 1) / kBlock;
  double* parts = static_cast<double*>(workspace);
  if (n > 0) {
    if (!pa->offsets || !masses_b || !densities_b || !accel_a || !accel_b || !source || !alpha || !pressure_in ||
        !pressure_out || !residual)
      return DSPH_ERR_ARG;
    if (pa->num_pairs > 0 && (!pa->j || !pa->gradients)) return DSPH_ERR_ARG;
    k_pp_update<<<nblk, kBlock, 0, st>>>(n, pa->num_ref, pa->num_pairs, pa->offsets, pa->j, pa->gradients, masses_b,
                                         densities_b, accel_a, accel_b, source, alpha, pressure_in, dt_pow, omega,
                                         clamp_pressure, rho0, threshold, pressure_out, residual, parts);
    DSPH_LAUNCH_CHECK();
  }
  k_pp_mean<<<1, kBlock, 0, st>>>(parts, nblk, n, error_device);
  return launch_status();
}
