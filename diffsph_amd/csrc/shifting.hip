// Implicit particle shifting (IPS): the kernel-Hessian system of
// computeShifting (v2/modules/shifting.py:192-284, 361-460) and the vector
// pieces of its BiCGStab (bicgstab_shifting, :318-...).
//
//   k_ips_assemble  per row i: B_i = sum_j A_j gradW_ij (the right-hand side,
//                   interleaved x/y) and, per list entry, the 2x2 block
//                   A_j H_ij of the Hessian2D of src/diffSPH/kernels.py:263-303
//                   (symmetric: h00, h01, h11 stored as a float4)
//   k_ips_spmv      y_i = sum_k H_k x_{j_k} over the same list
//   k_lincomb       out = x + a y + b z  (float32, the reference's order)
//   k_dot_*         float64 dot products, fixed-order two-level reduction
//   k_max_norm      max_i |(x_2i, x_2i+1)| (the solver threshold test)
//
// Pair geometry uses the reference's own quotients (pair_geom_exact), as
// the neighborDict the reference evaluates the Hessian on holds them.
#include "common.h"

namespace dsph {

__device__ __forceinline__ float cpow01(float t, int p) {
  t = clamp01(t);
  float r = t;
  for (int k = 1; k < p; ++k) r = r * t;
  return r;
}

__global__ void __launch_bounds__(kBlock) k_ips_assemble(PairsDev P, const float* __restrict__ areas,
                                                         float2* __restrict__ B, float4* __restrict__ Hs,
                                                         int64_t cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const RowRef rr = row_ref(P, i);
  float jx = 0.f, jy = 0.f;
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    const int j = P.nbr[k];
    const float h = pick_h(P.mode, hi, h_of(P.rh, j, P.h0));
    const Geom g = pair_geom_exact(xi, P.rpos[j], h, P.dom);
    const float A = areas[j];
    jx += g.gx * A;
    jy += g.gy * A;
    // Hessian2D with x_ij = xhat r, r = q h (kernels.py:264-303)
    const float q = g.q;
    const float r = q * h;
    const float x0 = g.ex * r, x1 = g.ey * r;
    const float factor = kC2 / (h * h);
    const float k2 = (20.0f * (4.0f * q - 1.0f)) * cpow01(1.0f - q, 2);
    const float k1 = (-20.0f * q) * cpow01(1.0f - q, 3);
    const float dA = (h * h) * (r * r) + 1e-5f * h;
    const float dB = h * ((r * r) * r) + 1e-10f * h;
    const float tAx = (x0 * x0) / dA * k2, tAy = (x1 * x1) / dA * k2;
    const float tBx = (x1 * x1) / dB * k1, tBy = (x0 * x0) / dB * k1;
    const float self = q < 1e-5f ? (1.0f / (h * h)) * -20.0f : 0.0f;  // d2kdq2(0) = -20
    const float dxx = factor * ((tAx + tBx) + self);
    const float dyy = factor * ((tAy + tBy) + self);
    const float dxy = (kC2 / (h * h)) * (((x0 * x1) / (h * (r * r) + 1e-5f * h)) *
                                         ((1.0f / h) * k2 - (1.0f / (r + 1e-3f * h)) * k1));
    const bool in = q <= 1.0f;
    if (k < cap) Hs[k] = make_float4(in ? A * dxx : 0.f, in ? A * dxy : 0.f, in ? A * dyy : 0.f, 0.f);
  }
  B[i] = make_float2(jx, jy);
}

__global__ void __launch_bounds__(kBlock) k_ips_spmv(PairsDev P, const float4* __restrict__ Hs, int64_t cap,
                                                     const float2* __restrict__ x, float2* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const RowRef rr = row_ref(P, i);
  float a = 0.f, b = 0.f;
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) {
    if (k >= cap) break;
    const float4 H = Hs[k];
    const float2 v = x[P.nbr[k]];
    a += H.x * v.x + H.y * v.y;
    b += H.y * v.x + H.z * v.y;
  }
  y[i] = make_float2(a, b);
}

__global__ void __launch_bounds__(kBlock) k_lincomb(int64_t n, const float* __restrict__ x, float a,
                                                    const float* __restrict__ y, float b,
                                                    const float* __restrict__ z, float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float v = x[k];
  if (y) v = v + a * y[k];
  if (z) v = v + b * z[k];
  out[k] = v;
}

__global__ void __launch_bounds__(kBlock) k_dot_partial(int64_t n, const float* __restrict__ a,
                                                        const float* __restrict__ b, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  sh[threadIdx.x] = k < n ? (double)a[k] * (double)b[k] : 0.0;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_dot_final(const double* __restrict__ part, int nb,
                                                      double* __restrict__ out) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) s += part[b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int t = kBlock / 2; t > 0; t >>= 1) {
    if ((int)threadIdx.x < t) sh[threadIdx.x] += sh[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_max_norm(int n, const float2* __restrict__ x, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const double a = x[i].x, b = x[i].y;
    v = sqrt(a * a + b * b);
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_max_final(const double* __restrict__ part, int nb,
                                                      double* __restrict__ out) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) s = fmax(s, part[b]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int t = kBlock / 2; t > 0; t >>= 1) {
    if ((int)threadIdx.x < t) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + t]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

}  // namespace dsph

using namespace dsph;

extern "C" int dsph_ips_assemble(const dsph_pairs* pairs, const float* areas, float* rhs, float* hessian,
                                 int64_t hessian_entries, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq == 0) return DSPH_OK;
  if (!areas || !rhs || !hessian || hessian_entries < 1) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  k_ips_assemble<<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(
      P, areas, reinterpret_cast<float2*>(rhs), reinterpret_cast<float4*>(hessian), hessian_entries);
  return launch_status();
}

extern "C" int dsph_ips_spmv(const dsph_pairs* pairs, const float* hessian, int64_t hessian_entries, const float* x,
                             float* y, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq == 0) return DSPH_OK;
  if (!hessian || !x || !y || hessian_entries < 1) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  k_ips_spmv<<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(P, reinterpret_cast<const float4*>(hessian),
                                                               hessian_entries, reinterpret_cast<const float2*>(x),
                                                               reinterpret_cast<float2*>(y));
  return launch_status();
}

extern "C" int dsph_vec_lincomb(int64_t n, const float* x, float a, const float* y, float b, const float* z,
                                float* out, void* stream) {
  if (n < 0 || (n > 0 && (!x || !out))) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  k_lincomb<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, as_stream(stream)>>>(n, x, a, y, b, z, out);
  return launch_status();
}

extern "C" size_t dsph_vec_reduce_workspace(int64_t n) {
  const int64_t nb = (n + kBlock - 1) / kBlock;
  return (size_t)(((nb > 0 ? nb : 1) * sizeof(double) + 255) / 256 * 256);
}

extern "C" int dsph_vec_dot(int64_t n, const float* a, const float* b, double* out_device, void* ws,
                            size_t ws_bytes, void* stream) {
  if (n < 0 || !out_device || (n > 0 && (!a || !b))) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_vec_reduce_workspace(n)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  const int64_t nb = (n + kBlock - 1) / kBlock;
  double* part = static_cast<double*>(ws);
  if (nb > 0) k_dot_partial<<<(unsigned)nb, kBlock, 0, st>>>(n, a, b, part);
  DSPH_LAUNCH_CHECK();
  k_dot_final<<<1, kBlock, 0, st>>>(part, (int)nb, out_device);
  return launch_status();
}

extern "C" int dsph_vec_max_norm2(int64_t n, const float* x, double* out_device, void* ws, size_t ws_bytes,
                                  void* stream) {
  if (n < 0 || n >= INT32_MAX || !out_device || (n > 0 && !x)) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_vec_reduce_workspace(n)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  const int64_t nb = (n + kBlock - 1) / kBlock;
  double* part = static_cast<double*>(ws);
  if (nb > 0) k_max_norm<<<(unsigned)nb, kBlock, 0, st>>>((int)n, reinterpret_cast<const float2*>(x), part);
  DSPH_LAUNCH_CHECK();
  k_max_final<<<1, kBlock, 0, st>>>(part, (int)nb, out_device);
  return launch_status();
}
