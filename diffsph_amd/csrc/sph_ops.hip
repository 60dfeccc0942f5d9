// SPH pair operators as gather kernels over a CSR neighbor list; kernel
// values are evaluated on the fly from positions (no per-pair arrays).
// One thread per query particle, neighbors in CSR order, no atomics.
#include "common.h"

namespace dsph {

int scan_exclusive(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws(int64_t n);

// ---------------------------------------------------------------------------
// Density: rho_i = sum_j m_j W_ij (v2/sphOps.py:19-29).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_density(PairsDev P, const float* __restrict__ rmass,
                                                    const float* __restrict__ qmass, float* __restrict__ rho,
                                                    float* __restrict__ vol, float* __restrict__ w_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  const KSel ks = ksel(P);
  float acc = 0.0f;
  const RowRef rr = row_ref(P, i);
  struct Nb {
    float2 x;
    float h, m;
  };
  auto load = [&](int j) { return Nb{P.rpos[j], h_of(P.rh, j, P.h0), rmass[j]}; };
  auto body = [&](int k, const Nb& d) {
    float dx = mod_offset(xi.x - d.x.x, P.dom, 0), dy = mod_offset(xi.y - d.x.y, P.dom, 1);
    const KConst kk = ks(pick_h(P.mode, hi, d.h));
    float q = sqrt_rn(dist2_exact(dx, dy)) * kk.ih;
    const float w = wendland_w(q, kk);
    if (w_out) w_out[k] = w;
    acc += d.m * w;
  };
  // four entries' gathers in flight; the sum runs in entry order
  int n = 0, k = rr.base;
  for (; n + 4 <= rr.cnt; n += 4, k += 4 * rr.step) {
    int j[4];
    Nb d[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = P.nbr[k + u * rr.step];
#pragma unroll
    for (int u = 0; u < 4; ++u) d[u] = load(j[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) body(k + u * rr.step, d[u]);
  }
  for (; n < rr.cnt; ++n, k += rr.step) body(k, load(P.nbr[k]));
  rho[i] = acc;
  if (vol) vol[i] = qmass[i] / acc;
}

// Density over neighbor tiles (uniform support): the workgroup's neighbors'
// (x, y, m) staged in LDS; arithmetic and order as k_density.
__global__ void __launch_bounds__(kBlock) k_density_tiled(PairsDev P, const float* __restrict__ rmass,
                                                          const float* __restrict__ qmass, float* __restrict__ rho,
                                                          float* __restrict__ vol, float* __restrict__ w_out) {
  extern __shared__ float4 tile[];
  const int b = blockIdx.x;
  const int i = b * kBlock + threadIdx.x;
  const bool tiled = load_tile_by(
      P, b,
      [&](int j) {
        const float2 x = P.rpos[j];
        return make_float4(x.x, x.y, rmass[j], 0.f);
      },
      tile);
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const KConst kk = ksel(P)(P.h0);
  float acc = 0.0f;
  const RowRef rr = row_ref(P, i);
  auto body = [&](int k, float4 d) {
    const float dx = mod_offset(xi.x - d.x, P.dom, 0), dy = mod_offset(xi.y - d.y, P.dom, 1);
    const float q = sqrt_rn(dist2_exact(dx, dy)) * kk.ih;
    const float w = wendland_w(q, kk);
    if (w_out) w_out[k] = w;
    acc += d.z * w;
  };
  auto rec = [&](int j) {
    const float2 x = P.rpos[j];
    return make_float4(x.x, x.y, rmass[j], 0.f);
  };
  int n = 0, k = rr.base;
  if (tiled) {
    const IdxSlot idx{P.tslot};
    for (; n + 4 <= rr.cnt; n += 4, k += 4 * rr.step) {
      int sl[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) sl[u] = idx(k + u * rr.step);
#pragma unroll
      for (int u = 0; u < 4; ++u) body(k + u * rr.step, tile[sl[u]]);
    }
    for (; n < rr.cnt; ++n, k += rr.step) body(k, tile[idx(k)]);
  } else {
    for (; n + 4 <= rr.cnt; n += 4, k += 4 * rr.step) {
      int j[4];
      float4 d[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) j[u] = P.nbr[k + u * rr.step];
#pragma unroll
      for (int u = 0; u < 4; ++u) d[u] = rec(j[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u) body(k + u * rr.step, d[u]);
    }
    for (; n < rr.cnt; ++n, k += rr.step) body(k, rec(P.nbr[k]));
  }
  rho[i] = acc;
  if (vol) vol[i] = qmass[i] / acc;
}

// The density launch: over neighbor tiles when the list carries them and
// the support is uniform.
static void launch_density(const PairsDev& P, const float* rmass, const float* qmass, float* rho, float* vol,
                           float* w_out, hipStream_t st) {
  if (P.tslot && !P.qh && !P.rh)
    k_density_tiled<<<grid_for(P.nq), kBlock, (size_t)P.tcap * sizeof(float4), st>>>(P, rmass, qmass, rho, vol,
                                                                                      w_out);
  else
    k_density<<<grid_for(P.nq), kBlock, 0, st>>>(P, rmass, qmass, rho, vol, w_out);
}

// ---------------------------------------------------------------------------
// Generic operation (v2/sphOps.py:6-149).
// ---------------------------------------------------------------------------
template <int W>
struct VecT {
  float v[W];
};

template <int W>
__global__ void __launch_bounds__(kBlock) k_sph_operation(PairsDev P, int op, int gmode,
                                                          const float* __restrict__ qm, const float* __restrict__ rm,
                                                          const float* __restrict__ qr, const float* __restrict__ rr,
                                                          const float* __restrict__ qq, const float* __restrict__ rq,
                                                          float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float hi = h_of(P.qh, i, P.h0);
  float qi[W];
#pragma unroll
  for (int c = 0; c < W; ++c) qi[c] = qq ? qq[(int64_t)i * W + c] : 0.0f;
  const float rho_i = qr ? qr[i] : 1.0f;
  float acc[W * 2];
#pragma unroll
  for (int c = 0; c < W * 2; ++c) acc[c] = 0.0f;
  const RowRef row = row_ref(P, i);
  const KSel ks = ksel(P);
  for (int n = 0, k = row.base; n < row.cnt; ++n, k += row.step) {
    const int j = P.nbr[k];
    const float hj = h_of(P.rh, j, P.h0);
    const float h = pick_h(P.mode, hi, hj);
    Geom g = pair_geom(xi, P.rpos[j], h, ks(h), P.dom);
    float qj[W];
#pragma unroll
    for (int c = 0; c < W; ++c) qj[c] = rq[(int64_t)j * W + c];
    if (op == DSPH_OP_INTERPOLATE) {
      const float kk = rm[j] / rr[j] * g.w;
#pragma unroll
      for (int c = 0; c < W; ++c) acc[c] += qj[c] * kk;
    } else if (gmode == DSPH_GRAD_SYMMETRIC) {
      const float kx = rm[j] * g.gx, ky = rm[j] * g.gy;
      const float irj = 1.0f / (rr[j] * rr[j]), iri = 1.0f / (rho_i * rho_i);
      if (op == DSPH_OP_GRADIENT) {
#pragma unroll
        for (int c = 0; c < W; ++c) {
          float a = qi[c] * iri + qj[c] * irj;
          acc[2 * c] += a * kx;
          acc[2 * c + 1] += a * ky;
        }
      } else {
        acc[0] += (qi[0] * iri + qj[0] * irj) * kx + (qi[W > 1 ? 1 : 0] * iri + qj[W > 1 ? 1 : 0] * irj) * ky;
      }
    } else {
      // divergenceConsistent (v2/sphOps.py:108-149) weighs with m_j / rho_i
      const float V = op == DSPH_OP_DIVERGENCE_CONSISTENT ? rm[j] / rho_i : rm[j] / rr[j];
      const float kx = V * g.gx, ky = V * g.gy;
      if (op == DSPH_OP_GRADIENT) {
#pragma unroll
        for (int c = 0; c < W; ++c) {
          float a = gmode == DSPH_GRAD_DIFFERENCE ? qi[c] - qj[c]
                    : gmode == DSPH_GRAD_SUMMATION ? qi[c] + qj[c]
                                                   : qj[c];
          acc[2 * c] += a * kx;
          acc[2 * c + 1] += a * ky;
        }
      } else {
        // divergence: 'difference' is q_j - q_i (v2/sphOps.py:95-96)
        float ax, ay;
        const int c1 = W > 1 ? 1 : 0;
        if (gmode == DSPH_GRAD_DIFFERENCE) {
          ax = qj[0] - qi[0];
          ay = qj[c1] - qi[c1];
        } else if (gmode == DSPH_GRAD_SUMMATION) {
          ax = qj[0] + qi[0];
          ay = qj[c1] + qi[c1];
        } else {
          ax = qj[0];
          ay = qj[c1];
        }
        acc[0] += ax * kx + ay * ky;
      }
    }
  }
  if (op == DSPH_OP_INTERPOLATE) {
#pragma unroll
    for (int c = 0; c < W; ++c) out[(int64_t)i * W + c] = acc[c];
  } else if (op == DSPH_OP_GRADIENT) {
    const float s = gmode == DSPH_GRAD_SYMMETRIC ? rho_i : 1.0f;
#pragma unroll
    for (int c = 0; c < 2 * W; ++c) out[(int64_t)i * 2 * W + c] = gmode == DSPH_GRAD_SYMMETRIC ? s * acc[c] : acc[c];
  } else {
    out[i] = gmode == DSPH_GRAD_SYMMETRIC ? rho_i * acc[0] : acc[0];
  }
}

// ---------------------------------------------------------------------------
// Export of the reference's per-pair neighborDict arrays.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_export_counts(PairsDev P, const int32_t* __restrict__ q_rank,
                                                          int32_t* __restrict__ cnt) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P.nq) return;
  int row = q_rank ? q_rank[q] : q;
  cnt[q] = row_ref(P, row).cnt;
}

// Sliced list -> CSR order.
__global__ void __launch_bounds__(kBlock) k_sell_to_csr(PairsDev P, int32_t* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const RowRef rr = row_ref(P, i);
  const int o = P.off[i];
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step) out[o + n] = P.nbr[k];
}

__global__ void __launch_bounds__(kBlock) k_export_pairs(PairsDev P, const int32_t* __restrict__ q_sort,
                                                         const int32_t* __restrict__ q_rank,
                                                         const int32_t* __restrict__ r_sort,
                                                         const int32_t* __restrict__ out_off, int64_t* out_i,
                                                         int64_t* out_j, float* out_r, float* out_x, float* out_h,
                                                         float* out_w, float* out_g) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P.nq) return;
  const int row = q_rank ? q_rank[q] : q;
  const float2 xi = P.qpos[row];
  const float hi = h_of(P.qh, row, P.h0);
  const RowRef rr = row_ref(P, row);
  int64_t o = out_off[q];
  for (int n = 0, k = rr.base; n < rr.cnt; ++n, k += rr.step, ++o) {
    const int j = P.nbr[k];
    const float hj = h_of(P.rh, j, P.h0);
    const float h = pick_h(P.mode, hi, hj);
    // exported per-pair arrays carry the reference's own quotients
    Geom g = pair_geom_exact(xi, P.rpos[j], h, P.dom);
    if (out_i) out_i[o] = q;
    if (out_j) out_j[o] = r_sort ? r_sort[j] : j;
    if (out_r) out_r[o] = g.q;
    if (out_x) {
      out_x[2 * o] = g.ex;
      out_x[2 * o + 1] = g.ey;
    }
    if (out_h) out_h[o] = h;
    if (out_w) out_w[o] = g.w;
    if (out_g) {
      out_g[2 * o] = g.gx;
      out_g[2 * o + 1] = g.gy;
    }
  }
  (void)q_sort;
}

// ---------------------------------------------------------------------------
// Gather / scatter of per-particle rows.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_gather(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                                   int64_t n, int w, float* __restrict__ dst) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * w) return;
  int64_t p = t / w;
  int c = (int)(t - p * w);
  dst[t] = src[(int64_t)idx[p] * w + c];
}
// Several per-particle arrays permuted by one index in one pass:
// dst_f[p] = src_f[idx[p]], words[f] 4-byte words per element (1, 2 or 4).
constexpr int kGatherFields = 8;
struct GatherSet {
  const uint32_t* src[kGatherFields];
  uint32_t* dst[kGatherFields];
  int words[kGatherFields];
  int nf;
};

__global__ void __launch_bounds__(kBlock) k_gather_set(GatherSet g, const int32_t* __restrict__ idx, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t j = idx[p];
  for (int f = 0; f < g.nf; ++f) {
    const int w = g.words[f];
    if (w == 1) {
      g.dst[f][p] = g.src[f][j];
    } else if (w == 2) {
      reinterpret_cast<uint2*>(g.dst[f])[p] = reinterpret_cast<const uint2*>(g.src[f])[j];
    } else {
      reinterpret_cast<uint4*>(g.dst[f])[p] = reinterpret_cast<const uint4*>(g.src[f])[j];
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_scatter(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                                    int64_t n, int w, float* __restrict__ dst) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * w) return;
  int64_t p = t / w;
  int c = (int)(t - p * w);
  dst[(int64_t)idx[p] * w + c] = src[t];
}

// ---------------------------------------------------------------------------
// Drop-in ops over the reference's per-pair arrays (CSR given by
// numNeighbors / neighborOffset).
// ---------------------------------------------------------------------------
// The caller's CSR may be inconsistent: each segment is clipped to
// [0, num_pairs) and neighbor indices outside [0, num_ref) are skipped, so a
// bad argument can never make the kernel read out of bounds.
__global__ void __launch_bounds__(kBlock) k_op_density(const float* __restrict__ mj, const int64_t* __restrict__ jj,
                                                       const float* __restrict__ W, int n, int64_t nref, int64_t np,
                                                       const int32_t* __restrict__ nn, const int32_t* __restrict__ off,
                                                       float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc = 0.0f;
  int64_t b, e;
  seg_bounds(off, nn, i, np, b, e);
  for (int64_t k = b; k < e; ++k) {
    const int64_t j = jj[k];
    if (j >= 0 && j < nref) acc += mj[j] * W[k];
  }
  out[i] = acc;
}

__global__ void __launch_bounds__(kBlock) k_op_interp(const float* __restrict__ mj, const float* __restrict__ rj,
                                                      const float* __restrict__ qj, int w,
                                                      const int64_t* __restrict__ jj, const float* __restrict__ W,
                                                      int n, int64_t nref, int64_t np, const int32_t* __restrict__ nn,
                                                      const int32_t* __restrict__ off, float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  int64_t b, e;
  seg_bounds(off, nn, i, np, b, e);
  for (int64_t k = b; k < e; ++k) {
    const int64_t j = jj[k];
    if (j < 0 || j >= nref) continue;
    const float f = mj[j] / rj[j] * W[k];
    for (int c = 0; c < w; ++c) acc[c] += qj[j * w + c] * f;
  }
  for (int c = 0; c < w; ++c) out[(int64_t)i * w + c] = acc[c];
}

__global__ void __launch_bounds__(kBlock) k_op_scatter_add(const float* __restrict__ in, int w, int n, int64_t np,
                                                           const int32_t* __restrict__ nn,
                                                           const int32_t* __restrict__ off, float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  int64_t b, e;
  seg_bounds(off, nn, i, np, b, e);
  for (int64_t k = b; k < e; ++k)
    for (int c = 0; c < w; ++c) acc[c] += in[k * w + c];
  for (int c = 0; c < w; ++c) out[(int64_t)i * w + c] = acc[c];
}

int check_pairs(const dsph_pairs* p) {
  if (!p || p->nq < 0 || p->nr < 0 || p->nq >= INT32_MAX || p->nr >= INT32_MAX) return DSPH_ERR_ARG;
  if (p->mode < 0 || p->mode > 2) return DSPH_ERR_ARG;
  if (p->nbr16 && !p->nbr16_overflow) return DSPH_ERR_ARG;
  if (p->tile_slot && (!p->tile_runs || !p->tile_n || p->tile_cap < 1 || p->tile_cap > DSPH_TILE_MAX_CAP ||
                       p->nq != p->nr))
    return DSPH_ERR_ARG;
  if (p->nq > 0 && (!p->qpos || (!p->slice_off && !p->offsets))) return DSPH_ERR_ARG;
  if (p->nq > 0 && p->slice_off && !p->counts) return DSPH_ERR_ARG;
  if (p->nr > 0 && !p->rpos) return DSPH_ERR_ARG;
  if (((p->mode != DSPH_MODE_GATHER && !p->rh) || (p->mode != DSPH_MODE_SCATTER && !p->qh)) && !(p->h_uniform > 0.0f))
    return DSPH_ERR_ARG;
  return DSPH_OK;
}

}  // namespace dsph

using namespace dsph;

extern "C" const char* dsph_status_string(int s) {
  switch (s) {
    case DSPH_OK: return "ok";
    case DSPH_ERR_ARG: return "invalid argument";
    case DSPH_ERR_DIM: return "unsupported dimension or grid shape";
    case DSPH_ERR_LAUNCH: return "HIP launch or runtime failure";
    case DSPH_ERR_NOT_CONVERGED: return "solver did not converge";
    case DSPH_ERR_CAPACITY: return "output buffer or workspace too small";
    case DSPH_ERR_DEVICE: return "no usable HIP device";
    default: return "unknown status";
  }
}

extern "C" int dsph_abi_version(void) { return 4; }

extern "C" int dsph_density(const dsph_pairs* pairs, const float* rmass, const float* qmass, float* rho, float* vol,
                            void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq == 0) return DSPH_OK;
  if (!rmass || !rho || (vol && !qmass)) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  launch_density(P, rmass, qmass, rho, vol, nullptr, as_stream(stream));
  return launch_status();
}

extern "C" int dsph_density_w(const dsph_pairs* pairs, const float* rmass, const float* qmass, float* rho,
                              float* vol, float* pair_w, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (pairs->nq == 0) return DSPH_OK;
  if (!rmass || !rho || (vol && !qmass) || !pair_w) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  launch_density(P, rmass, qmass, rho, vol, pair_w, as_stream(stream));
  return launch_status();
}

extern "C" int dsph_sph_operation(const dsph_pairs* pairs, int32_t op, int32_t grad_mode, const float* qmass,
                                  const float* rmass, const float* qrho, const float* rrho, const float* qq,
                                  const float* rq, int32_t width, float* out, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (op < 0 || op > 3 || grad_mode < 0 || grad_mode > 3) return DSPH_ERR_ARG;
  if (width < 1 || width > 2) return DSPH_ERR_DIM;
  if ((op == DSPH_OP_DIVERGENCE || op == DSPH_OP_DIVERGENCE_CONSISTENT) && width != 2) return DSPH_ERR_DIM;
  if (op == DSPH_OP_DIVERGENCE_CONSISTENT && !qrho) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!rmass || !rrho || !rq || !out) return DSPH_ERR_ARG;
  const bool needs_qi = op != DSPH_OP_INTERPOLATE && grad_mode != DSPH_GRAD_NAIVE;
  if (needs_qi && !qq) return DSPH_ERR_ARG;
  if (op != DSPH_OP_INTERPOLATE && grad_mode == DSPH_GRAD_SYMMETRIC && !qrho) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  hipStream_t st = as_stream(stream);
  if (width == 1)
    k_sph_operation<1><<<grid_for(P.nq), kBlock, 0, st>>>(P, op, grad_mode, qmass, rmass, qrho, rrho, qq, rq, out);
  else
    k_sph_operation<2><<<grid_for(P.nq), kBlock, 0, st>>>(P, op, grad_mode, qmass, rmass, qrho, rrho, qq, rq, out);
  return launch_status();
}

extern "C" size_t dsph_pairs_export_workspace(int64_t nq) {
  return ((size_t)(nq > 0 ? nq : 1) * 4 + 255) / 256 * 256 + scan_ws(nq);
}

extern "C" int dsph_pairs_export(const dsph_pairs* pairs, const int32_t* q_sort_index, const int32_t* q_rank,
                                 const int32_t* r_sort_index, int32_t* out_offsets, int64_t* out_i, int64_t* out_j,
                                 float* out_distances, float* out_vectors, float* out_supports, float* out_kernels,
                                 float* out_gradients, void* ws, size_t ws_bytes, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (!out_offsets) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_pairs_export_workspace(pairs->nq)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  PairsDev P = make_pairs(*pairs);
  int32_t* cnt = static_cast<int32_t*>(ws);
  char* scan_w = static_cast<char*>(ws) + ((size_t)(pairs->nq > 0 ? pairs->nq : 1) * 4 + 255) / 256 * 256;
  if (P.nq > 0) {
    k_export_counts<<<grid_for(P.nq), kBlock, 0, st>>>(P, q_rank, cnt);
    DSPH_LAUNCH_CHECK();
  }
  s = scan_exclusive(cnt, P.nq, out_offsets, scan_w, scan_ws(P.nq), st);
  if (s != DSPH_OK) return s;
  if (P.nq == 0) return DSPH_OK;
  k_export_pairs<<<grid_for(P.nq), kBlock, 0, st>>>(P, q_sort_index, q_rank, r_sort_index, out_offsets, out_i, out_j,
                                                   out_distances, out_vectors, out_supports, out_kernels,
                                                   out_gradients);
  return launch_status();
}

extern "C" int dsph_sell_to_csr(const dsph_pairs* pairs, int32_t* out_nbr, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (!pairs->offsets) return DSPH_ERR_ARG;
  if (pairs->nq == 0) return DSPH_OK;
  if (!out_nbr || !pairs->nbr) return DSPH_ERR_ARG;
  PairsDev P = make_pairs(*pairs);
  k_sell_to_csr<<<grid_for(P.nq), kBlock, 0, as_stream(stream)>>>(P, out_nbr);
  return launch_status();
}

extern "C" int dsph_gather_f32(const float* src, const int32_t* index, int64_t n, int32_t width, float* dst,
                               void* stream) {
  if (n < 0 || width < 1) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  if (!src || !index || !dst) return DSPH_ERR_ARG;
  k_gather<<<grid_for(n * width), kBlock, 0, as_stream(stream)>>>(src, index, n, width, dst);
  return launch_status();
}

extern "C" int dsph_gather_set(const int32_t* index, int64_t n, int32_t nfields, const void* const* src,
                               void* const* dst, const int32_t* words, void* stream) {
  if (n < 0 || n >= INT32_MAX || nfields < 0 || nfields > kGatherFields) return DSPH_ERR_ARG;
  if (n == 0 || nfields == 0) return DSPH_OK;
  if (!index || !src || !dst || !words) return DSPH_ERR_ARG;
  GatherSet g{};
  g.nf = nfields;
  for (int f = 0; f < nfields; ++f) {
    const int w = words[f];
    const size_t align = (size_t)w * 4;
    if ((w != 1 && w != 2 && w != 4) || !src[f] || !dst[f]) return DSPH_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(src[f]) % align || reinterpret_cast<uintptr_t>(dst[f]) % align)
      return DSPH_ERR_ARG;
    g.src[f] = static_cast<const uint32_t*>(src[f]);
    g.dst[f] = static_cast<uint32_t*>(dst[f]);
    g.words[f] = w;
  }
  k_gather_set<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(g, index, (int)n);
  return launch_status();
}

extern "C" int dsph_scatter_f32(const float* src, const int32_t* index, int64_t n, int32_t width, float* dst,
                                void* stream) {
  if (n < 0 || width < 1) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  if (!src || !index || !dst) return DSPH_ERR_ARG;
  k_scatter<<<grid_for(n * width), kBlock, 0, as_stream(stream)>>>(src, index, n, width, dst);
  return launch_status();
}

extern "C" int dsph_op_sph_density(const float* masses_j, int64_t num_ref, const int64_t* indices_j,
                                   const float* kernels, int64_t num_pairs, int32_t num_particles,
                                   const int32_t* num_neighbors, const int32_t* neighbor_offset, float* out,
                                   void* stream) {
  if (num_particles < 0 || num_ref < 0 || num_pairs < 0) return DSPH_ERR_ARG;
  if (num_particles == 0) return DSPH_OK;
  if (!num_neighbors || !neighbor_offset || !out) return DSPH_ERR_ARG;
  if ((num_pairs > 0 && (!indices_j || !kernels)) || (num_ref > 0 && !masses_j)) return DSPH_ERR_ARG;
  k_op_density<<<grid_for(num_particles), kBlock, 0, as_stream(stream)>>>(
      masses_j, indices_j, kernels, num_particles, num_ref, num_pairs, num_neighbors, neighbor_offset, out);
  return launch_status();
}

extern "C" int dsph_op_sph_interpolation(const float* masses_j, const float* densities_j, const float* quantities_j,
                                         int32_t width, int64_t num_ref, const int64_t* indices_j,
                                         const float* kernels, int64_t num_pairs, int32_t num_particles,
                                         const int32_t* num_neighbors, const int32_t* neighbor_offset, float* out,
                                         void* stream) {
  if (num_particles < 0 || width < 1 || width > 4 || num_ref < 0 || num_pairs < 0) return DSPH_ERR_ARG;
  if (num_particles == 0) return DSPH_OK;
  if (!num_neighbors || !neighbor_offset || !out) return DSPH_ERR_ARG;
  if ((num_pairs > 0 && (!indices_j || !kernels)) || (num_ref > 0 && (!masses_j || !densities_j || !quantities_j)))
    return DSPH_ERR_ARG;
  k_op_interp<<<grid_for(num_particles), kBlock, 0, as_stream(stream)>>>(masses_j, densities_j, quantities_j, width,
                                                                        indices_j, kernels, num_particles, num_ref,
                                                                        num_pairs, num_neighbors, neighbor_offset,
                                                                        out);
  return launch_status();
}

extern "C" int dsph_op_scatter_add(const float* input, int32_t width, int64_t num_pairs, int32_t num_particles,
                                   const int32_t* num_neighbors, const int32_t* neighbor_offset, float* out,
                                   void* stream) {
  if (num_particles < 0 || width < 1 || width > 4 || num_pairs < 0) return DSPH_ERR_ARG;
  if (num_particles == 0) return DSPH_OK;
  if ((num_pairs > 0 && !input) || !num_neighbors || !neighbor_offset || !out) return DSPH_ERR_ARG;
  k_op_scatter_add<<<grid_for(num_particles), kBlock, 0, as_stream(stream)>>>(input, width, num_particles, num_pairs,
                                                                             num_neighbors, neighbor_offset, out);
  return launch_status();
}
