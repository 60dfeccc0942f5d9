// Shared device helpers of the particle hot path (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <cmath>

#include "../../include/diffsph.h"

namespace dsph {

constexpr int kBlock = 256;  // 4 wave64s per workgroup

// Device form of a domain: L = fl32(max-min) and half = fl32(L/2) are what
// the reference's float32 tensor arithmetic produces
// (v2/modules/neighborhood.py:430, v2/math.py:4-6).
struct Dom {
  float min[2];
  float L[2];
  float half[2];
  int periodic[2];
};

inline Dom make_dom(const dsph_domain& d) {
  Dom o;
  for (int k = 0; k < 2; ++k) {
    o.min[k] = d.min[k];
    o.L[k] = d.max[k] - d.min[k];
    o.half[k] = (float)(((double)d.max[k] - (double)d.min[k]) / 2.0);
    o.periodic[k] = d.periodic[k] != 0;
  }
  return o;
}

// torch.remainder for float32: r = fmod(a,b); r += b when r and b differ in
// sign and r != 0.
__device__ __forceinline__ float py_remainder(float a, float b) {
  float r = fmodf(a, b);
  if (r != 0.0f && ((r < 0.0f) != (b < 0.0f))) r = r + b;
  return r;
}

// getPeriodicPosition (v2/modules/neighborhood.py:430).
__device__ __forceinline__ float wrap_coord(float x, const Dom& d, int k) {
  if (!d.periodic[k]) return x;
  return py_remainder(x - d.min[k], d.L[k]) + d.min[k];
}

// v2/math.py:4-6 `mod`, applied per axis to a pair offset.
__device__ __forceinline__ float mod_offset(float x, const Dom& d, int k) {
  if (!d.periodic[k]) return x;
  float ax = fabsf(x);
  if (ax > d.half[k]) {
    float s = x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
    return s * (py_remainder(ax + d.min[k], d.L[k]) + d.min[k]);
  }
  return x;
}

// Correctly rounded sqrt.  (HIP's __fsqrt_rn is the native approximate sqrt
// unless OCML_BASIC_ROUNDED_OPERATIONS is defined; sqrtf is IEEE under
// -fhip-fp32-correctly-rounded-divide-sqrt.)
__device__ __forceinline__ float sqrt_rn(float x) { return sqrtf(x); }

// |d|^2 with both products rounded before the add (no FMA contraction), so
// the neighbor predicate is reproducible bit for bit on the CPU.
__device__ __forceinline__ float dist2_exact(float dx, float dy) {
  return __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
}

// The neighbor predicate sqrt_rn(d2) /_rn h <= 1, decided from d2 alone when
// it lies clearly inside or outside h^2 (1 -/+ 1e-6): the rounding of the
// square root, of the division and of the thresholds themselves is below
// 3e-7 relative, so the shortcut returns exactly what the full expression
// returns; only the band between the thresholds evaluates it.
__device__ __forceinline__ bool within_support(float d2, float h) {
  const float hh = h * h;
  if (d2 < hh * 0.999999f) return true;
  if (d2 > hh * 1.000001f) return false;
  return __fdiv_rn(sqrt_rn(d2), h) <= 1.0f;
}

// Wendland C2 in 2-D (kernelFunctions/Wendland2.py): C_2 = 7/pi.
constexpr float kC2 = (float)(7.0 / 3.14159265358979323846);

__device__ __forceinline__ float clamp01(float t) { return fminf(fmaxf(t, 0.0f), 1.0f); }

// All float expressions below are evaluated exactly as written: the library
// is compiled with -ffp-contract=off and IEEE division/sqrt, so every value
// is reproduced bit for bit by a CPU implementation of the same expressions.

// Kernel constants of a support h: 1/h, C/h^2 (for W) and C/h^3 (for
// dW/dr).  Pair loops over a uniform support evaluate them once per row
// (KSel) instead of dividing by h, h^2 and h^3 for every pair.
struct KConst {
  float ih, cw, cd;
};
__device__ __forceinline__ KConst kconst(float h) {
  KConst k;
  k.ih = 1.0f / h;
  k.cw = kC2 / (h * h);
  k.cd = kC2 / ((h * h) * h);
  return k;
}

// W(q) = (1-q)^4 (1+4q) C/h^2, cpow(1-q,4) = ((t*t)*t)*t
__device__ __forceinline__ float wendland_w(float q, const KConst& k) {
  float t = clamp01(1.0f - q);
  float t4 = ((t * t) * t) * t;
  return (t4 * (1.0f + 4.0f * q)) * k.cw;
}
// dk/dq C/h^3 with dk/dq = -20 q (1-q)^3; grad W = xhat * this.
__device__ __forceinline__ float wendland_dw(float q, const KConst& k) {
  float t = clamp01(1.0f - q);
  float t3 = (t * t) * t;
  return ((-20.0f * q) * t3) * k.cd;
}

__device__ __forceinline__ float pick_h(int mode, float hi, float hj) {
  return mode == DSPH_MODE_SCATTER ? hj : (mode == DSPH_MODE_GATHER ? hi : 0.5f * (hi + hj));
}

// Geometry of one pair (query i, reference j) evaluated on the fly, with the
// reference's conventions: d = x_i - x_j (min-image), r = |d|,
// xhat = d/(r+1e-7), q = r/h_ij (v2/modules/neighborhood.py:445-451), in
// reciprocal form: q = r * (1/h), xhat = d * (1/(r+1e-7)), one division per
// pair (a few ulp from the quotients the reference forms).
struct Geom {
  float dx, dy;   // x_i - x_j
  float r, q, h;
  float ri;       // 1 / (r + 1e-7)
  float ex, ey;   // xhat
  float w;        // W_ij
  float dw;       // dW/dr
  float gx, gy;   // grad_i W_ij
};

__device__ __forceinline__ Geom pair_geom(float2 pi, float2 pj, float h, const KConst& k, const Dom& d) {
  Geom g;
  g.dx = mod_offset(pi.x - pj.x, d, 0);
  g.dy = mod_offset(pi.y - pj.y, d, 1);
  g.r = sqrt_rn(dist2_exact(g.dx, g.dy));
  g.h = h;
  g.q = g.r * k.ih;
  g.ri = 1.0f / (g.r + 1e-7f);
  g.ex = g.dx * g.ri;
  g.ey = g.dy * g.ri;
  g.w = wendland_w(g.q, k);
  g.dw = wendland_dw(g.q, k);
  g.gx = g.ex * g.dw;
  g.gy = g.ey * g.dw;
  return g;
}
__device__ __forceinline__ Geom pair_geom(float2 pi, float2 pj, float h, const Dom& d) {
  return pair_geom(pi, pj, h, kconst(h), d);
}

// The same geometry with the reference's own quotients (q = r/h, xhat =
// d/(r+1e-7), W and dW/dr divided by h^2 and h^3): the per-pair arrays
// handed to reference callers (neighborDict exports) use this form.
__device__ __forceinline__ Geom pair_geom_exact(float2 pi, float2 pj, float h, const Dom& d) {
  Geom g;
  g.dx = mod_offset(pi.x - pj.x, d, 0);
  g.dy = mod_offset(pi.y - pj.y, d, 1);
  g.r = sqrt_rn(dist2_exact(g.dx, g.dy));
  g.h = h;
  g.q = __fdiv_rn(g.r, h);
  const float den = g.r + 1e-7f;
  g.ri = 1.0f / den;
  g.ex = __fdiv_rn(g.dx, den);
  g.ey = __fdiv_rn(g.dy, den);
  const float t = clamp01(1.0f - g.q);
  const float t3 = (t * t) * t;
  g.w = ((t3 * t * (1.0f + 4.0f * g.q)) * kC2) / (h * h);
  g.dw = (((-20.0f * g.q) * t3) * kC2) / ((h * h) * h);
  g.gx = g.ex * g.dw;
  g.gy = g.ey * g.dw;
  return g;
}

// Kernel constants per pair: hoisted when both supports are the uniform h0
// (the branch is uniform across the wave), else evaluated for the pair's h.
struct KSel {
  KConst k0;
  bool uniform;
  __device__ __forceinline__ KConst operator()(float h) const { return uniform ? k0 : kconst(h); }
};

// The DFSPH pair coefficient V_j gradW_ij in the factored form
// d_ij * s_ij with s_ij = V_j * (dW/dr / (r + 1e-7)): the per-entry scalar
// s_ij is all the solver stores, d_ij is recomputed from the positions.
// (Algebraically the reference's V_j * (d/(r+1e-7)) * dW/dr; it differs
// from that product by a few ulp.)
__device__ __forceinline__ float dfsph_t(const Geom& g) { return g.dw * g.ri; }

struct PairsDev {
  int nq, nr;
  const int32_t* off;
  const int32_t* nbr;
  const float2* qpos;
  const float2* rpos;
  const float* qh;
  const float* rh;
  float h0;
  int mode;
  Dom dom;
  const int32_t* cnt;
  const int32_t* slice;
  const int16_t* nbr16;  // optional 16-bit deltas (pressure sweeps)
  const int32_t* ovf;    // nonzero: nbr16 invalid
  const uint16_t* tslot; // optional neighbor tiles (pressure sweeps)
  const int32_t* truns;
  const int32_t* tn;
  int tcap;
};

inline PairsDev make_pairs(const dsph_pairs& p) {
  PairsDev o;
  o.nq = (int)p.nq;
  o.nr = (int)p.nr;
  o.off = p.offsets;
  o.nbr = p.nbr;
  o.qpos = reinterpret_cast<const float2*>(p.qpos);
  o.rpos = reinterpret_cast<const float2*>(p.rpos);
  o.qh = p.qh;
  o.rh = p.rh;
  o.h0 = p.h_uniform;
  o.mode = p.mode;
  o.dom = make_dom(p.domain);
  o.cnt = p.counts;
  o.slice = p.slice_off;
  o.nbr16 = p.nbr16;
  o.ovf = p.nbr16_overflow;
  o.tslot = p.tile_slot;
  o.truns = p.tile_runs;
  o.tn = p.tile_n;
  o.tcap = p.tile_cap;
  return o;
}

// Neighbor tiles (the tile_* fields of dsph_pairs): tile b holds the
// particles of R disjoint runs, tile_runs[b * kTileRunStride ...] =
// {R, start_0, length_0, ...}, laid end to end; entry k of a row of
// workgroup b names its neighbor by its position in them, tile_slot[k].
constexpr int kTileMaxRuns = DSPH_TILE_MAX_RUNS;
constexpr int kTileRunStride = 2 * kTileMaxRuns + 1;

struct IdxSlot {
  const uint16_t* a;
  __device__ __forceinline__ int operator()(int k) const { return (int)__builtin_nontemporal_load(a + k); }
};

// Wave-wide inclusive scans (64 lanes, all active).
__device__ __forceinline__ int wave_incl_max(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(v, d, 64);
    if (lane >= d) v = v > y ? v : y;
  }
  return v;
}
__device__ __forceinline__ int wave_incl_sum(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(v, d, 64);
    if (lane >= d) v += y;
  }
  return v;
}

// The run table of tile b into LDS (arrays of kTileMaxRuns), loaded by wave
// 0: run u holds particles [run_s[u], run_e[u]), staged at positions
// j + run_d[u].  Returns R; the caller synchronises the workgroup before
// reading the table.  Only for a tile whose tile_n is valid.
__device__ __forceinline__ int load_run_table(const int32_t* __restrict__ tile_runs, int b, int* run_s, int* run_e,
                                              int* run_d) {
  const int32_t* tr = tile_runs + (size_t)b * kTileRunStride;
  const int R = tr[0];
  const int t = threadIdx.x;
  if (t < 64) {
    const int s0 = t < R ? tr[1 + 2 * t] : 0, len = t < R ? tr[2 + 2 * t] : 0;
    const int o = wave_incl_sum(len) - len;
    if (t < R) {
      run_s[t] = s0;
      run_e[t] = s0 + len;
      run_d[t] = o - s0;
    }
  }
  return R;
}

// Stages tile lb in LDS, run after run: put(o, j) for its particles j laid
// end to end (o = 0, 1, ...; contiguous loads within a run).  False
// (uniform over the workgroup, before any barrier) when the workgroup has
// no usable tile.
template <class Put>
__device__ __forceinline__ bool stage_tile(const PairsDev& P, int lb, const Put& put) {
  const int S = P.tn[lb];
  if (S < 0 || S > P.tcap) return false;
  const int32_t* tr = P.truns + (size_t)lb * kTileRunStride;
  const int R = tr[0];
  int o = 0;
  for (int u = 0; u < R; ++u) {
    const int s0 = tr[1 + 2 * u], len = tr[2 + 2 * u];
    for (int s = threadIdx.x; s < len; s += blockDim.x) put(o + s, s0 + s);
    o += len;
  }
  __syncthreads();
  return true;
}

// tile[o] = rec(j).
template <class T, class Rec>
__device__ __forceinline__ bool load_tile_by(const PairsDev& P, int lb, const Rec& rec, T* tile) {
  return stage_tile(P, lb, [&](int o, int j) { tile[o] = rec(j); });
}

// The kernel constants of a pair list (evaluated once per thread).
__device__ __forceinline__ KSel ksel(const PairsDev& P) {
  KSel s;
  s.uniform = !P.qh && !P.rh;
  s.k0 = kconst(P.h0);
  return s;
}

// Where the entries of row i live: entry n is at base + n*step (sliced ELL:
// step 64, consecutive lanes of a wave read consecutive words; CSR: step 1).
struct RowRef {
  int base;
  int step;
  int cnt;
};

__device__ __forceinline__ RowRef row_ref(const PairsDev& P, int i) {
  RowRef r;
  if (P.slice) {
    r.base = P.slice[i >> 6] + (i & 63);
    r.step = 64;
    r.cnt = P.cnt[i];
  } else {
    r.base = P.off[i];
    r.step = 1;
    r.cnt = P.off[i + 1] - P.off[i];
  }
  return r;
}

// slot = atomicAdd(&count[key], 1) for every valid lane of the wave, with one
// atomic per distinct key: lanes sharing a key take consecutive slots in lane
// order.  All 64 lanes must call it (invalid lanes pass valid = false).
__device__ __forceinline__ int wave_slot(int32_t* __restrict__ count, int key, bool valid) {
  const int lane = threadIdx.x & 63;
  unsigned long long remaining = __ballot(valid);
  int slot = 0;
  while (remaining) {
    const int leader = __ffsll((long long)remaining) - 1;
    const int lkey = __shfl(key, leader, 64);
    const unsigned long long same = __ballot(valid && key == lkey);
    int base = 0;
    if (lane == leader) base = atomicAdd(&count[lkey], __popcll(same));
    base = __shfl(base, leader, 64);
    if (valid && key == lkey) slot = base + __popcll(same & ((1ull << lane) - 1ull));
    remaining &= ~same;
  }
  return slot;
}

// body(j, load(j)) for j = b .. e-1 in order; the loads of U consecutive j
// are issued before the first body of the group, so U gathers are in
// flight at once.  The order of the bodies (and of any sums they form) is
// that of the plain loop.
template <int U, class Load, class Body>
__device__ __forceinline__ void scan_range(int b, int e, Load&& load, Body&& body) {
  int j = b;
  for (; j + U <= e; j += U) {
    decltype(load(j)) d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) d[u] = load(j + u);
#pragma unroll
    for (int u = 0; u < U; ++u) body(j + u, d[u]);
  }
  for (; j < e; ++j) body(j, load(j));
}

// body(j, load(j)) over the entries j of a neighbor row, in row order, with
// U entries and their gathers in flight (Stream: the entries are read with
// non-temporal loads, for lists read once).
template <int U, bool Stream, class Load, class Body>
__device__ __forceinline__ void scan_row(const int32_t* __restrict__ nbr, const RowRef& rr, Load&& load,
                                         Body&& body) {
  int n = 0, k = rr.base;
  for (; n + U <= rr.cnt; n += U, k += U * rr.step) {
    int j[U];
#pragma unroll
    for (int u = 0; u < U; ++u) j[u] = Stream ? __builtin_nontemporal_load(nbr + k + u * rr.step) : nbr[k + u * rr.step];
    decltype(load(0)) d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) d[u] = load(j[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) body(j[u], d[u]);
  }
  for (; n < rr.cnt; ++n, k += rr.step) {
    const int j = nbr[k];
    body(j, load(j));
  }
}

// CSR segment [b, e) of row i (offsets off, lengths nn) clipped to the
// pair arrays [0, np): inconsistent input cannot index out of bounds.
__device__ __forceinline__ void seg_bounds(const int32_t* off, const int32_t* nn, int i, int64_t np, int64_t& b,
                                           int64_t& e) {
  b = off[i];
  e = b + (int64_t)nn[i];
  if (b < 0) b = 0;
  if (e > np) e = np;
}

__device__ __forceinline__ float h_of(const float* arr, int idx, float h0) {
  return arr ? arr[idx] : h0;
}

// XCD-aware block order.  The dispatcher deals the workgroups of a launch
// round-robin over the 8 XCDs (block b runs on XCD b % 8), each with its own
// L2.  logical_block() renumbers them so that XCD x processes one contiguous
// range of logical blocks: particles in cell order are spatially local, so
// each XCD's L2 then holds the neighbor fields of its own 1/8 of the domain
// instead of all of it.  A bijection onto [0, gridDim.x); results do not
// depend on it.
#ifndef DSPH_XCD_REMAP
#define DSPH_XCD_REMAP 1
#endif
constexpr int kNumXcd = 8;

__device__ __forceinline__ int logical_block() {
#if DSPH_XCD_REMAP
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb / kNumXcd, r = nb % kNumXcd;
  const int x = b % kNumXcd, k = b / kNumXcd;
  return x * q + (x < r ? x : r) + k;
#else
  return blockIdx.x;
#endif
}

inline int grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : g);
}

inline int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DSPH_OK : DSPH_ERR_LAUNCH;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Validates a dsph_pairs descriptor (sizes, layout pointers, supports);
// defined in sph_ops.hip.
int check_pairs(const dsph_pairs* p);

}  // namespace dsph

#define DSPH_CHECK_HIP(expr)                          \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return DSPH_ERR_LAUNCH;     \
  } while (0)

#define DSPH_LAUNCH_CHECK()                           \
  do {                                                \
    int _s = ::dsph::launch_status();                 \
    if (_s != DSPH_OK) return _s;                     \
  } while (0)
