// Batched ensemble of independent periodic delta-SPH domains (gfx950).
//
// All domains live in one concatenated particle state; every pass of a step
// is one launch over all particles, each particle resolving its own box,
// grid and time step through dom_of[i].  Work is per particle, so 1k- and
// 64k-particle domains share the machine without per-domain launches or
// load-balancing logic.
//
// One step (per domain exactly the single-domain sequence of
// v2/simulationSchemes/deltaPlus.py:176-422 + v2/modules/integration.py:
// 231-307 + v2/modules/timestep.py:8-39 as the single-domain entry points
// run it):
//   sort:   cell keys (global cell = domain cell offset + local cell) ->
//           counting sort with per-cell insertion sort -> permute the state
//           A -> B (B is in cell order; domains stay contiguous)
//   stage s = 0..3: grid of the stage positions (stage 0 reuses the sort's)
//           -> fixed-width sliced-ELL rows (entries mapped to state order)
//           -> normalization + density gradient -> right-hand side
//           -> stage state s+1 = B + k_s c  (c = dt/2, dt/2, dt per domain)
//   RK4:    A = B + dt (k1 + 2k2 + 2k3 + k4)/6 per domain
//   dt:     per-domain float64 sum of squares of the averaged dudt in the
//           same reduction order as the single-domain reduction, then
//           computeTimestep in float32.
// No host synchronisation, no allocation, no memset nodes: a sequence of
// steps can be captured into one HIP graph.
#include <cstring>
#include <vector>

#include "common.h"
#include "deltasph_impl.h"

namespace dsph {

int scan_exclusive(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws(int64_t n);

namespace {

// Per-domain grid and offsets.
struct BGrid {
  int nc[2];
  float hcell;
  int cell_off;  // first global cell
  int first;     // first particle
  int count;     // particles
  int sq_off;    // first block of the domain's sum-of-squares partials
  int sq_nblk;   // blocks of the domain's sum of squares
};

// The single-domain sum of squares runs 1024 blocks of 256 threads striding
// over the array; the batched one reproduces that order per domain.
constexpr int kSqBlocks = 1024;
constexpr int64_t kSqStride = (int64_t)kSqBlocks * kBlock;

inline size_t a256(size_t b) { return (b + 255) / 256 * 256; }

struct Layout {
  size_t doms, grids, dom_of, slice, cnt, nbr, keys, slot, si, cell_count, cell_start, scan, wq, spos;
  size_t Bpos, Bvel, Bmass, Brho, Buid;
  size_t Spos[3], Svel[3], Srho[3];
  size_t ku[4], kr[4];
  size_t grad, vol, avg, sq_part, sq_dom, status;
  size_t total;
  int64_t npad, sq_cap;
};

Layout make_layout(int64_t n, int ndom, int64_t ncell, int width) {
  Layout L;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t r = o;
    o += a256(bytes);
    return r;
  };
  L.npad = (n + 63) / 64 * 64;
  L.sq_cap = (2 * n + kBlock - 1) / kBlock + ndom + 1;
  L.doms = take(sizeof(Dom) * ndom);
  L.grids = take(sizeof(BGrid) * ndom);
  L.dom_of = take(4 * n);
  L.slice = take(4 * (L.npad / 64 + 1));
  L.cnt = take(4 * n);
  L.nbr = take((size_t)4 * L.npad * width);
  L.keys = take(4 * n);
  L.slot = take(4 * n);
  L.si = take(4 * n);
  L.cell_count = take(4 * ncell);
  L.cell_start = take(4 * (ncell + 1));
  L.scan = take(scan_ws(ncell));
  L.wq = take(8 * n);
  L.spos = take(8 * n);
  L.Bpos = take(8 * n);
  L.Bvel = take(8 * n);
  L.Bmass = take(4 * n);
  L.Brho = take(4 * n);
  L.Buid = take(4 * n);
  for (int s = 0; s < 3; ++s) {
    L.Spos[s] = take(8 * n);
    L.Svel[s] = take(8 * n);
    L.Srho[s] = take(4 * n);
  }
  for (int s = 0; s < 4; ++s) {
    L.ku[s] = take(8 * n);
    L.kr[s] = take(4 * n);
  }
  L.grad = take(8 * n);
  L.vol = take(4 * n);
  L.avg = take(8 * n);
  L.sq_part = take(8 * L.sq_cap);
  L.sq_dom = take(4 * L.sq_cap);
  L.status = take(16);
  L.total = o;
  return L;
}

template <class T>
inline T* at(void* ws, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(ws) + off);
}

#ifndef DSPH_BFILL_STAGE
#define DSPH_BFILL_STAGE 48
#endif
constexpr int kBFillStage = DSPH_BFILL_STAGE;

__device__ __forceinline__ int bcell(float x, const Dom& d, int k, const BGrid& g) {
  // clamped in float before the conversion (NaN and out-of-range
  // coordinates land in an edge cell)
  const float f = fminf(fmaxf(floorf(__fdiv_rn(x - d.min[k], g.hcell)), 0.0f), (float)(g.nc[k] - 1));
  return (int)f;
}

__global__ void __launch_bounds__(kBlock) k_fill_i32(int32_t* __restrict__ a, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Wrapped position and global cell of every particle; provisional slot by
// atomics (the in-cell sort fixes the final order).
__global__ void __launch_bounds__(kBlock) k_bkeys(int n, const float2* __restrict__ x,
                                                  const int32_t* __restrict__ dom_of, const Dom* __restrict__ doms,
                                                  const BGrid* __restrict__ grids, float2* __restrict__ wq,
                                                  int32_t* __restrict__ keys, int32_t* __restrict__ slot,
                                                  int32_t* __restrict__ cell_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  int key = 0;
  if (valid) {
    const int b = dom_of[i];
    const Dom d = doms[b];
    const BGrid g = grids[b];
    const float2 p = x[i];
    const float2 w = make_float2(wrap_coord(p.x, d, 0), wrap_coord(p.y, d, 1));
    wq[i] = w;
    key = g.cell_off + bcell(w.y, d, 1, g) * g.nc[0] + bcell(w.x, d, 0, g);
    keys[i] = key;
  }
  // one atomic per distinct cell of the wave (the state is nearly cell-ordered)
  const int sl = wave_slot(cell_count, key, valid);
  if (valid) slot[i] = sl;
}

__global__ void __launch_bounds__(kBlock) k_bscatter(int n, const int32_t* __restrict__ keys,
                                                     const int32_t* __restrict__ slot,
                                                     const int32_t* __restrict__ cell_start,
                                                     int32_t* __restrict__ si) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  si[cell_start[keys[i]] + slot[i]] = i;
}

// Insertion sort of each cell by particle index (deterministic order).
__global__ void __launch_bounds__(kBlock) k_bcellsort(int ncell, const int32_t* __restrict__ cell_start,
                                                      int32_t* __restrict__ si) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int b = cell_start[c], e = cell_start[c + 1];
  for (int a = b + 1; a < e; ++a) {
    const int v = si[a];
    int k = a - 1;
    while (k >= b && si[k] > v) {
      si[k + 1] = si[k];
      --k;
    }
    si[k + 1] = v;
  }
}

// State A -> B in cell order; wout[p] = wrapped position of B's particle p.
__global__ void __launch_bounds__(kBlock) k_bpermute(int n, const int32_t* __restrict__ si,
                                                     const float2* __restrict__ pos, const float2* __restrict__ vel,
                                                     const float* __restrict__ mass, const float* __restrict__ rho,
                                                     const int32_t* __restrict__ uid,
                                                     const float2* __restrict__ win, float2* __restrict__ Bpos,
                                                     float2* __restrict__ Bvel, float* __restrict__ Bmass,
                                                     float* __restrict__ Brho, int32_t* __restrict__ Buid,
                                                     float2* __restrict__ wout) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = si[p];
  Bpos[p] = pos[i];
  Bvel[p] = vel[i];
  Bmass[p] = mass[i];
  Brho[p] = rho[i];
  Buid[p] = uid[i];
  wout[p] = win[i];
}

__global__ void __launch_bounds__(kBlock) k_bgather2(int n, const int32_t* __restrict__ si,
                                                     const float2* __restrict__ src, float2* __restrict__ dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  dst[p] = src[si[p]];
}

// A candidate: sorted wrapped position and the state index it maps to.
struct BCand {
  float2 x;
  int idx;
};

// Row i (state order) of the fixed-width sliced ELL: entry k at
// (i/64)*64*width + 64k + i%64, candidates in the single-domain search's
// order (cell rows dy = -1, 0, 1; cells dx = -1, 0, 1, a non-wrapping row as
// one contiguous slice), entries mapped through si to state indices (si ==
// NULL: the sorted order is the state order).  Rows longer than `width`
// keep their first `width` entries and are counted in status[0]; status[1]
// is the longest row, status[2..3] (uint64) the number of pairs found.
__global__ void __launch_bounds__(kBlock) k_bfill(int n, const int32_t* __restrict__ dom_of,
                                                  const Dom* __restrict__ doms, const BGrid* __restrict__ grids,
                                                  const float2* __restrict__ q, const float2* __restrict__ spos,
                                                  const int32_t* __restrict__ cs, const int32_t* __restrict__ si,
                                                  float h, int width, int32_t* __restrict__ nbr,
                                                  int32_t* __restrict__ cnt, int32_t* __restrict__ status) {
  __shared__ int sh_max[kBlock / 64];
  __shared__ int sh_over[kBlock / 64];
  __shared__ int sh_sum[kBlock / 64];
  // the first kBFillStage entries of a row are staged in LDS as 16-bit
  // offsets idx - i (a per-thread column) and written column by column after
  // the scan: one contiguous line per store of the wave instead of scattered
  // words; an offset beyond 16 bits is written at once (slot marked)
  __shared__ int16_t stage[kBFillStage + 1][kBlock];  // + the sink row of the unconditional writes
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int total = 0;
  if (i < n) {
    const int b = dom_of[i];
    const Dom d = doms[b];
    const BGrid g = grids[b];
    const float2 qi = q[i];
    const int64_t base = (int64_t)(i >> 6) * 64 * width + (i & 63);
    const int ncx = g.nc[0], ncy = g.nc[1];
    const int cx = bcell(qi.x, d, 0, g), cy = bcell(qi.y, d, 1, g);
    // every candidate writes the staging slot of the next entry (the sink
    // row past the staged ones); a kept candidate advances `total`, so its
    // value stays
    auto take = [&](bool kp, int idx) {
      const int di = idx - i;
      const bool fits = di > -32768 && di <= 32767;
      const bool st = total < kBFillStage && total < width;
      stage[st ? total : kBFillStage][threadIdx.x] = fits ? (int16_t)di : (int16_t)-32768;
      if (kp && !(st && fits) && total < width) nbr[base + (int64_t)total * 64] = idx;
      total += kp ? 1 : 0;
    };
    auto load = [&](int j) { return BCand{spos[j], si ? si[j] : j}; };
    // off a periodic seam (a periodic axis of 7+ cells, no wrapped row or
    // column) the plain difference is the min-image offset
    const bool small = (d.periodic[0] && g.nc[0] < 7) || (d.periodic[1] && g.nc[1] < 7);
    auto visit = [&](int jb, int je, bool seam) {
      if (seam)
        scan_range<4>(jb, je, load, [&](int, const BCand& c) {
          const float dx = mod_offset(qi.x - c.x.x, d, 0);
          const float dy = mod_offset(qi.y - c.x.y, d, 1);
          take(within_support(dist2_exact(dx, dy), h), c.idx);
        });
      else
        scan_range<4>(jb, je, load, [&](int, const BCand& c) {
          take(within_support(dist2_exact(qi.x - c.x.x, qi.y - c.x.y), h), c.idx);
        });
    };
    for (int oy = -1; oy <= 1; ++oy) {
      int yy = cy + oy;
      if (d.periodic[1]) {
        yy = yy < 0 ? yy + ncy : (yy >= ncy ? yy - ncy : yy);
      } else if (yy < 0 || yy >= ncy) {
        continue;
      }
      const int row = g.cell_off + yy * ncx;
      const bool wraps = d.periodic[0] && (cx == 0 || cx == ncx - 1);
      const bool seam = small || wraps || yy != cy + oy;
      if (!wraps) {
        const int x0 = cx - 1 < 0 ? 0 : cx - 1;
        const int x1 = cx + 1 > ncx - 1 ? ncx - 1 : cx + 1;
        visit(cs[row + x0], cs[row + x1 + 1], seam);
      } else {
        for (int ox = -1; ox <= 1; ++ox) {
          int xx = cx + ox;
          xx = xx < 0 ? xx + ncx : (xx >= ncx ? xx - ncx : xx);
          visit(cs[row + xx], cs[row + xx + 1], seam);
        }
      }
    }
    const int staged = total < kBFillStage ? total : kBFillStage;
    for (int t = 0; t < staged && t < width; ++t) {
      const int16_t v = stage[t][threadIdx.x];
      if (v != (int16_t)-32768) nbr[base + (int64_t)t * 64] = i + v;
    }
    cnt[i] = total < width ? total : width;
  }
  int m = total, over = total > width ? 1 : 0, sum = total;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int t = __shfl_xor(m, o, 64);
    m = t > m ? t : m;
    over += __shfl_xor(over, o, 64);
    sum += __shfl_xor(sum, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh_max[threadIdx.x >> 6] = m;
    sh_over[threadIdx.x >> 6] = over;
    sh_sum[threadIdx.x >> 6] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int bm = 0, bo = 0;
    unsigned long long bs = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      bm = sh_max[w] > bm ? sh_max[w] : bm;
      bo += sh_over[w];
      bs += (unsigned long long)sh_sum[w];
    }
    if (bo) atomicAdd(&status[0], bo);
    atomicMax(&status[1], bm);
    atomicAdd(reinterpret_cast<unsigned long long*>(status + 2), bs);
  }
}

// Stage state = B + k c with c = dt_d * f (f = 0.5 or 1), the single-domain
// rk_update with c formed as float32 dt * 0.5.
__global__ void __launch_bounds__(kBlock) k_brk_update(int n, const int32_t* __restrict__ dom_of,
                                                       const float* __restrict__ dt, int half,
                                                       const float2* __restrict__ x0, const float2* __restrict__ kx,
                                                       float2* __restrict__ x, const float2* __restrict__ u0,
                                                       const float2* __restrict__ ku, float2* __restrict__ u,
                                                       const float* __restrict__ r0, const float* __restrict__ kr,
                                                       float* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = dt[dom_of[i]];
  const float c = half ? t * 0.5f : t;
  const float2 a = x0[i], ka = kx[i], b = u0[i], kb = ku[i];
  x[i] = make_float2(a.x + ka.x * c, a.y + ka.y * c);
  u[i] = make_float2(b.x + kb.x * c, b.y + kb.y * c);
  r[i] = r0[i] + kr[i] * c;
}

__device__ __forceinline__ float rk4_avg_b(float a, float b, float c, float d) {
  return (((a + 2.0f * b) + 2.0f * c) + d) / 6.0f;
}

struct CombineArgs {
  const float2* kx[4];
  const float2* ku[4];
  const float* kr[4];
  const float2 *Bpos, *Bvel;
  const float *Bmass, *Brho;
  const int32_t* Buid;
  float2 *pos, *vel;
  float *mass, *rho;
  int32_t* uid;
  float2* avg;
  const int32_t* dom_of;
  const float* dt;
  int n;
};

__global__ void __launch_bounds__(kBlock) k_bcombine(CombineArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const float t = A.dt[A.dom_of[i]];
  {
    const float2 a = A.kx[0][i], b = A.kx[1][i], c = A.kx[2][i], d = A.kx[3][i];
    const float2 x = A.Bpos[i];
    const float dx = rk4_avg_b(a.x, b.x, c.x, d.x), dy = rk4_avg_b(a.y, b.y, c.y, d.y);
    A.pos[i] = make_float2(x.x + dx * t, x.y + dy * t);
  }
  {
    const float2 a = A.ku[0][i], b = A.ku[1][i], c = A.ku[2][i], d = A.ku[3][i];
    const float2 u = A.Bvel[i];
    const float dx = rk4_avg_b(a.x, b.x, c.x, d.x), dy = rk4_avg_b(a.y, b.y, c.y, d.y);
    A.vel[i] = make_float2(u.x + dx * t, u.y + dy * t);
    A.avg[i] = make_float2(dx, dy);
  }
  {
    const float dr = rk4_avg_b(A.kr[0][i], A.kr[1][i], A.kr[2][i], A.kr[3][i]);
    A.rho[i] = A.Brho[i] + dr * t;
  }
  A.mass[i] = A.Bmass[i];
  A.uid[i] = A.Buid[i];
}

// Per-domain float64 sum of squares of the 2*count floats of the averaged
// dudt: block lb of domain d covers elements lb*256 + t + k*262144, reduced
// by the same wave/block tree as the single-domain reduction.
__global__ void __launch_bounds__(kBlock) k_bsumsq(const int32_t* __restrict__ sq_dom,
                                                   const BGrid* __restrict__ grids, const float* __restrict__ v,
                                                   double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  const int b = sq_dom[blockIdx.x];
  if (b < 0) return;  // whole block exits together
  const BGrid g = grids[b];
  const int64_t lb = (int64_t)blockIdx.x - g.sq_off;
  const int64_t len = 2 * (int64_t)g.count;
  const float* src = v + 2 * (int64_t)g.first;
  double s = 0.0;
  for (int64_t i = lb * kBlock + threadIdx.x; i < len; i += kSqStride) {
    const double a = (double)src[i];
    if (a == a) s += a * a;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += sh[w];
    part[blockIdx.x] = t;
  }
}

// computeTimestep per domain (float32, the order of the single-domain host
// formula); time += dt.
__global__ void __launch_bounds__(kBlock) k_bdt(int ndom, const BGrid* __restrict__ grids,
                                                const double* __restrict__ part, dsph_timestep_params tp,
                                                float* __restrict__ dt, double* __restrict__ time) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ndom) return;
  const BGrid g = grids[b];
  double s = 0.0;
  for (int k = 0; k < g.sq_nblk; ++k) s += part[g.sq_off + k];
  const float used = dt[b];
  if (time) time[b] = time[b] + (double)used;
  if (!tp.adaptive) return;
  const float amax = (float)sqrt(s);
  const float dt_a = (0.25f * sqrtf(tp.support / (amax + 1e-7f))) / tp.kernel_scale;
  float nw = used;
  if (tp.viscosity_constraint) nw = tp.dt_visc;
  if (tp.acceleration_constraint) nw = dt_a < nw ? dt_a : nw;
  if (tp.acoustic_constraint) nw = tp.dt_acoustic < nw ? tp.dt_acoustic : nw;
  nw = tp.max_dt < nw ? tp.max_dt : nw;
  nw = tp.min_dt > nw ? tp.min_dt : nw;
  dt[b] = nw;
}

struct Ctx {
  const dsph_batch* b;
  Layout L;
  hipStream_t st;
  int n, ndom, ncell, block;
  void* ws;
  template <class T>
  T* p(size_t off) const {
    return at<T>(ws, off);
  }
  int grid(int64_t m) const { return grid_for(m, block); }
};

// Cell grid of positions x (state order): wrapped positions into wq, sorted
// slots into si, cell starts; with permute, also the state A -> B.
int build_grid(const Ctx& c, const float2* x, float2* wq) {
  const Layout& L = c.L;
  int32_t* cc = c.p<int32_t>(L.cell_count);
  k_fill_i32<<<grid_for(c.ncell), kBlock, 0, c.st>>>(cc, c.ncell, 0);
  k_bkeys<<<c.grid(c.n), c.block, 0, c.st>>>(c.n, x, c.p<int32_t>(L.dom_of), c.p<Dom>(L.doms),
                                             c.p<BGrid>(L.grids), wq, c.p<int32_t>(L.keys), c.p<int32_t>(L.slot),
                                             cc);
  DSPH_LAUNCH_CHECK();
  int s = scan_exclusive(cc, c.ncell, c.p<int32_t>(L.cell_start), c.p<void>(L.scan), scan_ws(c.ncell), c.st);
  if (s != DSPH_OK) return s;
  k_bscatter<<<c.grid(c.n), c.block, 0, c.st>>>(c.n, c.p<int32_t>(L.keys), c.p<int32_t>(L.slot),
                                                c.p<int32_t>(L.cell_start), c.p<int32_t>(L.si));
  k_bcellsort<<<c.grid(c.ncell), c.block, 0, c.st>>>(c.ncell, c.p<int32_t>(L.cell_start), c.p<int32_t>(L.si));
  return launch_status();
}

int fill_rows(const Ctx& c, const float2* q, const float2* spos, const int32_t* si) {
  const Layout& L = c.L;
  k_bfill<<<c.grid(c.n), c.block, 0, c.st>>>(c.n, c.p<int32_t>(L.dom_of), c.p<Dom>(L.doms), c.p<BGrid>(L.grids), q,
                                             spos, c.p<int32_t>(L.cell_start), si, c.b->support, c.b->width,
                                             c.p<int32_t>(L.nbr), c.p<int32_t>(L.cnt), c.p<int32_t>(L.status));
  return launch_status();
}

PairsDev batch_pairs(const Ctx& c, const float2* wpos) {
  PairsDev P;
  P.nq = P.nr = c.n;
  P.off = nullptr;
  P.nbr = c.p<int32_t>(c.L.nbr);
  P.qpos = P.rpos = wpos;
  P.qh = P.rh = nullptr;
  P.h0 = c.b->support;
  P.mode = DSPH_MODE_SCATTER;
  P.dom = Dom{};
  P.cnt = c.p<int32_t>(c.L.cnt);
  P.slice = c.p<int32_t>(c.L.slice);
  return P;
}

// One right-hand side at (wpos, vel, rho) into ku, kr.
int stage_rhs(const Ctx& c, const dsph_deltasph_params* prm, const float2* wpos, const float2* vel,
              const float* rho, float2* ku, float* kr) {
  const Layout& L = c.L;
  const PairsDev P = batch_pairs(c, wpos);
  const float* mass = c.p<float>(L.Bmass);
  float2* grad = c.p<float2>(L.grad);
  float* vol = c.p<float>(L.vol);
  int s;
  if (prm->diffusion) {
    // V_j = m_j / rho_j once per particle for the normalization pass
    s = launch_volume(c.n, mass, rho, vol, c.st);
    if (s != DSPH_OK) return s;
    s = launch_normalization_per_particle(P, c.p<Dom>(L.doms), c.p<int32_t>(L.dom_of), mass, rho, vol, grad, c.block,
                                          c.st);
    if (s != DSPH_OK) return s;
  }
  RhsArgs A;
  s = make_rhs_args(prm, reinterpret_cast<const float*>(vel), mass, rho, reinterpret_cast<const float*>(grad),
                    reinterpret_cast<float*>(ku), kr, nullptr, &A);
  if (s != DSPH_OK) return s;
  return launch_rhs_per_particle(P, A, c.p<Dom>(L.doms), c.p<int32_t>(L.dom_of), c.block, c.st);
}

int one_step(const Ctx& c, const dsph_deltasph_params* prm, const dsph_timestep_params* tp, float2* pos, float2* vel,
             float* mass, float* rho, int32_t* uid, float* dt, double* time) {
  const Layout& L = c.L;
  float2* wq = c.p<float2>(L.wq);
  float2* spos = c.p<float2>(L.spos);
  int32_t* si = c.p<int32_t>(L.si);
  const int32_t* dom_of = c.p<int32_t>(L.dom_of);
  float2* Bpos = c.p<float2>(L.Bpos);
  float2* Bvel = c.p<float2>(L.Bvel);
  float* Brho = c.p<float>(L.Brho);
  // sort A into cell order -> B; wq = wrapped B positions
  int s = build_grid(c, pos, spos);
  if (s != DSPH_OK) return s;
  k_bpermute<<<c.grid(c.n), c.block, 0, c.st>>>(c.n, si, pos, vel, mass, rho, uid, spos, Bpos, Bvel,
                                                c.p<float>(L.Bmass), Brho, c.p<int32_t>(L.Buid), wq);
  DSPH_LAUNCH_CHECK();
  const float2* sx[4] = {Bpos, c.p<float2>(L.Spos[0]), c.p<float2>(L.Spos[1]), c.p<float2>(L.Spos[2])};
  const float2* su[4] = {Bvel, c.p<float2>(L.Svel[0]), c.p<float2>(L.Svel[1]), c.p<float2>(L.Svel[2])};
  const float* sr[4] = {Brho, c.p<float>(L.Srho[0]), c.p<float>(L.Srho[1]), c.p<float>(L.Srho[2])};
  for (int k = 0; k < 4; ++k) {
    if (k == 0) {
      // the sorted state's grid is the sort's own (identity order)
      s = fill_rows(c, wq, wq, nullptr);
    } else {
      s = build_grid(c, sx[k], wq);
      if (s != DSPH_OK) return s;
      k_bgather2<<<c.grid(c.n), c.block, 0, c.st>>>(c.n, si, wq, spos);
      DSPH_LAUNCH_CHECK();
      s = fill_rows(c, wq, spos, si);
    }
    if (s != DSPH_OK) return s;
    s = stage_rhs(c, prm, wq, su[k], sr[k], c.p<float2>(L.ku[k]), c.p<float>(L.kr[k]));
    if (s != DSPH_OK) return s;
    if (k < 3) {
      k_brk_update<<<c.grid(c.n), c.block, 0, c.st>>>(
          c.n, dom_of, dt, k < 2 ? 1 : 0, Bpos, su[k], c.p<float2>(L.Spos[k]), Bvel, c.p<float2>(L.ku[k]),
          c.p<float2>(L.Svel[k]), Brho, c.p<float>(L.kr[k]), c.p<float>(L.Srho[k]));
      DSPH_LAUNCH_CHECK();
    }
  }
  CombineArgs A;
  for (int k = 0; k < 4; ++k) {
    A.kx[k] = su[k];
    A.ku[k] = c.p<float2>(L.ku[k]);
    A.kr[k] = c.p<float>(L.kr[k]);
  }
  A.Bpos = Bpos;
  A.Bvel = Bvel;
  A.Bmass = c.p<float>(L.Bmass);
  A.Brho = Brho;
  A.Buid = c.p<int32_t>(L.Buid);
  A.pos = pos;
  A.vel = vel;
  A.mass = mass;
  A.rho = rho;
  A.uid = uid;
  A.avg = c.p<float2>(L.avg);
  A.dom_of = dom_of;
  A.dt = dt;
  A.n = c.n;
  k_bcombine<<<c.grid(c.n), c.block, 0, c.st>>>(A);
  k_bsumsq<<<(int)L.sq_cap, kBlock, 0, c.st>>>(c.p<int32_t>(L.sq_dom), c.p<BGrid>(L.grids), c.p<float>(L.avg),
                                               c.p<double>(L.sq_part));
  k_bdt<<<grid_for(c.ndom), kBlock, 0, c.st>>>(c.ndom, c.p<BGrid>(L.grids), c.p<double>(L.sq_part), *tp, dt, time);
  return launch_status();
}

int check_batch(const dsph_batch* b) {
  if (!b || !b->workspace || b->n < 1 || b->ndom < 1 || b->ncell < 1 || b->width < 1) return DSPH_ERR_ARG;
  if (b->block != 64 && b->block != 128 && b->block != 256) return DSPH_ERR_ARG;
  if (b->n >= INT32_MAX || b->ncell >= INT32_MAX) return DSPH_ERR_ARG;
  const Layout L = make_layout(b->n, b->ndom, b->ncell, b->width);
  if (b->workspace_bytes < L.total) return DSPH_ERR_CAPACITY;
  return DSPH_OK;
}

Ctx make_ctx(const dsph_batch* b, hipStream_t st) {
  Ctx c;
  c.b = b;
  c.L = make_layout(b->n, b->ndom, b->ncell, b->width);
  c.st = st;
  c.n = (int)b->n;
  c.ndom = b->ndom;
  c.ncell = (int)b->ncell;
  c.block = b->block;
  c.ws = b->workspace;
  return c;
}

int run_steps(const dsph_batch* b, const dsph_deltasph_params* prm, const dsph_timestep_params* tp, float* pos,
              float* vel, float* mass, float* rho, int32_t* uid, float* dt, int32_t steps, double* time,
              hipStream_t st) {
  int s = check_batch(b);
  if (s != DSPH_OK) return s;
  if (!prm || !tp || !pos || !vel || !mass || !rho || !uid || !dt || steps < 0) return DSPH_ERR_ARG;
  if (prm->viscosity < -1 || prm->viscosity > 1) return DSPH_ERR_ARG;
  if (!(tp->support > 0.0f) || !(tp->kernel_scale > 0.0f)) return DSPH_ERR_ARG;
  const Ctx c = make_ctx(b, st);
  for (int k = 0; k < steps; ++k) {
    s = one_step(c, prm, tp, reinterpret_cast<float2*>(pos), reinterpret_cast<float2*>(vel), mass, rho, uid, dt,
                 time);
    if (s != DSPH_OK) return s;
  }
  return DSPH_OK;
}

// Domain tables of a batch (host side).
struct Plan {
  std::vector<Dom> doms;
  std::vector<BGrid> grids;
  int64_t n = 0, ncell = 0;
};

int make_plan(const dsph_domain* domains, const int64_t* counts, int32_t ndom, float hsearch, Plan* out) {
  if (!domains || !counts || ndom < 1 || !(hsearch > 0.0f)) return DSPH_ERR_ARG;
  Plan& P = *out;
  P.doms.resize(ndom);
  P.grids.resize(ndom);
  int64_t n = 0, ncell = 0, sq = 0;
  for (int b = 0; b < ndom; ++b) {
    const dsph_domain& d = domains[b];
    if (!d.periodic[0] || !d.periodic[1] || counts[b] < 1) return DSPH_ERR_ARG;
    dsph_grid g;
    int s = dsph_grid_dims(&d, hsearch, &g);
    if (s != DSPH_OK) return s;
    P.doms[b] = make_dom(d);
    BGrid& G = P.grids[b];
    G.nc[0] = g.ncell[0];
    G.nc[1] = g.ncell[1];
    G.hcell = g.hcell;
    G.cell_off = (int)ncell;
    G.first = (int)n;
    G.count = (int)counts[b];
    const int64_t len = 2 * counts[b];
    int64_t nb = (len + kBlock - 1) / kBlock;
    if (nb > kSqBlocks) nb = kSqBlocks;
    G.sq_off = (int)sq;
    G.sq_nblk = (int)nb;
    n += counts[b];
    ncell += (int64_t)g.ncell[0] * g.ncell[1];
    sq += nb;
    if (n >= INT32_MAX || ncell >= INT32_MAX) return DSPH_ERR_DIM;
  }
  P.n = n;
  P.ncell = ncell;
  return DSPH_OK;
}

}  // namespace
}  // namespace dsph

using namespace dsph;

extern "C" size_t dsph_batch_workspace(const dsph_domain* domains, const int64_t* counts, int32_t ndom,
                                       float hsearch, int32_t width) {
  Plan P;
  if (width < 1 || make_plan(domains, counts, ndom, hsearch, &P) != DSPH_OK) return 0;
  const int64_t npad = (P.n + 63) / 64 * 64;
  if (npad * (int64_t)width >= INT32_MAX) return 0;
  return make_layout(P.n, ndom, P.ncell, width).total;
}

extern "C" int dsph_batch_init(const dsph_domain* domains, const int64_t* counts, int32_t ndom, float support,
                               float hsearch, int32_t width, int32_t block, void* workspace, size_t workspace_bytes,
                               dsph_batch* out) {
  if (!out || !workspace || width < 1 || !(support > 0.0f) || !(hsearch >= support)) return DSPH_ERR_ARG;
  if (block != 64 && block != 128 && block != 256) return DSPH_ERR_ARG;
  Plan P;
  int s = make_plan(domains, counts, ndom, hsearch, &P);
  if (s != DSPH_OK) return s;
  const int64_t npad = (P.n + 63) / 64 * 64;
  if (npad * (int64_t)width >= INT32_MAX) return DSPH_ERR_CAPACITY;
  const Layout L = make_layout(P.n, ndom, P.ncell, width);
  if (workspace_bytes < L.total) return DSPH_ERR_CAPACITY;
  std::vector<int32_t> dom_of(P.n), slice(npad / 64 + 1), sq_dom(L.sq_cap, -1);
  for (int b = 0; b < ndom; ++b) {
    const BGrid& G = P.grids[b];
    for (int k = 0; k < G.count; ++k) dom_of[G.first + k] = b;
    for (int k = 0; k < G.sq_nblk; ++k) sq_dom[G.sq_off + k] = b;
  }
  for (int64_t k = 0; k <= npad / 64; ++k) slice[k] = (int32_t)(k * 64 * width);
  DSPH_CHECK_HIP(hipMemcpy(at<void>(workspace, L.doms), P.doms.data(), sizeof(Dom) * ndom, hipMemcpyHostToDevice));
  DSPH_CHECK_HIP(
      hipMemcpy(at<void>(workspace, L.grids), P.grids.data(), sizeof(BGrid) * ndom, hipMemcpyHostToDevice));
  DSPH_CHECK_HIP(hipMemcpy(at<void>(workspace, L.dom_of), dom_of.data(), 4 * P.n, hipMemcpyHostToDevice));
  DSPH_CHECK_HIP(hipMemcpy(at<void>(workspace, L.slice), slice.data(), 4 * slice.size(), hipMemcpyHostToDevice));
  DSPH_CHECK_HIP(hipMemcpy(at<void>(workspace, L.sq_dom), sq_dom.data(), 4 * sq_dom.size(), hipMemcpyHostToDevice));
  const int32_t zero[4] = {0, 0, 0, 0};
  DSPH_CHECK_HIP(hipMemcpy(at<void>(workspace, L.status), zero, sizeof(zero), hipMemcpyHostToDevice));
  out->n = P.n;
  out->ncell = P.ncell;
  out->ndom = ndom;
  out->width = width;
  out->block = block;
  out->support = support;
  out->hsearch = hsearch;
  out->pad_ = 0;
  out->workspace = workspace;
  out->workspace_bytes = workspace_bytes;
  return DSPH_OK;
}

extern "C" int dsph_batch_step(const dsph_batch* batch, const dsph_deltasph_params* prm,
                               const dsph_timestep_params* tp, float* pos, float* vel, float* mass, float* rho,
                               int32_t* uid, float* dt, double* time, int32_t steps, void* stream) {
  return run_steps(batch, prm, tp, pos, vel, mass, rho, uid, dt, steps, time, as_stream(stream));
}

extern "C" int dsph_batch_graph(const dsph_batch* batch, const dsph_deltasph_params* prm,
                                const dsph_timestep_params* tp, float* pos, float* vel, float* mass, float* rho,
                                int32_t* uid, float* dt, double* time, int32_t steps, void** graph) {
  if (!graph || steps < 1) return DSPH_ERR_ARG;
  *graph = nullptr;
  int s = check_batch(batch);
  if (s != DSPH_OK) return s;
  hipStream_t st;
  DSPH_CHECK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipStreamDestroy(st);
    return DSPH_ERR_LAUNCH;
  }
  s = run_steps(batch, prm, tp, pos, vel, mass, rho, uid, dt, steps, time, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  (void)hipStreamDestroy(st);
  if (s == DSPH_OK && (e != hipSuccess || !g)) s = DSPH_ERR_LAUNCH;
  if (s != DSPH_OK) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return DSPH_ERR_LAUNCH;
  *graph = ex;
  return DSPH_OK;
}

extern "C" int dsph_graph_launch(void* graph, void* stream) {
  if (!graph) return DSPH_ERR_ARG;
  DSPH_CHECK_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph), as_stream(stream)));
  return DSPH_OK;
}

extern "C" int dsph_graph_destroy(void* graph) {
  if (!graph) return DSPH_OK;
  DSPH_CHECK_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)));
  return DSPH_OK;
}

extern "C" int dsph_batch_status(const dsph_batch* batch, int32_t* status_host, int32_t reset, void* stream) {
  int s = check_batch(batch);
  if (s != DSPH_OK) return s;
  if (!status_host) return DSPH_ERR_ARG;
  hipStream_t st = as_stream(stream);
  const Layout L = make_layout(batch->n, batch->ndom, batch->ncell, batch->width);
  int32_t* dev = at<int32_t>(batch->workspace, L.status);
  DSPH_CHECK_HIP(hipMemcpyAsync(status_host, dev, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (reset) {
    k_fill_i32<<<1, kBlock, 0, st>>>(dev, 4, 0);
    DSPH_LAUNCH_CHECK();
  }
  DSPH_CHECK_HIP(hipStreamSynchronize(st));
  return DSPH_OK;
}

extern "C" int dsph_batch_replay_pass(const dsph_batch* batch, const dsph_deltasph_params* prm, int32_t pass,
                                      void* stream) {
  int s = check_batch(batch);
  if (s != DSPH_OK) return s;
  if (!prm || pass < 0 || pass > 2) return DSPH_ERR_ARG;
  const Ctx c = make_ctx(batch, as_stream(stream));
  const Layout& L = c.L;
  float2* wq = c.p<float2>(L.wq);
  if (pass == 0) return fill_rows(c, wq, c.p<float2>(L.spos), c.p<int32_t>(L.si));
  const PairsDev P = batch_pairs(c, wq);
  const float* rho = c.p<float>(L.Srho[2]);
  if (pass == 1)
    return launch_normalization_per_particle(P, c.p<Dom>(L.doms), c.p<int32_t>(L.dom_of), c.p<float>(L.Bmass), rho,
                                             c.p<float>(L.vol), c.p<float2>(L.grad), c.block, c.st);
  RhsArgs A;
  s = make_rhs_args(prm, c.p<float>(L.Svel[2]), c.p<float>(L.Bmass), rho, c.p<float>(L.grad), c.p<float>(L.ku[3]),
                    c.p<float>(L.kr[3]), nullptr, &A);
  if (s != DSPH_OK) return s;
  return launch_rhs_per_particle(P, A, c.p<Dom>(L.doms), c.p<int32_t>(L.dom_of), c.block, c.st);
}

extern "C" int dsph_batch_probe(const dsph_batch* batch, const float* pos, int32_t* max_count_host,
                                int64_t* pairs_host, void* stream) {
  int s = check_batch(batch);
  if (s != DSPH_OK) return s;
  if (!pos || !max_count_host) return DSPH_ERR_ARG;
  hipStream_t st = as_stream(stream);
  const Ctx c = make_ctx(batch, st);
  int32_t* status = c.p<int32_t>(c.L.status);
  k_fill_i32<<<1, kBlock, 0, st>>>(status, 4, 0);
  float2* wq = c.p<float2>(c.L.wq);
  s = build_grid(c, reinterpret_cast<const float2*>(pos), wq);
  if (s != DSPH_OK) return s;
  k_bgather2<<<c.grid(c.n), c.block, 0, st>>>(c.n, c.p<int32_t>(c.L.si), wq, c.p<float2>(c.L.spos));
  s = fill_rows(c, wq, c.p<float2>(c.L.spos), c.p<int32_t>(c.L.si));
  if (s != DSPH_OK) return s;
  int32_t h[4] = {0, 0, 0, 0};
  DSPH_CHECK_HIP(hipMemcpyAsync(h, status, sizeof(h), hipMemcpyDeviceToHost, st));
  k_fill_i32<<<1, kBlock, 0, st>>>(status, 4, 0);
  DSPH_LAUNCH_CHECK();
  DSPH_CHECK_HIP(hipStreamSynchronize(st));
  *max_count_host = h[1];
  if (pairs_host) {
    uint64_t t;
    memcpy(&t, h + 2, sizeof(t));
    *pairs_host = (int64_t)t;
  }
  return DSPH_OK;
}
