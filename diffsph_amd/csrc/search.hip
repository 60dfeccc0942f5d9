// Cell-list neighbor search on a dense grid (gfx950).
//
// Build: key = cell of the wrapped position -> per-cell histogram (atomics
// give a provisional slot) -> exclusive scan -> scatter -> per-cell sort by
// original index (makes the order deterministic) -> gather positions and
// supports into cell order.
// Query: one thread per query particle walks the 3x3 cell block; the three
// cells of a row are one contiguous slice of the sorted arrays unless the
// row wraps around a periodic x boundary.
#include "common.h"

namespace dsph {

int scan_exclusive(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws(int64_t n);
// Tile run tables of a square list over the grid (k_tile_runs).
int launch_tile_runs(int64_t nq, const dsph_grid* grid, const dsph_domain* domain, const int32_t* cell_start,
                     int32_t* tile_runs, int32_t* tile_n, hipStream_t st);
int scan_read_total(const void* ws, int64_t* total_host, hipStream_t st);

struct GridDev {
  int nc[2];
  float hcell;
};

__device__ __forceinline__ int cell_coord(float x, const Dom& d, int k, const GridDev& g) {
  // clamped in float before the conversion: NaN (fmaxf returns the other
  // operand) and out-of-range coordinates land in an edge cell
  const float f = fminf(fmaxf(floorf(__fdiv_rn(x - d.min[k], g.hcell)), 0.0f), (float)(g.nc[k] - 1));
  return (int)f;
}

__global__ void __launch_bounds__(kBlock) k_cell_keys(const float2* __restrict__ pos, int n, Dom d, GridDev g,
                                                      int32_t* __restrict__ keys, int32_t* __restrict__ slot,
                                                      int32_t* __restrict__ cell_count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  int key = 0;
  if (valid) {
    float2 p = pos[i];
    float wx = wrap_coord(p.x, d, 0), wy = wrap_coord(p.y, d, 1);
    int cx = cell_coord(wx, d, 0, g), cy = cell_coord(wy, d, 1, g);
    key = cy * g.nc[0] + cx;
    keys[i] = key;
  }
  const int sl = wave_slot(cell_count, key, valid);
  if (valid) slot[i] = sl;
}

__global__ void __launch_bounds__(kBlock) k_cell_scatter(int n, const int32_t* __restrict__ keys,
                                                         const int32_t* __restrict__ slot,
                                                         const int32_t* __restrict__ cell_start,
                                                         int32_t* __restrict__ sort_index) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sort_index[cell_start[keys[i]] + slot[i]] = i;
}

// Each cell's slice sorted by original index (cells hold ~16 particles),
// one thread per cell.  A slice of up to kSortLds entries is loaded into
// the thread's LDS column with its loads in flight together, and each entry
// goes to its rank (the count of smaller entries: the indices are
// distinct); a longer one is insertion-sorted in place.
constexpr int kSortLds = 32;
__global__ void __launch_bounds__(kBlock) k_cell_sort(int ncell, const int32_t* __restrict__ cell_start,
                                                      int32_t* __restrict__ sort_index) {
  __shared__ int32_t col[kSortLds][kBlock];
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int b = cell_start[c], e = cell_start[c + 1];
  const int m = e - b;
  const int t = threadIdx.x;
  if (m > DSPH_CELL_MAX) return;  // a diverged state: the search reports it
  if (m <= kSortLds) {
    for (int a0 = 0; a0 < m; a0 += 8) {
      int v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = a0 + u < m ? sort_index[b + a0 + u] : 0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (a0 + u < m) col[a0 + u][t] = v[u];
    }
    for (int a = 0; a < m; ++a) {
      const int v = col[a][t];
      int r = 0;
      for (int u = 0; u < m; ++u) r += col[u][t] < v ? 1 : 0;
      sort_index[b + r] = v;
    }
    return;
  }
  for (int a = b + 1; a < e; ++a) {
    int v = sort_index[a];
    int k = a - 1;
    while (k >= b && sort_index[k] > v) {
      sort_index[k + 1] = sort_index[k];
      --k;
    }
    sort_index[k + 1] = v;
  }
}

__global__ void __launch_bounds__(kBlock) k_sorted_gather(int n, Dom d, const int32_t* __restrict__ sort_index,
                                                          const float2* __restrict__ pos, const float* __restrict__ h,
                                                          float2* __restrict__ spos, float* __restrict__ sh,
                                                          int32_t* __restrict__ rank) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i = sort_index[p];
  float2 x = pos[i];
  spos[p] = make_float2(wrap_coord(x.x, d, 0), wrap_coord(x.y, d, 1));
  if (sh) sh[p] = h[i];
  if (rank) rank[i] = p;
}

__global__ void __launch_bounds__(kBlock) k_wrap(const float2* __restrict__ pos, int n, Dom d,
                                                 float2* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float2 p = pos[i];
  out[i] = make_float2(wrap_coord(p.x, d, 0), wrap_coord(p.y, d, 1));
}

struct SearchArgs {
  Dom d;
  GridDev g;
  const int32_t* cs;
  const float2* rpos;
  const float* rh;
  const float2* qpos;
  const float* qh;
  float h0;
  int mode;
  int wrapped;
  int nq;
  int nref;
};

// A candidate reference particle: wrapped position and support.
struct Cand {
  float2 x;
  float h;
};

__device__ __forceinline__ bool keep_pair(float2 qi, float hi, float2 rj, float hj, const SearchArgs& a) {
  float dx = mod_offset(qi.x - rj.x, a.d, 0);
  float dy = mod_offset(qi.y - rj.y, a.d, 1);
  return within_support(dist2_exact(dx, dy), pick_h(a.mode, hi, hj));
}

// keep_pair for a candidate range that does not cross a periodic seam: the
// plain difference is the min-image offset there (mod_offset returns it
// unchanged), so it is not evaluated.
__device__ __forceinline__ bool keep_pair_near(float2 qi, float hi, float2 rj, float hj, const SearchArgs& a) {
  return within_support(dist2_exact(qi.x - rj.x, qi.y - rj.y), pick_h(a.mode, hi, hj));
}

// A periodic axis of fewer than 7 cells: adjacent cells (the last one up to
// two cells wide) may hold pairs farther apart than half the domain, so
// every range takes mod_offset.  From 7 cells on, two adjacent cells span
// less than L - 5 L / 8 < L / 2, so only ranges across the seam can.
__device__ __forceinline__ bool small_periodic_grid(const SearchArgs& a) {
  return (a.d.periodic[0] && a.g.nc[0] < 7) || (a.d.periodic[1] && a.g.nc[1] < 7);
}

// Visit the candidate slices of query i in offset order (dy=-1,0,1;
// dx=-1,0,1): f(b, e, seam), seam true when the slice's pairs may cross a
// periodic seam (a wrapped row or column of cells, or a small periodic grid).
// A slice longer than three cells of DSPH_CELL_MAX is skipped and sets
// `over`, as does a query cell of more than DSPH_CELL_MAX particles (a
// diverged state: the work stays bounded, the caller reports it).
constexpr int kRangeMax = 3 * DSPH_CELL_MAX;
template <class F>
__device__ __forceinline__ void for_candidates(const SearchArgs& a, float2 qi, bool& over, F&& f) {
  const int ncx = a.g.nc[0], ncy = a.g.nc[1];
  int cx = cell_coord(qi.x, a.d, 0, a.g), cy = cell_coord(qi.y, a.d, 1, a.g);
  const bool small = small_periodic_grid(a);
  {
    const int c = cy * ncx + cx;
    if (a.cs[c + 1] - a.cs[c] > DSPH_CELL_MAX) over = true;  // the query's own cell
  }
  for (int oy = -1; oy <= 1; ++oy) {
    int yy = cy + oy;
    if (a.d.periodic[1]) {
      yy = yy < 0 ? yy + ncy : (yy >= ncy ? yy - ncy : yy);
    } else if (yy < 0 || yy >= ncy) {
      continue;
    }
    const int row = yy * ncx;
    bool wraps = a.d.periodic[0] && (cx == 0 || cx == ncx - 1);
    const bool seam = small || wraps || yy != cy + oy;
    auto visit = [&](int b, int e) {
      if (e - b > kRangeMax)
        over = true;
      else
        f(b, e, seam);
    };
    if (!wraps) {
      int x0 = cx - 1 < 0 ? 0 : cx - 1;
      int x1 = cx + 1 > ncx - 1 ? ncx - 1 : cx + 1;
      visit(a.cs[row + x0], a.cs[row + x1 + 1]);
    } else {
      for (int ox = -1; ox <= 1; ++ox) {
        int xx = cx + ox;
        xx = xx < 0 ? xx + ncx : (xx >= ncx ? xx - ncx : xx);
        visit(a.cs[row + xx], a.cs[row + xx + 1]);
      }
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_neighbor_count(SearchArgs a, int32_t* __restrict__ counts) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nq) return;
  float2 qi = a.qpos[i];
  if (!a.wrapped) qi = make_float2(wrap_coord(qi.x, a.d, 0), wrap_coord(qi.y, a.d, 1));
  float hi = h_of(a.qh, i, a.h0);
  int c = 0;
  bool over = false;  // counted without the skipped ranges; the host checks the cells
  auto load = [&](int j) { return Cand{a.rpos[j], h_of(a.rh, j, a.h0)}; };
  for_candidates(a, qi, over, [&](int b, int e, bool seam) {
    if (seam)
      scan_range<4>(b, e, load, [&](int, const Cand& r) { c += keep_pair(qi, hi, r.x, r.h, a) ? 1 : 0; });
    else
      scan_range<4>(b, e, load, [&](int, const Cand& r) { c += keep_pair_near(qi, hi, r.x, r.h, a) ? 1 : 0; });
  });
  counts[i] = c;
}

// Fill row i.  CSR (slice == NULL): entries at off[i] + n.  Sliced ELL:
// entries at slice[i/64] + 64 n + i%64, the rest of the slice's columns
// padded with -1.  Writes stay below `cap` and inside the row's own range.
// In the sliced layout the first kFillStage entries of every row are staged
// in LDS as 16-bit offsets j - i (a per-thread column: the staging costs 2
// bytes per entry, so a whole typical row fits) while the candidates are
// scanned, then written column by column, so each store of the wave is one
// contiguous line instead of 64 scattered words.  An entry whose offset does
// not fit in 16 bits, and every entry past kFillStage, is written as it is
// found.
#ifndef DSPH_FILL_STAGE
#define DSPH_FILL_STAGE 48
#endif
constexpr int kFillStage = DSPH_FILL_STAGE;
constexpr int16_t kStagedOut = -32768;  // slot already written to global memory

// With `d16` the 16-bit modular delta (j - i) mod nref of every entry is
// written beside it (rows = reference slots); a delta beyond int16 sets *ovf.
// With `map` (and no d16) the entry is map[j]: the reference slot's index in
// the caller's own order (its sort index), as dsph_remap_neighbors would
// write it afterwards.
__device__ __forceinline__ void put_entry(int32_t* __restrict__ nbr, int16_t* __restrict__ d16,
                                          const int32_t* __restrict__ map, int64_t o, int j, int i, int nref,
                                          bool& bad) {
  nbr[o] = map ? map[j] : j;
  if (d16) {
    int d = j - i;
    const int half = nref / 2;
    if (d > half) d -= nref;
    if (d < -half) d += nref;
    if (d > 32767 || d < -32767) {
      bad = true;
      d = 0;
    }
    d16[o] = (int16_t)d;
  }
}

// Tile slots written by the fill (fixed-width layout, rows = reference
// slots): workgroup b of the fill is tile b of the list; each entry also
// gets its neighbor's slot, j + run_d[u] for the run u holding j.  The tile
// runs come first (they depend on the cells only).
struct FillTiles {
  int32_t* runs;        // run tables, kTileRunStride ints per tile
  int32_t* n;           // particles per tile (set to -1 here when unusable)
  uint16_t* slot;       // per entry, beside nbr
  int cap;
};

// Largest tile whose staged (run << 12 | slot) values stay below the
// sentinel 0xffff.
constexpr int kFillTileMax = 4095;
constexpr int16_t kStagedNoRun = -1;  // slot mode: entry already written to global memory

// Row i of the fill.  kSlot: staged values are (u << 12) | slot, u the run
// of the candidate range in the tile's run table (rt_*), and the slots are
// written beside the entries; otherwise they are j - i.
template <bool kSlot>
__device__ __forceinline__ int64_t fill_row(const SearchArgs& a, int i, float2 qi, float hi, int64_t base,
                                            int64_t step, int64_t len, int64_t staged, int fixed_w,
                                            int16_t (*stage)[kBlock], int32_t* __restrict__ nbr, int64_t cap,
                                            int16_t* __restrict__ d16, const int32_t* __restrict__ map,
                                            const FillTiles& T, const int* rt_s, const int* rt_e, const int* rt_d,
                                            int R, bool& bad, bool& tbad, bool& over) {
  constexpr int16_t kOut = kSlot ? kStagedNoRun : kStagedOut;
  int64_t n = 0;
  // Every candidate writes the staging slot of the next entry (row
  // kFillStage, a sink, once the row is past the staged entries); a kept
  // candidate then advances n, so its value stays and the next candidates
  // write the following slot.  Only an entry past the staged ones, or whose
  // value does not fit the staging (a delta beyond 16 bits; a range in no
  // run), takes the branch to global memory.
  auto take = [&](bool kp, int j, int sb, bool range_ok) {
    const int v = j + sb;
    const bool fits = kSlot ? range_ok : (v > -32768 && v <= 32767);
    const bool st = n < staged;
    stage[st ? n : kFillStage][threadIdx.x] = fits ? (int16_t)v : kOut;
    if (kp && !(st && fits)) {
      const int64_t o = base + n * step;
      if (n < len && o < cap) {
        put_entry(nbr, d16, map, o, j, i, a.nref, bad);
        if (kSlot && fits) T.slot[o] = (uint16_t)(v & 0xfff);
      }
    }
    n += kp ? 1 : 0;
  };
  auto load = [&](int j) { return Cand{a.rpos[j], h_of(a.rh, j, a.h0)}; };
  for_candidates(a, qi, over, [&](int b, int e, bool seam) {
    // staged value of candidate j: j + sb
    int sb = -i;
    bool ok = true;
    if (kSlot && e > b) {
      int u = -1;
      for (int r = 0; r < R; ++r) u = (b >= rt_s[r] && e <= rt_e[r]) ? r : u;
      ok = u >= 0;
      tbad = tbad || !ok;
      sb = ok ? (u << 12) + rt_d[u] : 0;
    }
    if (seam)
      scan_range<4>(b, e, load, [&](int j, const Cand& r) { take(keep_pair(qi, hi, r.x, r.h, a), j, sb, ok); });
    else
      scan_range<4>(b, e, load, [&](int j, const Cand& r) { take(keep_pair_near(qi, hi, r.x, r.h, a), j, sb, ok); });
  });
  for (int64_t t = 0; t < staged; ++t) {
    const int64_t o = base + t * step;
    if (o >= cap) continue;
    if (t < n) {
      const int16_t v = stage[t][threadIdx.x];
      if (v == kOut) continue;
      if (kSlot) {
        const int w = (uint16_t)v, s = w & 0xfff;
        put_entry(nbr, d16, map, o, s - rt_d[w >> 12], i, a.nref, bad);
        T.slot[o] = (uint16_t)s;
      } else {
        put_entry(nbr, d16, map, o, i + v, i, a.nref, bad);
      }
    } else if (fixed_w == 0) {
      nbr[o] = -1;
    }
  }
  return n;
}

// fixed_w > 0: the one-pass fixed-width layout (slice s at s * 64 * fixed_w,
// no count pass): counts[i] = min(entries, fixed_w), *row_ovf set when a row
// has more entries than fixed_w (the list is then incomplete).  kTiles: the
// fixed-width layout with tile slots (T); a tile that does not fit T.cap
// (or kFillTileMax) is filled without them, and one whose rows name a
// candidate range outside its runs gets tile_n = -1.
template <bool kTiles>
__global__ void __launch_bounds__(kBlock) k_neighbor_fill(SearchArgs a, const int32_t* __restrict__ off,
                                                          const int32_t* __restrict__ slice,
                                                          int32_t* __restrict__ nbr, int64_t cap,
                                                          int16_t* __restrict__ d16, int32_t* __restrict__ ovf,
                                                          int fixed_w, int32_t* __restrict__ counts,
                                                          int32_t* __restrict__ row_ovf,
                                                          const int32_t* __restrict__ map, FillTiles T) {
  // one extra row: the sink of the unconditional staging writes
  __shared__ int16_t stage[kFillStage + 1][kBlock];
  __shared__ int rt_s[kTileMaxRuns], rt_e[kTileMaxRuns], rt_d[kTileMaxRuns];
  bool fit = false;
  int R = 0;
  if (kTiles) {
    // every thread reads tile_n before any may set it to -1 (barrier)
    const int S = T.n[blockIdx.x];
    fit = S >= 0 && S <= T.cap && S <= kFillTileMax;
    if (fit) R = load_run_table(T.runs, blockIdx.x, rt_s, rt_e, rt_d);
    __syncthreads();
    if (S > kFillTileMax && S <= T.cap && threadIdx.x == 0) T.n[blockIdx.x] = -1;
  }
  bool bad = false, tbad = false, over = false;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nq) return;
  float2 qi = a.qpos[i];
  if (!a.wrapped) qi = make_float2(wrap_coord(qi.x, a.d, 0), wrap_coord(qi.y, a.d, 1));
  float hi = h_of(a.qh, i, a.h0);
  int64_t base, step, len;
  if (fixed_w > 0) {
    base = (int64_t)(i >> 6) * 64 * fixed_w + (i & 63);
    step = 64;
    len = fixed_w;
  } else if (slice) {
    const int s = i >> 6;
    base = (int64_t)slice[s] + (i & 63);
    step = 64;
    len = ((int64_t)slice[s + 1] - slice[s]) >> 6;
  } else {
    base = off[i];
    step = 1;
    len = (int64_t)off[i + 1] - off[i];
  }
  const int64_t staged = (slice || fixed_w > 0) ? (kFillStage < len ? kFillStage : len) : 0;
  const int64_t n = (kTiles && fit) ? fill_row<true>(a, i, qi, hi, base, step, len, staged, fixed_w, stage, nbr,
                                                     cap, d16, map, T, rt_s, rt_e, rt_d, R, bad, tbad, over)
                                    : fill_row<false>(a, i, qi, hi, base, step, len, staged, fixed_w, stage, nbr,
                                                      cap, d16, map, T, rt_s, rt_e, rt_d, R, bad, tbad, over);
  if (kTiles && tbad) T.n[blockIdx.x] = -1;
  if (fixed_w > 0) {
    // entries past the row's count are never read: no padding
    counts[i] = n < len ? (int32_t)n : (int32_t)len;
    if (n > len) atomicMax(row_ovf, 1);
    if (over) atomicMax(row_ovf, 2);
  } else {
    for (int64_t t = n > staged ? n : staged; t < len; ++t) {
      const int64_t o = base + t * step;
      if (o < cap) nbr[o] = -1;
    }
  }
  if (bad) *ovf = 1;
}

__global__ void __launch_bounds__(kBlock) k_fixed_slices(int nslices, int width, int32_t* __restrict__ slice_off,
                                                         int32_t* __restrict__ row_ovf,
                                                         int32_t* __restrict__ d16_ovf) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) {
    *row_ovf = 0;
    if (d16_ovf) *d16_ovf = 0;
  }
  if (s <= nslices) slice_off[s] = s * 64 * width;
}

__global__ void __launch_bounds__(kBlock) k_zero_i32(int32_t* __restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_fill_flag_clear(int32_t* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

// Longest row of every 64-row slice, times 64 (one wave per slice).
__global__ void __launch_bounds__(kBlock) k_slice_len(const int32_t* __restrict__ counts, int nq,
                                                      int32_t* __restrict__ slen) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int v = i < nq ? counts[i] : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int t = __shfl_xor(v, o, 64);
    v = t > v ? t : v;
  }
  const int nslices = (nq + 63) >> 6;
  // 64 * (longest row) as int32; a row too long for that saturates, which
  // the scan then reports as an overflow of the int32 slice offsets
  if ((threadIdx.x & 63) == 0 && (i >> 6) < nslices)
    slen[i >> 6] = v > (INT32_MAX >> 6) ? INT32_MAX : (v < 0 ? -1 : v * 64);
}

inline GridDev make_grid(const dsph_grid& g) {
  GridDev o;
  o.nc[0] = g.ncell[0];
  o.nc[1] = g.ncell[1];
  o.hcell = g.hcell;
  return o;
}

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

static int check_grid(const dsph_grid* g) {
  if (!g || g->ncell[0] < 1 || g->ncell[1] < 1 || !(g->hcell > 0.0f)) return DSPH_ERR_ARG;
  if ((int64_t)g->ncell[0] * g->ncell[1] >= (int64_t)INT32_MAX) return DSPH_ERR_DIM;
  return DSPH_OK;
}

static int make_search(SearchArgs& a, const dsph_domain* domain, const dsph_grid* grid, const int32_t* cell_start,
                       const float* sorted_pos, const float* sorted_h, int64_t nref, const float* qpos,
                       const float* qh, int64_t nq, int32_t wrapped, float h0, int32_t mode) {
  if (!domain || check_grid(grid) != DSPH_OK) return DSPH_ERR_ARG;
  if (nq < 0 || nref < 0 || nq >= INT32_MAX || nref >= INT32_MAX) return DSPH_ERR_ARG;
  if (mode < 0 || mode > 2) return DSPH_ERR_ARG;
  if (nq > 0 && (!qpos || !cell_start)) return DSPH_ERR_ARG;
  if (nref > 0 && !sorted_pos) return DSPH_ERR_ARG;
  // Supports: a NULL array means the uniform h0, which must then be > 0.
  if (((mode != DSPH_MODE_GATHER && !sorted_h) || (mode != DSPH_MODE_SCATTER && !qh)) && !(h0 > 0.0f))
    return DSPH_ERR_ARG;
  a.d = make_dom(*domain);
  a.g = make_grid(*grid);
  a.cs = cell_start;
  a.rpos = reinterpret_cast<const float2*>(sorted_pos);
  a.rh = sorted_h;
  a.qpos = reinterpret_cast<const float2*>(qpos);
  a.qh = qh;
  a.nref = (int)nref;
  a.h0 = h0;
  a.mode = mode;
  a.wrapped = wrapped != 0;
  a.nq = (int)nq;
  return DSPH_OK;
}

}  // namespace dsph

using namespace dsph;

extern "C" int dsph_grid_dims(const dsph_domain* domain, float hsearch, dsph_grid* grid) {
  if (!domain || !grid || !(hsearch > 0.0f)) return DSPH_ERR_ARG;
  // Cells are kept 0.1% wider than the search radius so that float32
  // rounding of cell coordinates can never separate two particles closer
  // than the radius by two cells.
  const double hs = (double)hsearch * 1.001;
  double hc = 0.0;
  int nc[2];
  for (int k = 0; k < 2; ++k) {
    double L = (double)domain->max[k] - (double)domain->min[k];
    if (!(L > 0.0)) return DSPH_ERR_ARG;
    double n = std::floor(L / hs);
    if (n < 1.0) n = 1.0;
    double side = L / n;
    if (side > hc) hc = side;
  }
  // One square side for both axes: the larger one (reference compute_h).
  float hcell = (float)hc;
  if ((double)hcell < hs) hcell = nextafterf(hcell, INFINITY);
  for (int k = 0; k < 2; ++k) {
    double L = (double)domain->max[k] - (double)domain->min[k];
    double n;
    if (domain->periodic[k]) {
      // A periodic axis is tiled by floor(L/hcell) cells, each >= hcell wide
      // (the last cell absorbs the remainder), so the wrap keeps adjacency.
      // (+1e-4: hcell may have been rounded up by one float ulp; the last
      // cell is then at least (1-1e-4) hcell > the search radius wide.)
      n = std::floor(L / (double)hcell + 1e-4);
      if (n < 3.0) return DSPH_ERR_DIM;
    } else {
      n = std::ceil(L / (double)hcell);
      if (n < 1.0) n = 1.0;
    }
    if (n > 1e8) return DSPH_ERR_DIM;
    nc[k] = (int)n;
  }
  grid->ncell[0] = nc[0];
  grid->ncell[1] = nc[1];
  grid->hcell = hcell;
  grid->pad_ = 0;
  return DSPH_OK;
}

extern "C" size_t dsph_grid_build_workspace(int64_t n, const dsph_grid* grid) {
  if (!grid) return 0;
  int64_t ncell = (int64_t)grid->ncell[0] * grid->ncell[1];
  return align256((size_t)n * 4) * 2 + align256((size_t)ncell * 4) + scan_ws(ncell);
}

extern "C" int dsph_grid_build(const dsph_domain* domain, const dsph_grid* grid, const float* pos, const float* h,
                               int64_t n, int32_t* cell_start, int32_t* sort_index, int32_t* rank,
                               float* sorted_pos, float* sorted_h, void* ws, size_t ws_bytes, void* stream) {
  if (!domain || check_grid(grid) != DSPH_OK || n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if (!cell_start || (n > 0 && (!pos || !sort_index || !sorted_pos))) return DSPH_ERR_ARG;
  if (sorted_h && !h) return DSPH_ERR_ARG;
  // A periodic axis is only valid when the cells tile it (L/hcell >= nc).
  for (int k = 0; k < 2; ++k) {
    if (domain->periodic[k] && grid->ncell[k] < 3) return DSPH_ERR_DIM;
  }
  const int64_t ncell = (int64_t)grid->ncell[0] * grid->ncell[1];
  if (!ws || ws_bytes < dsph_grid_build_workspace(n, grid)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  char* w = static_cast<char*>(ws);
  int32_t* keys = reinterpret_cast<int32_t*>(w);
  w += align256((size_t)n * 4);
  int32_t* slot = reinterpret_cast<int32_t*>(w);
  w += align256((size_t)n * 4);
  int32_t* count = reinterpret_cast<int32_t*>(w);
  w += align256((size_t)ncell * 4);
  void* scan_w = w;
  size_t scan_b = scan_ws(ncell);

  Dom d = make_dom(*domain);
  GridDev g = make_grid(*grid);
  // a fill kernel rather than a memset node, so the call can be captured
  k_zero_i32<<<grid_for(ncell > 0 ? ncell : 1), kBlock, 0, st>>>(count, ncell);
  if (n > 0) {
    k_cell_keys<<<grid_for(n), kBlock, 0, st>>>(reinterpret_cast<const float2*>(pos), (int)n, d, g, keys, slot, count);
    DSPH_LAUNCH_CHECK();
  }
  int s = scan_exclusive(count, ncell, cell_start, scan_w, scan_b, st);
  if (s != DSPH_OK) return s;
  if (n > 0) {
    k_cell_scatter<<<grid_for(n), kBlock, 0, st>>>((int)n, keys, slot, cell_start, sort_index);
    k_cell_sort<<<grid_for(ncell), kBlock, 0, st>>>((int)ncell, cell_start, sort_index);
    k_sorted_gather<<<grid_for(n), kBlock, 0, st>>>((int)n, d, sort_index, reinterpret_cast<const float2*>(pos), h,
                                                    reinterpret_cast<float2*>(sorted_pos), sorted_h, rank);
    DSPH_LAUNCH_CHECK();
  }
  return DSPH_OK;
}

extern "C" int dsph_neighbor_count(const dsph_domain* domain, const dsph_grid* grid, const int32_t* cell_start,
                                   const float* sorted_pos, const float* sorted_h, int64_t nref, const float* qpos,
                                   const float* qh, int64_t nq, int32_t qpos_wrapped, float h_uniform, int32_t mode,
                                   int32_t* counts, void* stream) {
  SearchArgs a;
  int s = make_search(a, domain, grid, cell_start, sorted_pos, sorted_h, nref, qpos, qh, nq, qpos_wrapped,
                      h_uniform, mode);
  if (s != DSPH_OK) return s;
  if (nq > 0 && !counts) return DSPH_ERR_ARG;
  if (nq == 0) return DSPH_OK;
  k_neighbor_count<<<grid_for(nq), kBlock, 0, as_stream(stream)>>>(a, counts);
  return launch_status();
}

extern "C" int dsph_neighbor_fill(const dsph_domain* domain, const dsph_grid* grid, const int32_t* cell_start,
                                  const float* sorted_pos, const float* sorted_h, int64_t nref, const float* qpos,
                                  const float* qh, int64_t nq, int32_t qpos_wrapped, float h_uniform, int32_t mode,
                                  const int32_t* offsets, const int32_t* slice_off, int64_t layout_total,
                                  int32_t* nbr, int64_t capacity, void* stream) {
  SearchArgs a;
  int s = make_search(a, domain, grid, cell_start, sorted_pos, sorted_h, nref, qpos, qh, nq, qpos_wrapped,
                      h_uniform, mode);
  if (s != DSPH_OK) return s;
  if (nq == 0) return DSPH_OK;
  if ((!offsets && !slice_off) || (capacity > 0 && !nbr) || capacity < 0 || layout_total < 0) return DSPH_ERR_ARG;
  if (layout_total > capacity) return DSPH_ERR_CAPACITY;
  k_neighbor_fill<false><<<grid_for(nq), kBlock, 0, as_stream(stream)>>>(
      a, offsets, slice_off, nbr, capacity, nullptr, nullptr, 0, nullptr, nullptr, nullptr, FillTiles{});
  return launch_status();
}

extern "C" int dsph_neighbor_fill16(const dsph_domain* domain, const dsph_grid* grid, const int32_t* cell_start,
                                    const float* sorted_pos, const float* sorted_h, int64_t nref, const float* qpos,
                                    const float* qh, int64_t nq, int32_t qpos_wrapped, float h_uniform, int32_t mode,
                                    const int32_t* offsets, const int32_t* slice_off, int64_t layout_total,
                                    int32_t* nbr, int16_t* nbr16, int64_t capacity, int32_t* overflow,
                                    void* stream) {
  SearchArgs a;
  int s = make_search(a, domain, grid, cell_start, sorted_pos, sorted_h, nref, qpos, qh, nq, qpos_wrapped,
                      h_uniform, mode);
  if (s != DSPH_OK) return s;
  if (!overflow || (capacity > 0 && !nbr16)) return DSPH_ERR_ARG;
  // the deltas are j - i modulo n: row i must be reference slot i
  if (nq != nref || !qpos_wrapped) return DSPH_ERR_ARG;
  hipStream_t st = as_stream(stream);
  k_fill_flag_clear<<<1, 64, 0, st>>>(overflow);
  DSPH_LAUNCH_CHECK();
  if (nq == 0) return DSPH_OK;
  if ((!offsets && !slice_off) || (capacity > 0 && !nbr) || capacity < 0 || layout_total < 0) return DSPH_ERR_ARG;
  if (layout_total > capacity) return DSPH_ERR_CAPACITY;
  k_neighbor_fill<false><<<grid_for(nq), kBlock, 0, st>>>(a, offsets, slice_off, nbr, capacity, nbr16, overflow, 0,
                                                           nullptr, nullptr, nullptr, FillTiles{});
  return launch_status();
}

namespace dsph {
static int fill_fixed(const dsph_domain* domain, const dsph_grid* grid, const int32_t* cell_start,
                      const float* sorted_pos, const float* sorted_h, int64_t nref, const float* qpos, const float* qh,
                      int64_t nq, int32_t qpos_wrapped, float h_uniform, int32_t mode, int32_t width, int32_t* counts,
                      int32_t* slice_off, int32_t* nbr, int64_t capacity, int16_t* nbr16, int32_t* delta_overflow,
                      int32_t* row_overflow, const int32_t* map, const FillTiles* tiles, void* stream) {
  SearchArgs a;
  int s = make_search(a, domain, grid, cell_start, sorted_pos, sorted_h, nref, qpos, qh, nq, qpos_wrapped,
                      h_uniform, mode);
  if (s != DSPH_OK) return s;
  if (width < 1 || width > 4096 || !row_overflow || !slice_off || (nbr16 && !delta_overflow)) return DSPH_ERR_ARG;
  // 16-bit deltas are j - i modulo n: row i must be reference slot i
  if (nbr16 && (nq != nref || !qpos_wrapped || map)) return DSPH_ERR_ARG;
  const int64_t nslices = (nq + 63) / 64;
  if (nslices * 64 * (int64_t)width >= INT32_MAX) return DSPH_ERR_ARG;
  if (capacity < nslices * 64 * (int64_t)width) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  k_fixed_slices<<<grid_for(nslices + 1), kBlock, 0, st>>>((int)nslices, width, slice_off, row_overflow,
                                                          delta_overflow);
  DSPH_LAUNCH_CHECK();
  if (nq == 0) return DSPH_OK;
  if (!counts || !nbr) return DSPH_ERR_ARG;
  if (tiles) {
    int s2 = launch_tile_runs(nq, grid, domain, cell_start, tiles->runs, tiles->n, st);
    if (s2 != DSPH_OK) return s2;
    k_neighbor_fill<true><<<grid_for(nq), kBlock, 0, st>>>(a, nullptr, nullptr, nbr, capacity, nbr16, delta_overflow,
                                                          width, counts, row_overflow, map, *tiles);
  } else {
    k_neighbor_fill<false><<<grid_for(nq), kBlock, 0, st>>>(a, nullptr, nullptr, nbr, capacity, nbr16, delta_overflow,
                                                           width, counts, row_overflow, map, FillTiles{});
  }
  return launch_status();
}
}  // namespace dsph

extern "C" int dsph_neighbor_fill_fixed(const dsph_domain* domain, const dsph_grid* grid, const int32_t* cell_start,
                                        const float* sorted_pos, const float* sorted_h, int64_t nref,
                                        const float* qpos, const float* qh, int64_t nq, int32_t qpos_wrapped,
                                        float h_uniform, int32_t mode, int32_t width, int32_t* counts,
                                        int32_t* slice_off, int32_t* nbr, int64_t capacity, int16_t* nbr16,
                                        int32_t* delta_overflow, int32_t* row_overflow, void* stream) {
  return fill_fixed(domain, grid, cell_start, sorted_pos, sorted_h, nref, qpos, qh, nq, qpos_wrapped, h_uniform, mode,
                    width, counts, slice_off, nbr, capacity, nbr16, delta_overflow, row_overflow, nullptr, nullptr,
                    stream);
}

extern "C" int dsph_neighbor_fill_fixed_map(const dsph_domain* domain, const dsph_grid* grid,
                                            const int32_t* cell_start, const float* sorted_pos,
                                            const float* sorted_h, int64_t nref, const float* qpos, const float* qh,
                                            int64_t nq, int32_t qpos_wrapped, float h_uniform, int32_t mode,
                                            int32_t width, int32_t* counts, int32_t* slice_off, int32_t* nbr,
                                            int64_t capacity, const int32_t* map, int32_t* row_overflow,
                                            void* stream) {
  if (nref > 0 && !map) return DSPH_ERR_ARG;
  return fill_fixed(domain, grid, cell_start, sorted_pos, sorted_h, nref, qpos, qh, nq, qpos_wrapped, h_uniform, mode,
                    width, counts, slice_off, nbr, capacity, nullptr, nullptr, row_overflow, map, nullptr, stream);
}

extern "C" int dsph_neighbor_fill_fixed_tiles(const dsph_domain* domain, const dsph_grid* grid,
                                              const int32_t* cell_start, const float* sorted_pos,
                                              const float* sorted_h, int64_t nref, const float* qpos,
                                              const float* qh, int64_t nq, int32_t qpos_wrapped, float h_uniform,
                                              int32_t mode, int32_t width, int32_t* counts, int32_t* slice_off,
                                              int32_t* nbr, int64_t capacity, int16_t* nbr16,
                                              int32_t* delta_overflow, int32_t tile_cap, uint16_t* slot,
                                              int64_t slot_capacity, int32_t* tile_runs, int32_t* tile_n,
                                              int32_t* row_overflow, void* stream) {
  // the tiles of a list whose rows are the reference slots
  if (nq != nref || !qpos_wrapped) return DSPH_ERR_ARG;
  if (tile_cap < 1 || tile_cap > DSPH_TILE_MAX_CAP || slot_capacity < 0) return DSPH_ERR_ARG;
  if (nq > 0 && (!slot || !tile_runs || !tile_n)) return DSPH_ERR_ARG;
  if (width >= 1 && slot_capacity < ((nq + 63) / 64) * 64 * (int64_t)width) return DSPH_ERR_CAPACITY;
  FillTiles t{tile_runs, tile_n, slot, tile_cap};
  return fill_fixed(domain, grid, cell_start, sorted_pos, sorted_h, nref, qpos, qh, nq, qpos_wrapped, h_uniform, mode,
                    width, counts, slice_off, nbr, capacity, nbr16, delta_overflow, row_overflow, nullptr, &t,
                    stream);
}

extern "C" size_t dsph_sell_workspace(int64_t nq) {
  int64_t ns = (nq + 63) / 64;
  return align256((size_t)(ns > 0 ? ns : 1) * 4) + scan_ws(ns);
}

extern "C" int dsph_sell_layout(const int32_t* counts, int64_t nq, int32_t* slice_off, int64_t* total_host,
                                void* ws, size_t ws_bytes, void* stream) {
  if (nq < 0 || nq >= INT32_MAX || !slice_off || (nq > 0 && !counts)) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_sell_workspace(nq)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  const int64_t ns = (nq + 63) / 64;
  int32_t* slen = static_cast<int32_t*>(ws);
  char* scan_w = static_cast<char*>(ws) + align256((size_t)(ns > 0 ? ns : 1) * 4);
  if (nq > 0) {
    k_slice_len<<<grid_for(nq), kBlock, 0, st>>>(counts, (int)nq, slen);
    DSPH_LAUNCH_CHECK();
  }
  int s = scan_exclusive(slen, ns, slice_off, scan_w, scan_ws(ns), st);
  if (s != DSPH_OK) return s;
  if (total_host) return scan_read_total(scan_w, total_host, st);
  return DSPH_OK;
}

extern "C" int dsph_wrap_positions(const dsph_domain* domain, const float* pos, int64_t n, float* out, void* stream) {
  if (!domain || n < 0 || n >= INT32_MAX) return DSPH_ERR_ARG;
  if (n == 0) return DSPH_OK;
  if (!pos || !out) return DSPH_ERR_ARG;
  k_wrap<<<grid_for(n), kBlock, 0, as_stream(stream)>>>(reinterpret_cast<const float2*>(pos), (int)n,
                                                        make_dom(*domain), reinterpret_cast<float2*>(out));
  return launch_status();
}

namespace dsph {
__global__ void __launch_bounds__(kBlock) k_remap(int32_t* __restrict__ nbr, int64_t entries,
                                                  const int32_t* __restrict__ map, int nmap) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= entries) return;
  const int32_t j = nbr[e];
  nbr[e] = (j >= 0 && j < nmap) ? map[j] : -1;
}
}  // namespace dsph

extern "C" int dsph_remap_neighbors(int32_t* nbr, int64_t entries, const int32_t* map, int64_t nmap, void* stream) {
  if (entries < 0 || nmap < 0 || nmap >= INT32_MAX || (entries > 0 && (!nbr || !map))) return DSPH_ERR_ARG;
  if (entries == 0) return DSPH_OK;
  const int64_t nb = (entries + kBlock - 1) / kBlock;
  if (nb >= INT32_MAX) return DSPH_ERR_ARG;
  k_remap<<<(int)nb, kBlock, 0, as_stream(stream)>>>(nbr, entries, map, (int)nmap);
  return launch_status();
}
