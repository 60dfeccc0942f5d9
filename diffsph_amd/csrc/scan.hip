// Exclusive prefix sums over int32 with wave64 shuffles (three phases:
// block totals, one-block scan of the totals, block-local scan + offset).
// Totals are carried in int64 and an overflow of the int32 offsets is
// flagged in the workspace header.
#include "common.h"

namespace dsph {

constexpr int kScanPerThread = 8;
constexpr int kScanTile = kBlock * kScanPerThread;  // 2048 elements per block

size_t scan_ws(int64_t n);

// Inclusive scan over a 256-thread block; returns the block total in *total.
template <class T>
__device__ __forceinline__ T block_incl_scan(T v, T* total) {
  __shared__ T wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  T s = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T t = __shfl_up(s, o, 64);
    if (lane >= o) s += t;
  }
  if (lane == 63) wsum[wid] = s;
  __syncthreads();
  T add = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    T x = wsum[w];
    if (w < wid) add += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return s + add;
}

// Workspace: a 16-byte header {int64 total, int32 overflow} and one int64
// sum per 2048-element tile.  Tile sums and the running offsets are int64,
// so a total beyond INT32_MAX is detected instead of wrapping.
struct ScanHeader {
  int64_t total;
  int32_t overflow;
  int32_t pad_;
};

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                        int64_t* __restrict__ block_sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t v = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    int64_t idx = base + (int64_t)k * kBlock + threadIdx.x;
    if (idx < n) v += in[idx];
  }
  int64_t tot;
  block_incl_scan<int64_t>(v, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Exclusive scan of block_sums[0..nb) in place with a single 256-thread
// block; writes the total and the int32 overflow flag into the header.
__global__ void __launch_bounds__(kBlock) k_scan_partials(int64_t* __restrict__ sums, int nb,
                                                          ScanHeader* __restrict__ hdr) {
  int64_t carry = 0;
  int bad = 0;
  for (int base = 0; base < nb; base += kBlock) {
    int idx = base + threadIdx.x;
    int64_t v = idx < nb ? sums[idx] : 0;
    if (v < 0) bad = 1;
    int64_t tot;
    int64_t incl = block_incl_scan<int64_t>(v, &tot);
    if (idx < nb) sums[idx] = carry + incl - v;
    carry += tot;
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) {
    hdr->total = carry;
    hdr->overflow = (bad || carry > (int64_t)INT32_MAX) ? 1 : 0;
  }
}

// Tile-local scan plus the tile's offset.  When the header flags an
// overflow the int32 output is not meaningful (the caller reports the
// error); the arithmetic is still well defined (unsigned wrap).
__global__ void __launch_bounds__(kBlock) k_scan_final(const int32_t* __restrict__ in, int64_t n,
                                                       const int64_t* __restrict__ block_offsets,
                                                       int32_t* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
  int64_t vals[kScanPerThread];
  int64_t local = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    int64_t idx = base + k;
    vals[k] = idx < n ? in[idx] : 0;
    local += vals[k];
  }
  int64_t tot;
  int64_t incl = block_incl_scan<int64_t>(local, &tot);
  int64_t run = block_offsets[blockIdx.x] + incl - local;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    int64_t idx = base + k;
    if (idx < n) out[idx] = (int32_t)(uint32_t)(uint64_t)run;
    run += vals[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1)
    out[n] = (int32_t)(uint32_t)(uint64_t)(block_offsets[blockIdx.x] + tot);
}

__global__ void k_scan_empty(int32_t* out, ScanHeader* hdr) {
  out[0] = 0;
  hdr->total = 0;
  hdr->overflow = 0;
}

int scan_exclusive(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st) {
  if (n < 0 || !out || (n > 0 && !in)) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < scan_ws(n)) return DSPH_ERR_CAPACITY;
  ScanHeader* hdr = static_cast<ScanHeader*>(ws);
  if (n == 0) {
    k_scan_empty<<<1, 1, 0, st>>>(out, hdr);
    return launch_status();
  }
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb >= INT32_MAX) return DSPH_ERR_DIM;
  int64_t* sums = reinterpret_cast<int64_t*>(static_cast<char*>(ws) + sizeof(ScanHeader));
  k_scan_reduce<<<(int)nb, kBlock, 0, st>>>(in, n, sums);
  k_scan_partials<<<1, kBlock, 0, st>>>(sums, (int)nb, hdr);
  k_scan_final<<<(int)nb, kBlock, 0, st>>>(in, n, sums, out);
  return launch_status();
}

size_t scan_ws(int64_t n) {
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb < 1) nb = 1;
  return (size_t)((sizeof(ScanHeader) + nb * sizeof(int64_t) + 255) / 256 * 256);
}

// Blocking read of the scan's int64 total; DSPH_ERR_DIM when it does not
// fit the int32 offsets.
int scan_read_total(const void* ws, int64_t* total_host, hipStream_t st) {
  ScanHeader h;
  DSPH_CHECK_HIP(hipMemcpyAsync(&h, ws, sizeof(ScanHeader), hipMemcpyDeviceToHost, st));
  DSPH_CHECK_HIP(hipStreamSynchronize(st));
  if (h.overflow) return DSPH_ERR_DIM;
  *total_host = h.total;
  return DSPH_OK;
}

}  // namespace dsph

extern "C" size_t dsph_scan_workspace(int64_t n) { return dsph::scan_ws(n); }

extern "C" int dsph_scan_exclusive(const int32_t* in, int64_t n, int32_t* out, int64_t* total_host,
                                   void* ws, size_t ws_bytes, void* stream) {
  hipStream_t st = dsph::as_stream(stream);
  int s = dsph::scan_exclusive(in, n, out, ws, ws_bytes, st);
  if (s != DSPH_OK) return s;
  if (total_host) return dsph::scan_read_total(ws, total_host, st);
  return DSPH_OK;
}
