// Exclusive prefix sums over int32 with wave64 shuffles (three phases:
// block totals, one-block scan of the totals, block-local scan + offset).
#include "common.h"

namespace dsph {

constexpr int kScanPerThread = 8;
constexpr int kScanTile = kBlock * kScanPerThread;  // 2048 elements per block

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Inclusive scan over a 256-thread block; returns the block total in *total.
__device__ __forceinline__ int block_incl_scan(int v, int* total) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int s = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = s;
  __syncthreads();
  int add = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    int x = wsum[w];
    if (w < wid) add += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return s + add;
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                        int32_t* __restrict__ block_sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    int64_t idx = base + (int64_t)k * kBlock + threadIdx.x;
    if (idx < n) v += in[idx];
  }
  int tot;
  block_incl_scan(v, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Exclusive scan of block_sums[0..nb) in place with a single 256-thread block.
__global__ void __launch_bounds__(kBlock) k_scan_partials(int32_t* __restrict__ sums, int nb) {
  int carry = 0;
  for (int base = 0; base < nb; base += kBlock) {
    int idx = base + threadIdx.x;
    int v = idx < nb ? sums[idx] : 0;
    int tot;
    int incl = block_incl_scan(v, &tot);
    if (idx < nb) sums[idx] = carry + incl - v;
    carry += tot;
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_final(const int32_t* __restrict__ in, int64_t n,
                                                       const int32_t* __restrict__ block_offsets,
                                                       int32_t* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
  int vals[kScanPerThread];
  int local = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    int64_t idx = base + k;
    vals[k] = idx < n ? in[idx] : 0;
    local += vals[k];
  }
  int tot;
  int incl = block_incl_scan(local, &tot);
  int run = block_offsets[blockIdx.x] + incl - local;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    int64_t idx = base + k;
    if (idx < n) out[idx] = run;
    run += vals[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) out[n] = block_offsets[blockIdx.x] + tot;
}

__global__ void k_scan_empty(int32_t* out) { out[0] = 0; }

int scan_exclusive(const int32_t* in, int64_t n, int32_t* out, void* ws, size_t ws_bytes, hipStream_t st) {
  if (n < 0 || !out || (n > 0 && !in)) return DSPH_ERR_ARG;
  if (n == 0) {
    k_scan_empty<<<1, 1, 0, st>>>(out);
    return launch_status();
  }
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (ws_bytes < (size_t)nb * sizeof(int32_t) || !ws) return DSPH_ERR_CAPACITY;
  int32_t* sums = static_cast<int32_t*>(ws);
  k_scan_reduce<<<(int)nb, kBlock, 0, st>>>(in, n, sums);
  k_scan_partials<<<1, kBlock, 0, st>>>(sums, (int)nb);
  k_scan_final<<<(int)nb, kBlock, 0, st>>>(in, n, sums, out);
  return launch_status();
}

size_t scan_ws(int64_t n) {
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb < 1) nb = 1;
  return (size_t)((nb * sizeof(int32_t) + 255) / 256 * 256);
}

}  // namespace dsph

extern "C" size_t dsph_scan_workspace(int64_t n) { return dsph::scan_ws(n); }

extern "C" int dsph_scan_exclusive(const int32_t* in, int64_t n, int32_t* out, int64_t* total_host,
                                   void* ws, size_t ws_bytes, void* stream) {
  hipStream_t st = dsph::as_stream(stream);
  int s = dsph::scan_exclusive(in, n, out, ws, ws_bytes, st);
  if (s != DSPH_OK) return s;
  if (total_host) {
    int32_t t = 0;
    DSPH_CHECK_HIP(hipMemcpyAsync(&t, out + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DSPH_CHECK_HIP(hipStreamSynchronize(st));
    *total_host = t;
  }
  return DSPH_OK;
}
