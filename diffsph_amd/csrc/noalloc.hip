// An allocation-free variant of the DFSPH pressure solve for callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  Its scratch (solve
// control block, per-block float64 partials, source term and the two 16-byte
// record arrays of the sweeps) lives in one file-scope device array of
// kNoAllocScratchBytes that every GPU running the library holds until the
// process ends.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU (two streams, two host threads) race.  Everything the
// solve reads from the scratch is written by the same call first (the
// control block by a kernel), so the contents a previous call left do not
// matter.
#include <mutex>

#include "common.h"

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;  // 64 MiB

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current (host code never uses the array's name as an
// address).
int scratch_ptr(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return DSPH_ERR_LAUNCH;
  if (dev < 0 || dev >= kMaxDevices) return DSPH_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return DSPH_ERR_LAUNCH;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return DSPH_OK;
}

}  // namespace

extern "C" size_t dsph_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int64_t dsph_dfsph_solve_noalloc_max_particles(void) {
  // the largest n whose solve workspace fits the scratch (the workspace
  // grows monotonically with n)
  int64_t lo = 0, hi = int64_t(1) << 31;
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (dsph_dfsph_workspace(mid) <= kNoAllocScratchBytes)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

extern "C" int dsph_dfsph_solve_noalloc(const dsph_pairs* pairs, const dsph_dfsph_params* prm, const float* kgrad,
                                        const float* vel, const float* adv, const float* rho, const float* vol,
                                        const float* areas, const float* alpha, const float* src0, float* p,
                                        float* accel, float* residual, dsph_solve_stats* stats, void* stream) {
  if (!pairs) return DSPH_ERR_ARG;
  if (dsph_dfsph_workspace(pairs->nq) > kNoAllocScratchBytes) return DSPH_ERR_CAPACITY;
  void* ws = nullptr;
  const int s = scratch_ptr(&ws);
  if (s != DSPH_OK) return s;
  return dsph_dfsph_solve(pairs, prm, kgrad, vel, adv, rho, vol, areas, alpha, src0, p, accel, residual, stats, ws,
                          kNoAllocScratchBytes, stream);
}
