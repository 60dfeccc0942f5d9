// Drop-ins for the reference's compiled pair ops that take per-pair arrays
// (viscosityKernel, mDBC, computeNeighborhood), the Verlet-list pair filter,
// the Verlet reuse test and the compact-hash-map neighbor search
// (countNeighbors / buildNeighborList).
//
// Row work: one thread per row where a row is summed (viscosity, mDBC,
// hash-map search: the reference's order of summation and of output is the
// row order); one wave64 per row where a row is filtered (computeNeighborhood
// and the Verlet filter: 64 consecutive entries per step, coalesced, the
// kept entries compacted in order with a ballot prefix).
#include <cstring>

#include "common.h"
#include "mdbc_op.h"

namespace dsph {

// ---------------------------------------------------------------------------
// viscosityKernel (partiBench/viscosityKernel.h:98-145)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_op_viscosity(
    const float* __restrict__ vi, const float* __restrict__ vj, const float* __restrict__ mj,
    const float* __restrict__ rhoi, const float* __restrict__ rhoj, const float* __restrict__ hi_arr,
    const int64_t* __restrict__ jj, const float* __restrict__ hij, const float* __restrict__ grad,
    const float* __restrict__ dist, const float* __restrict__ vec, int n, int64_t nref, int64_t np,
    const int32_t* __restrict__ nn, const int32_t* __restrict__ off, float cs, float rho0, float alpha, float eps,
    int pi_switch, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float vx = vi[2 * (int64_t)i], vy = vi[2 * (int64_t)i + 1];
  const float rho_i = rhoi[i];
  float rx = 0.f, ry = 0.f;
  int64_t b, e;
  seg_bounds(off, nn, i, np, b, e);
  for (int64_t k = b; k < e; ++k) {
    const int64_t j = jj[k];
    if (j < 0 || j >= nref) continue;
    const float h = hij[k];
    const float dvx = vx - vj[2 * j], dvy = vy - vj[2 * j + 1];
    const float r = dist[k] * h;
    const float vr = dvx * vec[2 * k] + dvy * vec[2 * k + 1];
    float pi = vr / (r + eps + h * h);
    if (pi_switch) pi = vr < 0.f ? pi : 0.f;
    const float V = mj[j] / (rho_i + rhoj[j]);
    rx += V * pi * grad[2 * k];
    ry += V * pi * grad[2 * k + 1];
  }
  const float factor = alpha * hi_arr[i] * cs * rho0 / rho_i;
  out[2 * (int64_t)i] = rx * factor;
  out[2 * (int64_t)i + 1] = ry * factor;
}

// ---------------------------------------------------------------------------
// mDBC on per-pair arrays (v2/cppSrc/mDBC.h:129-244)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_op_mdbc(
    const float* __restrict__ mj, const float* __restrict__ rhoj, int64_t nref, const float* __restrict__ xb,
    const float* __restrict__ xg, int nb, const int64_t* __restrict__ jj, const float* __restrict__ W,
    const float* __restrict__ grad, const float* __restrict__ dist, const float* __restrict__ vec,
    const float* __restrict__ hij, int64_t np, const int32_t* __restrict__ nn, const int32_t* __restrict__ off,
    float rho0, float* __restrict__ rho_b, float* __restrict__ shepard) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  MdbcOpAcc acc;
  int64_t b, e;
  seg_bounds(off, nn, i, np, b, e);
  for (int64_t k = b; k < e; ++k) {
    const int64_t j = jj[k];
    if (j < 0 || j >= nref) continue;
    const float rh = -dist[k] * hij[k];
    acc.add(mj[j], rhoj[j], W[k], grad[2 * k], grad[2 * k + 1], rh * vec[2 * k], rh * vec[2 * k + 1]);
  }
  const int64_t i2 = 2 * (int64_t)i;
  float shep;
  rho_b[i] = acc.finish((int)(e - b), xb[i2] - xg[i2], xb[i2 + 1] - xg[i2 + 1], rho0, &shep);
  shepard[i] = shep;
}

// ---------------------------------------------------------------------------
// Row filters: count, then fill in row order.
// ---------------------------------------------------------------------------
struct CsrIn {
  const int64_t* jj;
  const int32_t* nn;
  const int32_t* off;
  int64_t np;
  int64_t nref;
  int n;
};

struct PairOut {
  int64_t* i;
  int64_t* j;
  float* rij;
  float* xij;
  float* hij;
  float* w;
  float* gw;
  int64_t cap;
};

// C++ computeNeighborhood geometry (v2/cppSrc/computeNeighborhood.h:110-113,
// 115-187): per-axis offset in double, |d|^2 <= h_j^2.
struct GeoCpp {
  const float* pi;
  const float* hi;
  const float* pj;
  const float* hj;
  const float* dmin;
  const float* dmax;
  const uint8_t* per;

  struct Q {
    float x, y, h, Lx, Ly;
    bool px, py;
  };
  struct P {
    double dx, dy, sum;
    float hj;
  };

  // ((p - q + L/2) - floor((p - q + L/2) / L) * L) - L/2, p - q in float
  __device__ static double modulo_op(float p, float q, float L) {
    const double a = (double)(p - q) + (double)L / 2.0;
    return (a - floor(a / (double)L) * (double)L) - (double)L / 2.0;
  }
  __device__ Q query(int i) const {
    Q q;
    q.x = pi[2 * (int64_t)i];
    q.y = pi[2 * (int64_t)i + 1];
    q.h = hi[i];
    q.Lx = dmax[0] - dmin[0];
    q.Ly = dmax[1] - dmin[1];
    q.px = per[0] != 0;
    q.py = per[1] != 0;
    return q;
  }
  __device__ bool eval(const Q& q, int64_t j, P& p) const {
    const float xj = pj[2 * j], yj = pj[2 * j + 1];
    p.hj = hj[j];
    p.dx = q.px ? modulo_op(q.x, xj, q.Lx) : (double)(q.x - xj);
    p.dy = q.py ? modulo_op(q.y, yj, q.Ly) : (double)(q.y - yj);
    p.sum = p.dx * p.dx + p.dy * p.dy;
    return p.sum <= (double)(p.hj * p.hj);
  }
  __device__ void write(const Q& q, int row, int64_t j, const P& p, const PairOut& o, int64_t k) const {
    const double d = sqrt(p.sum);
    const double den = d + (double)(1e-6f * q.h);
    o.i[k] = row;
    o.j[k] = j;
    o.rij[k] = (float)(d / (double)p.hj);
    o.xij[2 * k] = (float)(p.dx / den);
    o.xij[2 * k + 1] = (float)(p.dy / den);
    o.hij[k] = p.hj;
  }
};

// Python computeNeighborhood geometry (v2/modules/neighborhood.py:433-460)
// plus the Wendland C2 kernel values (:585-589), all float32.
struct GeoPy {
  const float* pi;
  const float* hi;
  const float* pj;
  const float* hj;
  Dom dom;
  int mode;

  struct Q {
    float2 x;
    float h;
  };
  using P = Geom;

  __device__ Q query(int i) const {
    Q q;
    q.x = make_float2(pi[2 * (int64_t)i], pi[2 * (int64_t)i + 1]);
    q.h = hi ? hi[i] : 0.f;
    return q;
  }
  __device__ bool eval(const Q& q, int64_t j, P& p) const {
    const float h = pick_h(mode, q.h, hj ? hj[j] : 0.f);
    p = pair_geom_exact(q.x, make_float2(pj[2 * j], pj[2 * j + 1]), h, dom);
    return p.q <= 1.0f;
  }
  __device__ void write(const Q&, int row, int64_t j, const P& p, const PairOut& o, int64_t k) const {
    o.i[k] = row;
    o.j[k] = j;
    o.rij[k] = p.q;
    o.xij[2 * k] = p.ex;
    o.xij[2 * k + 1] = p.ey;
    o.hij[k] = p.h;
    if (o.w) o.w[k] = p.w;
    if (o.gw) {
      o.gw[2 * k] = p.gx;
      o.gw[2 * k + 1] = p.gy;
    }
  }
};

constexpr int kWavesPerBlock = kBlock / 64;

template <class G>
__global__ void __launch_bounds__(kBlock) k_filter_count(G geo, CsrIn C, int32_t* __restrict__ counts) {
  const int row = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= C.n) return;  // whole wave
  int64_t b, e;
  seg_bounds(C.off, C.nn, row, C.np, b, e);
  const typename G::Q q = geo.query(row);
  int cnt = 0;
  for (int64_t t = b; t < e; t += 64) {
    const int64_t k = t + lane;
    bool keep = false;
    if (k < e) {
      const int64_t j = C.jj[k];
      typename G::P p;
      keep = j >= 0 && j < C.nref && geo.eval(q, j, p);
    }
    cnt += __popcll(__ballot(keep));
  }
  if (lane == 0) counts[row] = cnt;
}

template <class G>
__global__ void __launch_bounds__(kBlock) k_filter_fill(G geo, CsrIn C, const int32_t* __restrict__ new_off,
                                                         PairOut o) {
  const int row = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= C.n) return;  // whole wave
  int64_t b, e;
  seg_bounds(C.off, C.nn, row, C.np, b, e);
  const typename G::Q q = geo.query(row);
  int64_t pos = new_off[row];
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int64_t t = b; t < e; t += 64) {
    const int64_t k = t + lane;
    bool keep = false;
    int64_t j = -1;
    typename G::P p;
    if (k < e) {
      j = C.jj[k];
      keep = j >= 0 && j < C.nref && geo.eval(q, j, p);
    }
    const unsigned long long mask = __ballot(keep);
    if (keep) {
      const int64_t slot = pos + __popcll(mask & below);
      if (slot >= 0 && slot < o.cap) geo.write(q, row, j, p, o, slot);
    }
    pos += __popcll(mask);
  }
}

// ---------------------------------------------------------------------------
// Verlet reuse test: max |x_now - x_prior| and min support.
// ---------------------------------------------------------------------------
__global__ void k_verlet_init(unsigned* __restrict__ acc) {
  if (threadIdx.x == 0) {
    acc[0] = 0u;            // max displacement (non-negative float bits)
    acc[1] = 0x7f800000u;   // min support (+inf)
  }
}

__global__ void __launch_bounds__(kBlock) k_verlet_reduce(const float* __restrict__ a, const float* __restrict__ b,
                                                          int64_t n, const float* __restrict__ h, int64_t nh,
                                                          unsigned* __restrict__ acc) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float dmax = 0.f, hmin = __builtin_inff();
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n || k < nh; k += stride) {
    if (k < n) {
      const float dx = b[2 * k] - a[2 * k], dy = b[2 * k + 1] - a[2 * k + 1];
      dmax = fmaxf(dmax, sqrt_rn(dist2_exact(dx, dy)));
    }
    if (k < nh) hmin = fminf(hmin, h[k]);
  }
  for (int s = 32; s > 0; s >>= 1) {
    dmax = fmaxf(dmax, __shfl_down(dmax, s, 64));
    hmin = fminf(hmin, __shfl_down(hmin, s, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&acc[0], __float_as_uint(dmax));
    // supports are positive: their bit patterns order as the values
    if (hmin >= 0.f) atomicMin(&acc[1], __float_as_uint(hmin));
  }
}

// ---------------------------------------------------------------------------
// countNeighbors / buildNeighborList over a compact hash map
// (neighborhood.h:100-325)
// ---------------------------------------------------------------------------
struct HashIn {
  const float* qpos;
  const float* qh;
  const float* spos;
  const float* sh;
  const int32_t* table;
  const int32_t* ncell;
  const int64_t* cells;
  const float* dmin;
  const float* dmax;
  const uint8_t* per;
  int64_t ncell_rows;
  int nq, ns, hlen, range, mode;
  float hcell;
};

__device__ __forceinline__ int pymod_i(int n, int m) { return n >= 0 ? n % m : ((n % m) + m) % m; }

// Calls body(j) for every sorted particle j of every occupied cell around
// query i in the reference's order, and returns nothing.
template <class Body>
__device__ __forceinline__ void hash_walk(const HashIn& H, int i, Body&& body) {
  const float x = H.qpos[2 * (int64_t)i], y = H.qpos[2 * (int64_t)i + 1];
  // the float cell coordinate is clamped to +-1e9 before the conversion
  // (NaN -> -1e9): no value of the reference's range changes
  const int cx = (int)fminf(fmaxf(floorf((x - H.dmin[0]) / H.hcell), -1e9f), 1e9f);
  const int cy = (int)fminf(fmaxf(floorf((y - H.dmin[1]) / H.hcell), -1e9f), 1e9f);
  const int ncx = H.ncell[0], ncy = H.ncell[1];
  if (ncx < 1 || ncy < 1) return;  // no grid: nothing to visit
  const bool px = H.per[0] != 0, py = H.per[1] != 0;
  for (int ox = -H.range; ox <= H.range; ++ox) {
    for (int oy = -H.range; oy <= H.range; ++oy) {
      const int c0 = px ? pymod_i(cx + ox, ncx) : cx + ox;
      const int c1 = py ? pymod_i(cy + oy, ncy) : cy + oy;
      const int64_t lin = (int64_t)(int32_t)((uint32_t)c0 + (uint32_t)c1 * (uint32_t)ncx);
      const int32_t hsh = (int32_t)((uint32_t)c0 * 73856093u + (uint32_t)c1 * 19349663u);
      const int bucket = pymod_i(hsh, H.hlen);
      const int hb = H.table[2 * bucket], hl = H.table[2 * bucket + 1];
      if (hb == -1) continue;
      int64_t t0 = hb, t1 = (int64_t)hb + hl;
      if (t0 < 0) t0 = 0;
      if (t1 > H.ncell_rows) t1 = H.ncell_rows;
      for (int64_t t = t0; t < t1; ++t) {
        if (H.cells[3 * t] != lin) continue;
        int64_t j0 = H.cells[3 * t + 1], j1 = j0 + H.cells[3 * t + 2];
        if (j0 < 0) j0 = 0;
        if (j1 > H.ns) j1 = H.ns;
        for (int64_t j = j0; j < j1; ++j) body((int)j, x, y);
        break;
      }
    }
  }
}

// modDistance (neighborhood.h:93-99): float sum of double per-axis offsets.
__device__ __forceinline__ bool hash_keep(const HashIn& H, int i, int j, float x, float y) {
  const float xj = H.spos[2 * (int64_t)j], yj = H.spos[2 * (int64_t)j + 1];
  const double dx = H.per[0] != 0 ? GeoCpp::modulo_op(x, xj, H.dmax[0] - H.dmin[0]) : (double)(x - xj);
  const double dy = H.per[1] != 0 ? GeoCpp::modulo_op(y, yj, H.dmax[1] - H.dmin[1]) : (double)(y - yj);
  float sum = 0.f;
  sum = (float)((double)sum + dx * dx);
  sum = (float)((double)sum + dy * dy);
  const float dist = sqrt_rn(sum);
  if (H.mode == DSPH_MODE_SCATTER) return dist < H.sh[j];
  if (H.mode == DSPH_MODE_GATHER) return dist < H.qh[i];
  return dist < (H.qh[i] + H.sh[j]) / 2.f;
}

__global__ void __launch_bounds__(kBlock) k_hash_count(HashIn H, int32_t* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= H.nq) return;
  int cnt = 0;
  hash_walk(H, i, [&](int j, float x, float y) { cnt += hash_keep(H, i, j, x, y) ? 1 : 0; });
  counts[i] = cnt;
}

__global__ void __launch_bounds__(kBlock) k_hash_build(HashIn H, const int32_t* __restrict__ offsets, int64_t cap,
                                                       int32_t* __restrict__ oi, int32_t* __restrict__ oj) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= H.nq) return;
  int64_t pos = offsets[i];
  hash_walk(H, i, [&](int j, float x, float y) {
    if (!hash_keep(H, i, j, x, y)) return;
    if (pos >= 0 && pos < cap) {
      oi[pos] = i;
      oj[pos] = j;
    }
    ++pos;
  });
}

}  // namespace dsph

using namespace dsph;

namespace {

bool csr_args_ok(int32_t n, int64_t nref, int64_t np, const int64_t* jj, const int32_t* nn, const int32_t* off) {
  if (n < 0 || nref < 0 || np < 0) return false;
  if (n > 0 && (!nn || !off)) return false;
  if (np > 0 && !jj) return false;
  return true;
}

int grid_rows(int64_t n) { return grid_for(n * 64); }

HashIn make_hash(const float* qpos, const float* qh, int32_t nq, int32_t range, const float* spos, const float* sh,
                 int32_t ns, const int32_t* table, int32_t hlen, const int32_t* ncell, const int64_t* cells,
                 int64_t ncell_rows, float hcell, const float* dmin, const float* dmax, const uint8_t* per,
                 int32_t mode) {
  HashIn H;
  H.qpos = qpos;
  H.qh = qh;
  H.spos = spos;
  H.sh = sh;
  H.table = table;
  H.ncell = ncell;
  H.cells = cells;
  H.dmin = dmin;
  H.dmax = dmax;
  H.per = per;
  H.ncell_rows = ncell_rows;
  H.nq = nq;
  H.ns = ns;
  H.hlen = hlen;
  H.range = range;
  H.mode = mode;
  H.hcell = hcell;
  return H;
}

int hash_args_status(const HashIn& H) {
  if (H.nq < 0 || H.ns < 0 || H.range < 0 || H.range > 8 || H.hlen < 1 || H.ncell_rows < 0) return DSPH_ERR_ARG;
  if (H.mode < 0 || H.mode > 2 || !(H.hcell > 0.f)) return DSPH_ERR_ARG;
  if (H.nq == 0) return DSPH_OK;
  if (!H.qpos || !H.table || !H.ncell || !H.dmin || !H.dmax || !H.per) return DSPH_ERR_ARG;
  if (H.ns > 0 && !H.spos) return DSPH_ERR_ARG;
  if (H.ncell_rows > 0 && !H.cells) return DSPH_ERR_ARG;
  if (H.mode != DSPH_MODE_SCATTER && !H.qh) return DSPH_ERR_ARG;
  if (H.mode != DSPH_MODE_GATHER && H.ns > 0 && !H.sh) return DSPH_ERR_ARG;
  return DSPH_OK;
}

}  // namespace

extern "C" int dsph_op_viscosity(const float* velocities_i, const float* velocities_j, const float* masses_j,
                                 const float* densities_i, const float* densities_j, const float* supports_i,
                                 int64_t num_ref, const int64_t* indices_j, const float* supports,
                                 const float* gradients, const float* distances, const float* vectors,
                                 int64_t num_pairs, int32_t num_particles, const int32_t* num_neighbors,
                                 const int32_t* neighbor_offset, float cs, float rho0, float alpha, float eps,
                                 int32_t pi_switch, float* out, void* stream) {
  if (!csr_args_ok(num_particles, num_ref, num_pairs, indices_j, num_neighbors, neighbor_offset)) return DSPH_ERR_ARG;
  if (num_particles == 0) return DSPH_OK;
  if (!velocities_i || !densities_i || !supports_i || !out) return DSPH_ERR_ARG;
  if (num_pairs > 0 && (!supports || !gradients || !distances || !vectors)) return DSPH_ERR_ARG;
  if (num_ref > 0 && (!velocities_j || !masses_j || !densities_j)) return DSPH_ERR_ARG;
  k_op_viscosity<<<grid_for(num_particles), kBlock, 0, as_stream(stream)>>>(
      velocities_i, velocities_j, masses_j, densities_i, densities_j, supports_i, indices_j, supports, gradients,
      distances, vectors, num_particles, num_ref, num_pairs, num_neighbors, neighbor_offset, cs, rho0, alpha, eps,
      pi_switch != 0, out);
  return launch_status();
}

extern "C" int dsph_op_mdbc(const float* masses_j, const float* densities_j, int64_t num_ref,
                            const float* positions_boundary, const float* positions_ghost, int32_t num_boundary,
                            const int64_t* indices_j, const float* kernels, const float* gradients,
                            const float* distances, const float* vectors, const float* supports, int64_t num_pairs,
                            const int32_t* num_neighbors, const int32_t* neighbor_offset, float rho0,
                            float* mdbc_density, float* shepard_density, void* stream) {
  if (!csr_args_ok(num_boundary, num_ref, num_pairs, indices_j, num_neighbors, neighbor_offset)) return DSPH_ERR_ARG;
  if (num_boundary == 0) return DSPH_OK;
  if (!positions_boundary || !positions_ghost || !mdbc_density || !shepard_density) return DSPH_ERR_ARG;
  if (num_pairs > 0 && (!kernels || !gradients || !distances || !vectors || !supports)) return DSPH_ERR_ARG;
  if (num_ref > 0 && (!masses_j || !densities_j)) return DSPH_ERR_ARG;
  k_op_mdbc<<<grid_for(num_boundary), kBlock, 0, as_stream(stream)>>>(
      masses_j, densities_j, num_ref, positions_boundary, positions_ghost, num_boundary, indices_j, kernels,
      gradients, distances, vectors, supports, num_pairs, num_neighbors, neighbor_offset, rho0, mdbc_density,
      shepard_density);
  return launch_status();
}

namespace {

int cpp_geo(const float* pi, const float* hi, const float* pj, const float* hj, int32_t n, int64_t nref,
            const float* dmin, const float* dmax, const uint8_t* per, GeoCpp& g) {
  if (n > 0 && (!pi || !hi || !dmin || !dmax || !per)) return DSPH_ERR_ARG;
  if (nref > 0 && (!pj || !hj)) return DSPH_ERR_ARG;
  g.pi = pi;
  g.hi = hi;
  g.pj = pj;
  g.hj = hj;
  g.dmin = dmin;
  g.dmax = dmax;
  g.per = per;
  return DSPH_OK;
}

int py_geo(const dsph_domain* dom, int32_t mode, const float* pi, const float* hi, const float* pj, const float* hj,
           int32_t n, int64_t nref, GeoPy& g) {
  if (!dom || mode < 0 || mode > 2) return DSPH_ERR_ARG;
  if (n > 0 && !pi) return DSPH_ERR_ARG;
  if (nref > 0 && !pj) return DSPH_ERR_ARG;
  if (mode != DSPH_MODE_SCATTER && n > 0 && !hi) return DSPH_ERR_ARG;
  if (mode != DSPH_MODE_GATHER && nref > 0 && !hj) return DSPH_ERR_ARG;
  g.pi = pi;
  g.hi = hi;
  g.pj = pj;
  g.hj = hj;
  g.dom = make_dom(*dom);
  g.mode = mode;
  return DSPH_OK;
}

CsrIn make_csr(const int64_t* jj, const int32_t* nn, const int32_t* off, int64_t np, int64_t nref, int32_t n) {
  CsrIn c;
  c.jj = jj;
  c.nn = nn;
  c.off = off;
  c.np = np;
  c.nref = nref;
  c.n = n;
  return c;
}

}  // namespace

extern "C" int dsph_op_neighborhood_count(const float* positions_i, const float* supports_i, int32_t num_particles,
                                          const float* positions_j, const float* supports_j, int64_t num_ref,
                                          const int64_t* indices_j, int64_t num_pairs, const int32_t* num_neighbors,
                                          const int32_t* neighbor_offset, const float* min_domain,
                                          const float* max_domain, const uint8_t* periodicity, int32_t* counts,
                                          void* stream) {
  if (!csr_args_ok(num_particles, num_ref, num_pairs, indices_j, num_neighbors, neighbor_offset)) return DSPH_ERR_ARG;
  if (num_particles == 0) return DSPH_OK;
  GeoCpp g;
  int s = cpp_geo(positions_i, supports_i, positions_j, supports_j, num_particles, num_ref, min_domain, max_domain,
                  periodicity, g);
  if (s != DSPH_OK) return s;
  if (!counts) return DSPH_ERR_ARG;
  k_filter_count<GeoCpp><<<grid_rows(num_particles), kBlock, 0, as_stream(stream)>>>(
      g, make_csr(indices_j, num_neighbors, neighbor_offset, num_pairs, num_ref, num_particles), counts);
  return launch_status();
}

extern "C" int dsph_op_neighborhood_fill(const float* positions_i, const float* supports_i, int32_t num_particles,
                                         const float* positions_j, const float* supports_j, int64_t num_ref,
                                         const int64_t* indices_j, int64_t num_pairs, const int32_t* num_neighbors,
                                         const int32_t* neighbor_offset, const float* min_domain,
                                         const float* max_domain, const uint8_t* periodicity,
                                         const int32_t* new_offsets, int64_t capacity, int64_t* out_i,
                                         int64_t* out_j, float* out_rij, float* out_xij, float* out_hij,
                                         void* stream) {
  if (!csr_args_ok(num_particles, num_ref, num_pairs, indices_j, num_neighbors, neighbor_offset)) return DSPH_ERR_ARG;
  if (capacity < 0) return DSPH_ERR_ARG;
  if (num_particles == 0 || capacity == 0) return DSPH_OK;
  GeoCpp g;
  int s = cpp_geo(positions_i, supports_i, positions_j, supports_j, num_particles, num_ref, min_domain, max_domain,
                  periodicity, g);
  if (s != DSPH_OK) return s;
  if (!new_offsets || !out_i || !out_j || !out_rij || !out_xij || !out_hij) return DSPH_ERR_ARG;
  PairOut o{out_i, out_j, out_rij, out_xij, out_hij, nullptr, nullptr, capacity};
  k_filter_fill<GeoCpp><<<grid_rows(num_particles), kBlock, 0, as_stream(stream)>>>(
      g, make_csr(indices_j, num_neighbors, neighbor_offset, num_pairs, num_ref, num_particles), new_offsets, o);
  return launch_status();
}

extern "C" int dsph_pairs_filter_count(const dsph_domain* domain, int32_t mode, const float* positions_i,
                                       const float* supports_i, int32_t num_particles, const float* positions_j,
                                       const float* supports_j, int64_t num_ref, const int64_t* indices_j,
                                       int64_t num_pairs, const int32_t* num_neighbors,
                                       const int32_t* neighbor_offset, int32_t* counts, void* stream) {
  if (!csr_args_ok(num_particles, num_ref, num_pairs, indices_j, num_neighbors, neighbor_offset)) return DSPH_ERR_ARG;
  GeoPy g;
  int s = py_geo(domain, mode, positions_i, supports_i, positions_j, supports_j, num_particles, num_ref, g);
  if (s != DSPH_OK) return s;
  if (num_particles == 0) return DSPH_OK;
  if (!counts) return DSPH_ERR_ARG;
  k_filter_count<GeoPy><<<grid_rows(num_particles), kBlock, 0, as_stream(stream)>>>(
      g, make_csr(indices_j, num_neighbors, neighbor_offset, num_pairs, num_ref, num_particles), counts);
  return launch_status();
}

extern "C" int dsph_pairs_filter_fill(const dsph_domain* domain, int32_t mode, const float* positions_i,
                                      const float* supports_i, int32_t num_particles, const float* positions_j,
                                      const float* supports_j, int64_t num_ref, const int64_t* indices_j,
                                      int64_t num_pairs, const int32_t* num_neighbors,
                                      const int32_t* neighbor_offset, const int32_t* new_offsets, int64_t capacity,
                                      int64_t* out_i, int64_t* out_j, float* out_rij, float* out_xij,
                                      float* out_hij, float* out_kernels, float* out_gradients, void* stream) {
  if (!csr_args_ok(num_particles, num_ref, num_pairs, indices_j, num_neighbors, neighbor_offset)) return DSPH_ERR_ARG;
  if (capacity < 0) return DSPH_ERR_ARG;
  GeoPy g;
  int s = py_geo(domain, mode, positions_i, supports_i, positions_j, supports_j, num_particles, num_ref, g);
  if (s != DSPH_OK) return s;
  if (num_particles == 0 || capacity == 0) return DSPH_OK;
  if (!new_offsets || !out_i || !out_j || !out_rij || !out_xij || !out_hij) return DSPH_ERR_ARG;
  PairOut o{out_i, out_j, out_rij, out_xij, out_hij, out_kernels, out_gradients, capacity};
  k_filter_fill<GeoPy><<<grid_rows(num_particles), kBlock, 0, as_stream(stream)>>>(
      g, make_csr(indices_j, num_neighbors, neighbor_offset, num_pairs, num_ref, num_particles), new_offsets, o);
  return launch_status();
}

extern "C" size_t dsph_verlet_check_workspace(void) { return 256; }

extern "C" int dsph_verlet_check(const float* pos_prior, const float* pos_now, int64_t n, const float* supports,
                                 int64_t num_supports, float* out_host, void* workspace, size_t workspace_bytes,
                                 void* stream) {
  if (n < 0 || num_supports < 0 || !out_host) return DSPH_ERR_ARG;
  if ((n > 0 && (!pos_prior || !pos_now)) || (num_supports > 0 && !supports)) return DSPH_ERR_ARG;
  if (!workspace || workspace_bytes < dsph_verlet_check_workspace()) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  unsigned* acc = static_cast<unsigned*>(workspace);
  k_verlet_init<<<1, 64, 0, st>>>(acc);
  DSPH_LAUNCH_CHECK();
  const int64_t m = n > num_supports ? n : num_supports;
  if (m > 0) {
    int64_t g = (m + kBlock - 1) / kBlock;
    if (g > 2048) g = 2048;
    k_verlet_reduce<<<(int)g, kBlock, 0, st>>>(pos_prior, pos_now, n, supports, num_supports, acc);
    DSPH_LAUNCH_CHECK();
  }
  unsigned host[2];
  DSPH_CHECK_HIP(hipMemcpyAsync(host, acc, sizeof(host), hipMemcpyDeviceToHost, st));
  DSPH_CHECK_HIP(hipStreamSynchronize(st));
  memcpy(&out_host[0], &host[0], sizeof(float));
  memcpy(&out_host[1], &host[1], sizeof(float));
  return DSPH_OK;
}

extern "C" int dsph_hash_count_neighbors(const float* query_positions, const float* query_supports,
                                         int32_t num_query, int32_t search_range, const float* sorted_positions,
                                         const float* sorted_supports, int32_t num_sorted,
                                         const int32_t* hash_table, int32_t hash_map_length,
                                         const int32_t* num_cells, const int64_t* cell_table,
                                         int64_t num_cell_rows, float h_cell, const float* min_domain,
                                         const float* max_domain, const uint8_t* periodicity, int32_t mode,
                                         int32_t* counts, void* stream) {
  HashIn H = make_hash(query_positions, query_supports, num_query, search_range, sorted_positions, sorted_supports,
                       num_sorted, hash_table, hash_map_length, num_cells, cell_table, num_cell_rows, h_cell,
                       min_domain, max_domain, periodicity, mode);
  int s = hash_args_status(H);
  if (s != DSPH_OK || num_query == 0) return s;
  if (!counts) return DSPH_ERR_ARG;
  k_hash_count<<<grid_for(num_query), kBlock, 0, as_stream(stream)>>>(H, counts);
  return launch_status();
}

extern "C" int dsph_hash_build_neighbors(const int32_t* offsets, int64_t capacity, const float* query_positions,
                                         const float* query_supports, int32_t num_query, int32_t search_range,
                                         const float* sorted_positions, const float* sorted_supports,
                                         int32_t num_sorted, const int32_t* hash_table, int32_t hash_map_length,
                                         const int32_t* num_cells, const int64_t* cell_table,
                                         int64_t num_cell_rows, float h_cell, const float* min_domain,
                                         const float* max_domain, const uint8_t* periodicity, int32_t mode,
                                         int32_t* out_i, int32_t* out_j, void* stream) {
  HashIn H = make_hash(query_positions, query_supports, num_query, search_range, sorted_positions, sorted_supports,
                       num_sorted, hash_table, hash_map_length, num_cells, cell_table, num_cell_rows, h_cell,
                       min_domain, max_domain, periodicity, mode);
  int s = hash_args_status(H);
  if (s != DSPH_OK) return s;
  if (capacity < 0) return DSPH_ERR_ARG;
  if (num_query == 0 || capacity == 0) return DSPH_OK;
  if (!offsets || !out_i || !out_j) return DSPH_ERR_ARG;
  k_hash_build<<<grid_for(num_query), kBlock, 0, as_stream(stream)>>>(H, offsets, capacity, out_i, out_j);
  return launch_status();
}
