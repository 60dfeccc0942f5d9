// CPU oracle of the diffSPH particle hot path.
//
// TEST INFRASTRUCTURE ONLY: this library is the parity checker and the timed
// CPU baseline.  Only tests/, the smoke check and the CPU-baseline leg of the
// benchmark load it; the product path (diffsph_amd) never does.
//
// It restates the reference algorithms in plain C++17 + OpenMP, one OpenMP
// loop over particles walking a CSR neighbor list, as the reference's own
// CPU twins do (partiBench/cppOp.cpp:57-69, partiBench/viscosityKernel.cpp:
// 74-77, neighborhood.cpp:204-223).  Each function names the reference lines
// it follows.  Arithmetic is templated on the scalar type so the same code
// gives the float32 oracle and the float64 error reference.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).  FP
// contraction is off so that the float32 neighbor predicate is evaluated
// with separately rounded products, the contract the GPU search follows.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

// Portable float32 atan2/sin/cos (float64 series, IEEE arithmetic only) and
// the closed-form 2x2 pseudo-inverse, shared with the HIP build so that the
// float32 normalization matrices agree bit for bit.  The functions are
// pinned against numpy (atan2, sin, cos, linalg.pinv) in tests/test_oracle.py;
// the float64 path below restates the pseudo-inverse with <cmath>.
#include "../diffsph_amd/csrc/pmath.h"

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

struct orc_domain {
  float min[2];
  float max[2];
  int32_t periodic[2];
};

}  // extern "C"

namespace {

// torch.remainder for floats (the reference's `%` and torch.remainder).
template <class T>
T py_remainder(T a, T b) {
  T r = std::fmod(a, b);
  if (r != T(0) && ((r < T(0)) != (b < T(0)))) r = r + b;
  return r;
}

template <class T>
struct DomT {
  T min[2], L[2], half[2];
  int periodic[2];
  explicit DomT(const orc_domain& d) {
    for (int k = 0; k < 2; ++k) {
      min[k] = (T)d.min[k];
      // float32 tensor arithmetic: max - min in the working type.
      L[k] = (T)d.max[k] - (T)d.min[k];
      half[k] = (T)(((double)d.max[k] - (double)d.min[k]) / 2.0);
      periodic[k] = d.periodic[k] != 0;
    }
  }
};

// getPeriodicPosition, v2/modules/neighborhood.py:407-430.
template <class T>
T wrap(T x, const DomT<T>& d, int k) {
  if (!d.periodic[k]) return x;
  return py_remainder<T>(x - d.min[k], d.L[k]) + d.min[k];
}

// mod, v2/math.py:4-6.
template <class T>
T modoff(T x, const DomT<T>& d, int k) {
  if (!d.periodic[k]) return x;
  T ax = std::fabs(x);
  if (ax > d.half[k]) {
    T s = x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0));
    return s * (py_remainder<T>(ax + d.min[k], d.L[k]) + d.min[k]);
  }
  return x;
}

template <class T>
T pick_h(int mode, T hi, T hj) {
  return mode == 0 ? hj : (mode == 1 ? hi : T(0.5) * (hi + hj));
}

// Wendland C2, 2-D (kernelFunctions/Wendland2.py:5-36); cpow clamps to [0,1]
// (kernels.py:69-70).
template <class T>
T cpow(T q, int p) {
  T t = std::min(std::max(q, T(0)), T(1));
  T r = T(1);
  for (int k = 0; k < p; ++k) r *= t;
  return r;
}
template <class T>
T C2() { return (T)(7.0 / 3.14159265358979323846); }
// k(q) C / h^2 and dk/dq C / h^3, evaluated left to right as written in
// Wendland2.py:30-36 (cpow(x,p) = x*x*...*x).
template <class T>
T kernelW_exact(T q, T h) { return ((cpow<T>(T(1) - q, 4) * (T(1) + T(4) * q)) * C2<T>()) / (h * h); }
template <class T>
T kernelDW_exact(T q, T h) { return (((T(-20) * q) * cpow<T>(T(1) - q, 3)) * C2<T>()) / ((h * h) * h); }

// The solvers' reciprocal form: the constants 1/h, C/h^2 and C/h^3 of a
// support are evaluated once and multiplied in (the GPU kernels hoist them
// out of their pair loops), a few ulp from the quotients above.
template <class T>
struct KC {
  T ih, cw, cd;
};
template <class T>
KC<T> kconst(T h) {
  KC<T> k;
  k.ih = T(1) / h;
  k.cw = C2<T>() / (h * h);
  k.cd = C2<T>() / ((h * h) * h);
  return k;
}
template <class T>
T kernelW(T q, const KC<T>& k) { return (cpow<T>(T(1) - q, 4) * (T(1) + T(4) * q)) * k.cw; }
template <class T>
T kernelDW(T q, const KC<T>& k) { return ((T(-20) * q) * cpow<T>(T(1) - q, 3)) * k.cd; }
template <class T>
T kernelW(T q, T h) { return kernelW<T>(q, kconst<T>(h)); }
template <class T>
T kernelDW(T q, T h) { return kernelDW<T>(q, kconst<T>(h)); }

// Pair geometry with the reference's conventions
// (v2/modules/neighborhood.py:445-451, 587-591): d = x_i - x_j,
// xhat = d / (r + 1e-7), q = r / h, in the solvers' reciprocal form
// q = r * (1/h), xhat = d * (1/(r + 1e-7)) (one division per pair, as the
// GPU evaluates it).
template <class T>
struct G {
  T dx, dy, r, q, h, ri, ex, ey, w, dw, gx, gy;
};

template <class T>
G<T> geom(const T* xi, const T* xj, T h, const DomT<T>& d) {
  G<T> g;
  g.dx = modoff<T>(xi[0] - xj[0], d, 0);
  g.dy = modoff<T>(xi[1] - xj[1], d, 1);
  T a = g.dx * g.dx;
  T b = g.dy * g.dy;
  g.r = std::sqrt(a + b);
  g.h = h;
  const KC<T> k = kconst<T>(h);
  g.q = g.r * k.ih;
  g.ri = T(1) / (g.r + (T)1e-7);
  g.ex = g.dx * g.ri;
  g.ey = g.dy * g.ri;
  g.w = kernelW<T>(g.q, k);
  g.dw = kernelDW<T>(g.q, k);
  g.gx = g.ex * g.dw;
  g.gy = g.ey * g.dw;
  return g;
}

// The reference's own quotients, for the per-pair arrays handed to
// reference callers (neighborDict exports).
template <class T>
G<T> geom_exact(const T* xi, const T* xj, T h, const DomT<T>& d) {
  G<T> g;
  g.dx = modoff<T>(xi[0] - xj[0], d, 0);
  g.dy = modoff<T>(xi[1] - xj[1], d, 1);
  T a = g.dx * g.dx;
  T b = g.dy * g.dy;
  g.r = std::sqrt(a + b);
  g.h = h;
  g.q = g.r / h;
  g.ri = T(1) / (g.r + (T)1e-7);
  g.ex = g.dx / (g.r + (T)1e-7);
  g.ey = g.dy / (g.r + (T)1e-7);
  g.w = kernelW_exact<T>(g.q, h);
  g.dw = kernelDW_exact<T>(g.q, h);
  g.gx = g.ex * g.dw;
  g.gy = g.ey * g.dw;
  return g;
}

// The DFSPH pair coefficient V_j gradW_ij in the factored form the GPU
// solver stores: t = dW/dr / (r + 1e-7), gradW = d * t, s = V_j * t,
// V_j gradW = d * s (algebraically the reference's V_j * xhat * dW/dr,
// v2/modules/dfsph.py:13-20, 35-50; a few ulp apart).
template <class T>
T dfsph_t(const G<T>& g) {
  return g.dw * g.ri;
}

// The neighbor predicate in float32: r = sqrt(dx*dx + dy*dy) with separately
// rounded products (this file is compiled with -ffp-contract=off), q = r/h,
// keep iff q <= 1 (v2/modules/neighborhood.py:445-453).
bool keep_f32(const float* qi, float hi, const float* rj, float hj, const DomT<float>& d, int mode) {
  float dx = modoff<float>(qi[0] - rj[0], d, 0);
  float dy = modoff<float>(qi[1] - rj[1], d, 1);
  float a = dx * dx;
  float b = dy * dy;
  float r = std::sqrt(a + b);
  float h = pick_h<float>(mode, hi, hj);
  return r / h <= 1.0f;
}

int cell_of(float x, const DomT<float>& d, int k, int nc, float hcell) {
  // clamped before the conversion (NaN -> cell 0), as the GPU search does
  float f = std::floor((x - d.min[k]) / hcell);
  f = std::fmin(std::fmax(f, 0.0f), (float)(nc - 1));
  return (int)f;
}

int nthreads_override = 0;

void set_threads() {
#ifdef _OPENMP
  if (nthreads_override > 0) omp_set_num_threads(nthreads_override);
#endif
}

template <class T>
const T* row2(const T* a, int64_t i) { return a + 2 * i; }

}  // namespace

extern "C" {

void orc_set_threads(int n) { nthreads_override = n; }

int orc_max_threads(void) {
#ifdef _OPENMP
  return nthreads_override > 0 ? nthreads_override : omp_get_max_threads();
#else
  return 1;
#endif
}

// Grid of the cell-list search: one square side for both axes, periodic
// axes tiled by floor(L/hcell) cells (the reference's compute_h,
// neighborhoodDebug.ipynb cell 2), cells 0.1% wider than the search radius.
int orc_grid_dims(const orc_domain* dom, float hsearch, int32_t* ncx, int32_t* ncy, float* hcell_out) {
  const double hs = (double)hsearch * 1.001;
  double hc = 0.0;
  for (int k = 0; k < 2; ++k) {
    double L = (double)dom->max[k] - (double)dom->min[k];
    if (!(L > 0.0)) return 1;
    double n = std::floor(L / hs);
    if (n < 1.0) n = 1.0;
    hc = std::max(hc, L / n);
  }
  float hcell = (float)hc;
  if ((double)hcell < hs) hcell = std::nextafter(hcell, INFINITY);
  int nc[2];
  for (int k = 0; k < 2; ++k) {
    double L = (double)dom->max[k] - (double)dom->min[k];
    double n = dom->periodic[k] ? std::floor(L / (double)hcell + 1e-4) : std::ceil(L / (double)hcell);
    if (dom->periodic[k] && n < 3.0) return 2;
    if (n < 1.0) n = 1.0;
    nc[k] = (int)n;
  }
  *ncx = nc[0];
  *ncy = nc[1];
  *hcell_out = hcell;
  return 0;
}

// Counting sort by cell key, stable in the original index
// (sortReferenceParticles + buildCompactHashMap, neighborhoodDebug.ipynb
// cells 3-4, with a dense table instead of a hash map).
int orc_grid_build(const orc_domain* dom, int32_t ncx, int32_t ncy, float hcell, const float* pos, int64_t n,
                   int32_t* cell_start, int32_t* sort_index, float* sorted_pos) {
  DomT<float> d(*dom);
  const int64_t ncell = (int64_t)ncx * ncy;
  std::vector<int32_t> key(n);
  std::vector<int32_t> cnt(ncell + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    float wx = wrap<float>(pos[2 * i], d, 0), wy = wrap<float>(pos[2 * i + 1], d, 1);
    int cx = cell_of(wx, d, 0, ncx, hcell), cy = cell_of(wy, d, 1, ncy, hcell);
    key[i] = cy * ncx + cx;
    cnt[key[i]]++;
  }
  cell_start[0] = 0;
  for (int64_t c = 0; c < ncell; ++c) cell_start[c + 1] = cell_start[c] + cnt[c];
  std::vector<int32_t> cur(cell_start, cell_start + ncell);
  for (int64_t i = 0; i < n; ++i) sort_index[cur[key[i]]++] = (int32_t)i;
  for (int64_t p = 0; p < n; ++p) {
    int32_t i = sort_index[p];
    sorted_pos[2 * p] = wrap<float>(pos[2 * i], d, 0);
    sorted_pos[2 * p + 1] = wrap<float>(pos[2 * i + 1], d, 1);
  }
  return 0;
}

// Cell-list neighbor search in the canonical order (rows = queries, cells
// dy-major then dx over offsets -1..1, particles in sorted order).  Pass
// nbr == NULL to count only.  Queries are wrapped here (qpos raw).
int orc_neighbors_cell(const orc_domain* dom, int32_t ncx, int32_t ncy, float hcell, const int32_t* cell_start,
                       const float* spos, const float* sh, const float* qpos, const float* qh, int64_t nq,
                       float h0, int32_t mode, int32_t* counts, const int32_t* offsets, int32_t* nbr) {
  DomT<float> d(*dom);
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; ++i) {
    float qi[2] = {wrap<float>(qpos[2 * i], d, 0), wrap<float>(qpos[2 * i + 1], d, 1)};
    float hi = qh ? qh[i] : h0;
    int cx = cell_of(qi[0], d, 0, ncx, hcell), cy = cell_of(qi[1], d, 1, ncy, hcell);
    int c = 0;
    int64_t o = offsets ? offsets[i] : 0;
    for (int oy = -1; oy <= 1; ++oy) {
      int yy = cy + oy;
      if (d.periodic[1]) yy = (yy + ncy) % ncy;
      else if (yy < 0 || yy >= ncy) continue;
      for (int ox = -1; ox <= 1; ++ox) {
        int xx = cx + ox;
        if (d.periodic[0]) xx = (xx + ncx) % ncx;
        else if (xx < 0 || xx >= ncx) continue;
        int cell = yy * ncx + xx;
        for (int j = cell_start[cell]; j < cell_start[cell + 1]; ++j) {
          float hj = sh ? sh[j] : h0;
          if (keep_f32(qi, hi, spos + 2 * (int64_t)j, hj, d, mode)) {
            if (nbr) nbr[o++] = j;
            ++c;
          }
        }
      }
    }
    if (counts) counts[i] = c;
  }
  return 0;
}

// Brute force O(nq*nr) (v2/neighborhoodFallback/naive.py:9-35 with the
// filter predicate of v2/modules/neighborhood.py:453).  Rows in query order,
// j ascending.  nbr == NULL counts only.
int orc_neighbors_brute(const orc_domain* dom, const float* qpos, const float* qh, int64_t nq, const float* rpos,
                        const float* rh, int64_t nr, float h0, int32_t mode, int32_t* counts, const int32_t* offsets,
                        int32_t* nbr) {
  DomT<float> d(*dom);
  set_threads();
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nq; ++i) {
    float qi[2] = {wrap<float>(qpos[2 * i], d, 0), wrap<float>(qpos[2 * i + 1], d, 1)};
    float hi = qh ? qh[i] : h0;
    int c = 0;
    int64_t o = offsets ? offsets[i] : 0;
    for (int64_t j = 0; j < nr; ++j) {
      float rj[2] = {wrap<float>(rpos[2 * j], d, 0), wrap<float>(rpos[2 * j + 1], d, 1)};
      float hj = rh ? rh[j] : h0;
      if (keep_f32(qi, hi, rj, hj, d, mode)) {
        if (nbr) nbr[o++] = (int32_t)j;
        ++c;
      }
    }
    if (counts) counts[i] = c;
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Pair operators over a CSR list; positions are the wrapped positions.
// ---------------------------------------------------------------------------
namespace {

template <class T>
struct Csr {
  int64_t nq;
  const int32_t* off;
  const int32_t* nbr;
  const T* qpos;
  const T* rpos;
  const T* qh;
  const T* rh;
  T h0;
  int mode;
};

template <class T>
G<T> pg(const Csr<T>& c, int64_t i, int32_t j, const DomT<T>& d) {
  T hi = c.qh ? c.qh[i] : c.h0;
  T hj = c.rh ? c.rh[j] : c.h0;
  return geom<T>(c.qpos + 2 * i, c.rpos + 2 * (int64_t)j, pick_h<T>(c.mode, hi, hj), d);
}

// sphDensityInterpolation, v2/sphOps.py:19-29.
template <class T>
void density(const Csr<T>& c, const DomT<T>& d, const T* rmass, const T* qmass, T* rho, T* vol) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    T acc = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      acc += rmass[j] * g.w;
    }
    rho[i] = acc;
    if (vol) vol[i] = qmass[i] / acc;
  }
}

// v2/sphOps.py:6-149 (interpolate / gradient / divergence).
template <class T>
void sph_operation(const Csr<T>& c, const DomT<T>& d, int op, int gm, const T* qm, const T* rm, const T* qr,
                   const T* rr, const T* qq, const T* rq, int W, T* out) {
  (void)qm;
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    T acc[4] = {0, 0, 0, 0};
    T qi[2] = {0, 0};
    for (int cc = 0; cc < W; ++cc) qi[cc] = qq ? qq[i * W + cc] : T(0);
    T rho_i = qr ? qr[i] : T(1);
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      T qj[2] = {0, 0};
      for (int cc = 0; cc < W; ++cc) qj[cc] = rq[(int64_t)j * W + cc];
      if (op == 0) {
        T kk = rm[j] / rr[j] * g.w;
        for (int cc = 0; cc < W; ++cc) acc[cc] += qj[cc] * kk;
      } else if (gm == 3) {
        T kx = rm[j] * g.gx, ky = rm[j] * g.gy;
        T irj = T(1) / (rr[j] * rr[j]), iri = T(1) / (rho_i * rho_i);
        if (op == 1) {
          for (int cc = 0; cc < W; ++cc) {
            T a = qi[cc] * iri + qj[cc] * irj;
            acc[2 * cc] += a * kx;
            acc[2 * cc + 1] += a * ky;
          }
        } else {
          acc[0] += (qi[0] * iri + qj[0] * irj) * kx + (qi[1] * iri + qj[1] * irj) * ky;
        }
      } else {
        // op 3: divergenceConsistent (sphOps.py:108-149), m_j / rho_i
        T V = op == 3 ? rm[j] / rho_i : rm[j] / rr[j];
        T kx = V * g.gx, ky = V * g.gy;
        if (op == 1) {
          for (int cc = 0; cc < W; ++cc) {
            T a = gm == 1 ? qi[cc] - qj[cc] : (gm == 2 ? qi[cc] + qj[cc] : qj[cc]);
            acc[2 * cc] += a * kx;
            acc[2 * cc + 1] += a * ky;
          }
        } else {
          T ax = gm == 1 ? qj[0] - qi[0] : (gm == 2 ? qj[0] + qi[0] : qj[0]);
          T ay = gm == 1 ? qj[1] - qi[1] : (gm == 2 ? qj[1] + qi[1] : qj[1]);
          acc[0] += ax * kx + ay * ky;
        }
      }
    }
    if (op == 0) {
      for (int cc = 0; cc < W; ++cc) out[i * W + cc] = acc[cc];
    } else if (op == 1) {
      for (int cc = 0; cc < 2 * W; ++cc) out[i * 2 * W + cc] = gm == 3 ? rho_i * acc[cc] : acc[cc];
    } else {
      out[i] = gm == 3 ? rho_i * acc[0] : acc[0];
    }
  }
}

struct ViscP {
  int on, scheme, pi_switch;
  double alpha, cs, rho0, eps;
};

// computeAlpha (v2/modules/dfsph.py:7-27) for both solves, divergence of v
// (computeSourceTerm, :30-33) and the delta-SPH viscosity
// (v2/modules/viscosity.py:123-170 inviscid, :200-218 viscid).
template <class T>
void dfsph_prepare(const Csr<T>& c, const DomT<T>& d, const T* vel, const T* areas, const T* vol, const T* rho, T dt,
                   T rho0, const ViscP& vp, T* alpha_div, T* alpha_den, T* div_src, T* visc) {
  const T dt2 = (T)((double)dt * (double)dt);
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    T k1x = 0, k1y = 0, k2 = 0, dv = 0, vsx = 0, vsy = 0;
    const T hi = c.qh ? c.qh[i] : c.h0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      T Vj = vol[j];
      const T t = dfsph_t(g);
      const T sv = Vj * t;
      T kx = g.dx * sv, ky = g.dy * sv;
      k1x += kx;
      k1y += ky;
      T c2 = (Vj * Vj) / (areas[j] * rho0);
      const T gx = g.dx * t, gy = g.dy * t;
      k2 += c2 * (gx * gx + gy * gy);
      dv += (vel[2 * j] - vel[2 * i]) * kx + (vel[2 * j + 1] - vel[2 * i + 1]) * ky;
      if (vp.on) {
        T vx = vel[2 * i] - vel[2 * j], vy = vel[2 * i + 1] - vel[2 * j + 1];
        T rij = g.q * g.h;
        T vr = vx * g.ex + vy * g.ey;
        T pi = vr / (rij + (T)vp.eps * g.h * g.h);
        if (vp.pi_switch) pi = vr < T(0) ? pi : T(0);
        vsx += (Vj * pi) * g.gx;
        vsy += (Vj * pi) * g.gy;
      }
    }
    T Vi = vol[i];
    T mass_i = areas[i] * rho0;
    T k1sq = k1x * k1x + k1y * k1y;
    T fac = -dt2 * Vi;
    alpha_den[i] = fac / mass_i * k1sq + fac * k2;
    fac = -dt * Vi;
    alpha_div[i] = fac / mass_i * k1sq + fac * k2;
    div_src[i] = dv;
    if (vp.on && visc) {
      T s;
      if (vp.scheme == 0)
        s = (T)vp.alpha * hi * (T)vp.cs * (T)vp.rho0 / rho[i];
      else
        s = (T)(2.0 * 4.0 * vp.alpha * vp.rho0) / rho[i];
      visc[2 * i] = s * vsx;
      visc[2 * i + 1] = s * vsy;
    }
  }
}

// computePressureAcceleration (v2/modules/dfsph.py:35-37).
// isph: the 'difference' form of solveIncompressible_relaxedJacobi
// (incompressible.py:276-285), sum_j (p_i - p_j) V_j gradW_ij.
template <class T>
void sweep_a(const Csr<T>& c, const DomT<T>& d, const T* vol, const T* rho, const T* p, T* acc, int isph = 0) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    T ax = 0, ay = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      T s = isph ? p[i] - p[j] : p[i] + p[j];
      const T sv = vol[j] * dfsph_t(g);
      ax += s * (g.dx * sv);
      ay += s * (g.dy * sv);
    }
    acc[2 * i] = -ax / rho[i];
    acc[2 * i + 1] = -ay / rho[i];
  }
}

// updatePressure (v2/modules/dfsph.py:39-50); returns sum of
// clamp(r/rho0, min=-thr) (accumulated in double).
// isph: error |r| (incompressible.py:357) instead of max(r/rho0, -thr).
template <class T>
double sweep_b(const Csr<T>& c, const DomT<T>& d, const T* vol, const T* acc, const T* src, const T* alpha, T dtf,
               T omega, T rho0, T thr, int clampp, T* p, T* resid, int isph = 0) {
  double tot = 0.0;
  set_threads();
#pragma omp parallel for schedule(static) reduction(+ : tot)
  for (int64_t i = 0; i < c.nq; ++i) {
    T sum = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      const T sv = vol[j] * dfsph_t(g);
      sum += (acc[2 * j] - acc[2 * i]) * (g.dx * sv) + (acc[2 * j + 1] - acc[2 * i + 1]) * (g.dy * sv);
    }
    T K = -dtf * sum;
    T s = src[i];
    T res = K - s;
    T pn = p[i] + omega * (s - K) / alpha[i];
    if (clampp) pn = std::max(pn, T(0));
    p[i] = pn;
    if (resid) resid[i] = res;
    tot += isph ? (double)std::fabs(res) : (double)std::max(res / rho0, -thr);
  }
  return tot;
}

// Source terms (v2/modules/dfsph.py:57-65).
template <class T>
void source(const Csr<T>& c, const DomT<T>& d, const T* vel, const T* adv, const T* vol, const T* areas, T dt,
            int density_src, T* src) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    T ax = adv ? adv[2 * i] : T(0), ay = adv ? adv[2 * i + 1] : T(0);
    T vix = vel[2 * i] + dt * ax, viy = vel[2 * i + 1] + dt * ay;
    T dv = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      T bx = adv ? adv[2 * j] : T(0), by = adv ? adv[2 * j + 1] : T(0);
      T vjx = vel[2 * j] + dt * bx, vjy = vel[2 * j + 1] + dt * by;
      const T sv = vol[j] * dfsph_t(g);
      T kx = g.dx * sv, ky = g.dy * sv;
      dv += (vjx - vix) * kx + (vjy - viy) * ky;
    }
    src[i] = density_src ? (T(1) - areas[i] / vol[i]) - dt * dv : dv;
  }
}

template <class T>
Csr<T> make_csr(int64_t nq, const int32_t* off, const int32_t* nbr, const T* qpos, const T* rpos, const T* qh,
                const T* rh, T h0, int mode) {
  Csr<T> c;
  c.nq = nq;
  c.off = off;
  c.nbr = nbr;
  c.qpos = qpos;
  c.rpos = rpos;
  c.qh = qh;
  c.rh = rh;
  c.h0 = h0;
  c.mode = mode;
  return c;
}

template <class T>
int dfsph_solve(const Csr<T>& c, const DomT<T>& d, T dt, T rho0, T omega, T thr, int min_iters, int max_iters,
                int clampp, int density_src, const T* vel, const T* adv, const T* rho, const T* vol, const T* areas,
                const T* alpha, const T* src0, T* p, T* acc, T* resid, int32_t* iters_out, double* err_out,
                double* err0_out, int isph = 0) {
  const int64_t n = c.nq;
  std::vector<T> src(n);
  if (density_src || adv)
    source<T>(c, d, vel, adv, vol, areas, dt, density_src, src.data());
  else
    std::copy(src0, src0 + n, src.begin());
  const T dtf = density_src ? (T)((double)dt * (double)dt) : dt;
  // while i < maxIters and (i < minIters or error > errorThreshold)
  // (v2/modules/dfsph.py:88-103)
  int it = 0;
  double err = 0.0, err0 = 0.0;
  while (it < max_iters && (it < min_iters || err > (double)thr)) {
    sweep_a<T>(c, d, vol, rho, p, acc, isph);
    double tot = sweep_b<T>(c, d, vol, acc, src.data(), alpha, dtf, omega, rho0, thr, clampp, p, resid, isph);
    err = tot / (double)n;
    if (it == 0) err0 = err;
    ++it;
  }
  sweep_a<T>(c, d, vol, rho, p, acc, isph);
  *iters_out = it;
  *err_out = err;
  *err0_out = err0;
  return 0;
}

// ---------------------------------------------------------------------------
// ISPH (v2/modules/incompressible.py, v2/sparse.py).
// ---------------------------------------------------------------------------

// The GPU's fixed float64 reduction: rows in blocks of 256; in a block each
// 64-lane wave adds by the shuffle-down tree (lane l += lane l+o for
// o = 32..1) and the four wave sums are added in order; then one block adds
// the block partials: lane t sums partials t, t+256, ... in order and the
// same block tree combines the lanes.
static double block_tree256(double* v) {
  double t = 0.0;
  for (int w = 0; w < 4; ++w) {
    double* a = v + 64 * w;
    for (int o = 32; o > 0; o >>= 1)
      for (int l = 0; l < o; ++l) a[l] = a[l] + a[l + o];
    t += a[0];
  }
  return t;
}

static double final_sum_tree(const std::vector<double>& part) {
  double lanes[256];
  for (int t = 0; t < 256; ++t) {
    double v = 0.0;
    for (size_t k = t; k < part.size(); k += 256) v += part[k];
    lanes[t] = v;
  }
  return block_tree256(lanes);
}

// sum_i f(i) over n rows with the GPU's tree.
template <class F>
double tree_sum(int64_t n, F&& f) {
  const int64_t nb = (n + 255) / 256;
  std::vector<double> part((size_t)(nb > 0 ? nb : 1), 0.0);
  for (int64_t b = 0; b < nb; ++b) {
    double v[256];
    for (int t = 0; t < 256; ++t) {
      const int64_t i = b * 256 + t;
      v[t] = i < n ? f(i) : 0.0;
    }
    part[b] = block_tree256(v);
  }
  return final_sum_tree(part);
}

// computeLaplacian (incompressible.py:102-133); values in CSR order.
template <class T>
void isph_laplacian(const Csr<T>& c, const DomT<T>& d, const T* mass, const T* rho, T c2dt, T* val, T* aii, T* pre) {
  for (int64_t i = 0; i < c.nq; ++i) {
    T a = 0;
    int64_t self_k = -1;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      const int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      const T gx2 = g.gx * g.gx, gy2 = g.gy * g.gy;
      const T gn = std::sqrt(gx2 + gy2);
      const T quot = g.q * g.h + (T)1e-7 * g.h;
      const T t = (c2dt * (mass[j] / rho[j])) * (gn / quot);
      val[k] = t;
      a += t;
      if (j == i) self_k = k;
    }
    if (self_k >= 0) val[self_k] = -a;
    aii[i] = a;
    pre[i] = T(-1) / a;
  }
}

template <class T>
void spmv_csr(const Csr<T>& c, const T* val, const T* x, T* y) {
  for (int64_t i = 0; i < c.nq; ++i) {
    T acc = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) acc += val[k] * x[c.nbr[k]];
    y[i] = acc;
  }
}

// BiCGStab (sparse.py:141-230) in the GPU's arithmetic: float32 vectors,
// float64 dots and scalars, the scalars rounded to float32 where they
// scale a vector.  Returns the status; *iters the completed iterations.
template <class V>
void spmv_promote(const Csr<float>& c, const float* val, const V* x, V* y) {
  for (int64_t i = 0; i < c.nq; ++i) {
    V acc = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) acc += (V)val[k] * x[c.nbr[k]];
    y[i] = acc;
  }
}

// Vectors of type T (float, or double for solverPrecision = float64); the
// float inputs are promoted and x is rounded back.
template <class T>
int bicgstab(const Csr<float>& c, const float* val, const float* pre_f, const float* b_f, float* x_f, float x0_scale,
             double rtol, double atol, int maxiter, int32_t* iters, double* rnorm_out, double* bnorm_out) {
  const int64_t n = c.nq;
  auto dot = [&](const T* u, const T* w) { return tree_sum(n, [&](int64_t i) { return (double)u[i] * (double)w[i]; }); };
  std::vector<T> r(n), rt(n), p(n), ph(n), v(n), sh(n), t(n), ax(n), bv(n), xv(n), prev(n);
  for (int64_t i = 0; i < n; ++i) {
    bv[i] = (T)b_f[i];
    xv[i] = (T)(x_f[i] * x0_scale);
    prev[i] = (T)pre_f[i];
  }
  const T* b = bv.data();
  const T* pre = prev.data();
  T* x = xv.data();
  struct Store {
    std::vector<T>& xv;
    float* x_f;
    ~Store() {
      for (size_t i = 0; i < xv.size(); ++i) x_f[i] = (float)xv[i];
    }
  } store{xv, x_f};
  spmv_promote<T>(c, val, x, ax.data());
  for (int64_t i = 0; i < n; ++i) r[i] = rt[i] = b[i] - ax[i];
  const double bnorm = std::sqrt(dot(b, b));
  double rr = dot(r.data(), r.data());
  const double atol2 = std::max(atol, rtol * bnorm);
  const double rhotol = 1.1920928955078125e-07 * 1.1920928955078125e-07;
  double rnorm = std::sqrt(rr);
  *bnorm_out = bnorm;
  *iters = 0;
  *rnorm_out = rnorm;
  if (bnorm == 0.0) {
    for (int64_t i = 0; i < n; ++i) x[i] = b[i];
    return 0;
  }
  if (maxiter <= 0) return 2;
  if (rnorm < atol2) return 0;
  double rho = rr, rho_prev = 0.0, alpha = 0.0, omega = 0.0, beta = 0.0;
  if (std::fabs(rho) < rhotol) return -10;
  int it = 0;
  for (;;) {
    for (int64_t i = 0; i < n; ++i) {
      const T pi = it == 0 ? r[i] : ((p[i] - (T)omega * v[i]) * (T)beta) + r[i];
      p[i] = pi;
      ph[i] = pre[i] * pi;
    }
    spmv_promote<T>(c, val, ph.data(), v.data());
    const double rv = dot(rt.data(), v.data());
    if (rv == 0.0) {
      *iters = it;
      return -11;
    }
    alpha = rho / rv;
    for (int64_t i = 0; i < n; ++i) {
      r[i] = r[i] - (T)alpha * v[i];
      sh[i] = pre[i] * r[i];
    }
    const double ss = dot(r.data(), r.data());
    if (std::sqrt(ss) < atol2) {
      for (int64_t i = 0; i < n; ++i) x[i] = x[i] + (T)alpha * ph[i];
      *iters = it;
      *rnorm_out = std::sqrt(ss);
      return 1;
    }
    spmv_promote<T>(c, val, sh.data(), t.data());
    omega = dot(t.data(), r.data()) / dot(t.data(), t.data());
    for (int64_t i = 0; i < n; ++i) {
      T xi = x[i] + (T)alpha * ph[i];
      xi = xi + (T)omega * sh[i];
      x[i] = xi;
      r[i] = r[i] - (T)omega * t[i];
    }
    rr = dot(r.data(), r.data());
    const double rtr = dot(rt.data(), r.data());
    rho_prev = rho;
    ++it;
    *iters = it;
    if (it >= maxiter) return 2;
    rnorm = std::sqrt(rr);
    *rnorm_out = rnorm;
    if (rnorm < atol2) return 0;
    rho = rtr;
    if (std::fabs(rho) < rhotol) return -10;
    if (std::fabs(omega) < rhotol) return -11;
    beta = (rho / rho_prev) * (alpha / omega);
  }
}

// computeDivergenceSourceTerm (incompressible.py:148-165).
template <class T>
void isph_source(const Csr<T>& c, const DomT<T>& d, const T* vel, const T* adv, const T* mass, const T* rho, T dt,
                 T rho0, T* B) {
  const int64_t n = c.nq;
  for (int64_t i = 0; i < n; ++i) {
    const T aix = adv ? adv[2 * i] : T(0), aiy = adv ? adv[2 * i + 1] : T(0);
    const T vix = vel[2 * i] + aix * dt, viy = vel[2 * i + 1] + aiy * dt;
    T acc = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      const int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      const T ajx = adv ? adv[2 * j] : T(0), ajy = adv ? adv[2 * j + 1] : T(0);
      const T vjx = vel[2 * j] + ajx * dt, vjy = vel[2 * j + 1] + ajy * dt;
      const T V = mass[j] / rho[i];
      const T kx = V * g.gx, ky = V * g.gy;
      acc += (vjx - vix) * kx + (vjy - viy) * ky;
    }
    B[i] = rho0 * acc;
  }
  const T meanf = (T)(tree_sum(n, [&](int64_t i) { return (double)B[i]; }) / (double)n);
  for (int64_t i = 0; i < n; ++i) B[i] = B[i] - meanf;
}

// Closing passes of the ISPH step (incompressible.py:398-460).
template <class T>
void isph_close(const Csr<T>& c, const DomT<T>& d, const T* mass, const T* rho, const T* vel_in, const T* adv,
                const T* pstar, const T* pss, T dt, T rho0, T* vprime, T* shift, T* pos, T* vel) {
  const int64_t n = c.nq;
  const T neg_dt2 = (T)(-((double)dt * (double)dt));
  for (int64_t i = 0; i < n; ++i) {
    const T psi = pstar[i], pqi = pss[i], rhoi = rho[i];
    const T iri = T(1) / (rhoi * rhoi);
    T gsx = 0, gsy = 0, gyx = 0, gyy = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      const int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      const T mj = mass[j], rhoj = rho[j];
      const T V = mj / rhoj;
      const T kx = V * g.gx, ky = V * g.gy;
      const T s = psi + pstar[j];
      gsx += s * kx;
      gsy += s * ky;
      const T irj = T(1) / (rhoj * rhoj);
      const T a = pqi * iri + pss[j] * irj;
      gyx += a * (mj * g.gx);
      gyy += a * (mj * g.gy);
    }
    const T ax0 = adv ? adv[2 * i] : T(0), ay0 = adv ? adv[2 * i + 1] : T(0);
    const T vsx = vel_in[2 * i] + ax0 * dt, vsy = vel_in[2 * i + 1] + ay0 * dt;
    const T ax = (-gsx) / rho0, ay = (-gsy) / rho0;
    vprime[2 * i] = vsx + ax * dt;
    vprime[2 * i + 1] = vsy + ay * dt;
    shift[2 * i] = (neg_dt2 * (rhoi * gyx)) / rho0;
    shift[2 * i + 1] = (neg_dt2 * (rhoi * gyy)) / rho0;
  }
  for (int64_t i = 0; i < n; ++i) {
    const T vix = vprime[2 * i], viy = vprime[2 * i + 1];
    T g00 = 0, g01 = 0, g10 = 0, g11 = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      const int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      const T V = mass[j] / rho[j];
      const T kx = V * g.gx, ky = V * g.gy;
      const T dx = vix - vprime[2 * j], dy = viy - vprime[2 * j + 1];
      g00 += dx * kx;
      g01 += dx * ky;
      g10 += dy * kx;
      g11 += dy * ky;
    }
    const T sx = shift[2 * i], sy = shift[2 * i + 1];
    vel[2 * i] = vix + (g00 * sx + g01 * sy);
    vel[2 * i + 1] = viy + (g10 * sx + g11 * sy);
    pos[2 * i] = (pos[2 * i] + vix * dt) + sx;
    pos[2 * i + 1] = (pos[2 * i + 1] + viy * dt) + sy;
  }
}

// ---------------------------------------------------------------------------
// delta-SPH right-hand side (v2/simulationSchemes/deltaPlus.py:176-422).
// ---------------------------------------------------------------------------
inline float atan2_t(float y, float x) { return dsph_pm::atan2f_pm(y, x); }
inline double atan2_t(double y, double x) { return std::atan2(y, x); }
inline void sincos_t(float x, float* s, float* c) { dsph_pm::sincosf_pm(x, s, c); }
inline void sincos_t(double x, double* s, double* c) {
  *s = std::sin(x);
  *c = std::cos(x);
}
template <class T>
T sgn_t(T v) { return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0)); }

// pinv2x2, v2/math.py:10-57, in the working type.
template <class T>
void pinv2(T a, T b, T c, T d, T* L, T* ev) {
  const T a2 = a * a, b2 = b * b, c2 = c * c, d2 = d * d;
  const T theta = T(0.5) * atan2_t(T(2) * a * c + T(2) * b * d, a2 + b2 - c2 - d2);
  T st, ct;
  sincos_t(theta, &st, &ct);
  const T S1 = a2 + b2 + c2 + d2;
  const T e = a2 + b2 - c2 - d2;
  const T f = a * c + b * d;
  const T S2 = std::sqrt(e * e + T(4) * (f * f));
  const T o1 = std::sqrt((S1 + S2) / T(2));
  const T o2 = std::sqrt(std::max(S1 - S2 + (T)1e-7, (T)1e-7) / T(2));
  const T phi = T(0.5) * atan2_t(T(2) * a * b + T(2) * c * d, a2 - b2 + c2 - d2);
  T sp, cp;
  sincos_t(phi, &sp, &cp);
  const T s11 = sgn_t<T>((a * ct + c * st) * cp + (b * ct + d * st) * sp);
  const T s22 = sgn_t<T>((a * st - c * ct) * sp + (-b * st + d * ct) * cp);
  const T v00 = cp * s11, v01 = -sp * s22, v10 = sp * s11, v11 = cp * s22;
  const T i1 = std::fabs(o1) > (T)1e-5 ? T(1) / o1 : T(0);
  const T i2 = std::fabs(o2) > (T)1e-5 ? T(1) / o2 : T(0);
  const T w00 = v00 * i1, w01 = v01 * i2, w10 = v10 * i1, w11 = v11 * i2;
  L[0] = w00 * ct + w01 * (-st);
  L[1] = w00 * st + w01 * ct;
  L[2] = w10 * ct + w11 * (-st);
  L[3] = w10 * st + w11 * ct;
  const bool swap = std::fabs(o2) > std::fabs(o1);
  ev[0] = swap ? o2 : o1;
  ev[1] = swap ? o1 : o2;
}

// Wall particles as seen by the fluid rows (rows = fluid, entries = wall
// indices, rpos = wall positions).
template <class T>
struct BSet {
  Csr<T> c;
  const T* mass;
  const T* rho;
  const T* vel;
  double visc_coeff, eps;
  int viscosity, pi_switch;
};

// computeCovarianceMatrices (normalizationMatrices.py:18-23) over fluid
// (+ wall: 'fluidwBoundary', deltaPlus.py:241), identity for rows with < 4
// fluid neighbors (deltaPlus.py:245), pinv2x2, then
// renormalizedDensityGradient (densityDiffusion.py:7-26) + the plain wall
// gradient (densityGradient, deltaPlus.py:262-263).  One pass over each row:
// L_i is the same for every pair of row i, so
// sum_j V_j (rho_i - rho_j) L_i gradW_ij = L_i sum_j V_j (rho_i - rho_j) gradW_ij
// (the sums are accumulated with M and L applied once; a few ulp from the
// reference's per-pair product).
template <class T>
void dsph_normalization(const Csr<T>& c, const DomT<T>& d, const T* mass, const T* rho, T* Lout, T* evout,
                        T* grad_rho, const BSet<T>* B) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    T m00 = 0, m01 = 0, m10 = 0, m11 = 0, sx = 0, sy = 0;
    const T rho_i = rho[i];
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      T Vj = mass[j] / rho[j];
      T s = -(g.q * g.h);
      T qx = s * g.ex, qy = s * g.ey;
      T kx = Vj * g.gx, ky = Vj * g.gy;
      m00 += qx * kx;
      m01 += qx * ky;
      m10 += qy * kx;
      m11 += qy * ky;
      T dr = rho_i - rho[j];
      sx += dr * kx;
      sy += dr * ky;
    }
    T bx = 0, by = 0;
    if (B) {
      T b00 = 0, b01 = 0, b10 = 0, b11 = 0;
      for (int k = B->c.off[i]; k < B->c.off[i + 1]; ++k) {
        int32_t j = B->c.nbr[k];
        G<T> g = pg(B->c, i, j, d);
        T Vj = B->mass[j] / B->rho[j];
        T s = -(g.q * g.h);
        T qx = s * g.ex, qy = s * g.ey;
        T kx = Vj * g.gx, ky = Vj * g.gy;
        b00 += qx * kx;
        b01 += qx * ky;
        b10 += qy * kx;
        b11 += qy * ky;
        T dr = rho_i - B->rho[j];
        bx += dr * kx;
        by += dr * ky;
      }
      m00 = m00 + b00;
      m01 = m01 + b01;
      m10 = m10 + b10;
      m11 = m11 + b11;
    }
    if (c.off[i + 1] - c.off[i] < 4) {
      m00 = 1;
      m01 = 0;
      m10 = 0;
      m11 = 1;
    }
    T L[4], ev[2];
    pinv2<T>(m00, m01, m10, m11, L, ev);
    if (Lout)
      for (int q = 0; q < 4; ++q) Lout[4 * i + q] = L[q];
    if (evout) {
      evout[2 * i] = ev[0];
      evout[2 * i + 1] = ev[1];
    }
    T ax = L[0] * sx + L[1] * sy;
    T ay = L[2] * sx + L[3] * sy;
    if (B) {
      ax = ax + bx;
      ay = ay + by;
    }
    grad_rho[2 * i] = ax;
    grad_rho[2 * i + 1] = ay;
  }
}

struct DeltaP {
  double rho0, cs, delta, kscale, visc_coeff, eps, kappa, gx, gy;
  int viscosity, pi_switch, eos, diffusion;
};

// Fused delta-SPH terms: densityDiffusion.py:49-73 (deltaSPH scheme),
// viscosity.py:123-170 / 200-218, momentumEquation.py:6-39,
// pressureEOS.py:28-64, pressureForce.py:10-12, combined as
// deltaPlus.py:349-350.
template <class T>
void dsph_rhs(const Csr<T>& c, const DomT<T>& d, const DeltaP& P, const T* vel, const T* mass, const T* rho,
              const T* grad_rho, T* dudt, T* drhodt, T* terms, const BSet<T>* B) {
  const T rho0 = (T)P.rho0;
  const T eos_c = P.eos == 0 ? (T)P.cs * (T)P.cs : (T)P.kappa;
  const T cs = (T)P.cs;
  const T visc_c = P.viscosity == 1 ? (T)(2.0 * 4.0 * P.visc_coeff * P.rho0) : (T)P.visc_coeff;
  const T eps = (T)P.eps;
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    const T hi = c.qh ? c.qh[i] : c.h0;
    const T uix = vel[2 * i], uiy = vel[2 * i + 1];
    const T rho_i = rho[i];
    const T grix = P.diffusion ? grad_rho[2 * i] : T(0), griy = P.diffusion ? grad_rho[2 * i + 1] : T(0);
    const T p_i = eos_c * (rho_i - rho0);
    T accD = 0, accVx = 0, accVy = 0, accC = 0, accPx = 0, accPy = 0;
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      const T ujx = vel[2 * j], ujy = vel[2 * j + 1];
      const T rho_j = rho[j];
      const T m_j = mass[j];
      const T Vj = m_j / rho_j;
      const T kx = Vj * g.gx, ky = Vj * g.gy;
      const T rij = g.q * g.h;
      if (P.diffusion) {
        const T r_ij = (T(2) * (rho_j - rho_i)) / (rij + (T)1e-6 * g.h);
        const T psx = (-r_ij) * g.ex + (grix + grad_rho[2 * j]);
        const T psy = (-r_ij) * g.ey + (griy + grad_rho[2 * j + 1]);
        accD += psx * kx + psy * ky;
      }
      if (P.viscosity >= 0) {
        const T vx = uix - ujx, vy = uiy - ujy;
        const T vr = vx * g.ex + vy * g.ey;
        T pi = vr / (rij + eps * (g.h * g.h));
        if (P.pi_switch) pi = vr < T(0) ? pi : T(0);
        const T vs = Vj * pi;
        accVx += vs * g.gx;
        accVy += vs * g.gy;
      }
      {
        const T qx = (ujx - uix) * m_j, qy = (ujy - uiy) * m_j;
        T kq = qx * g.gx + qy * g.gy;
        kq = (kq * rho_i) / rho_j;
        accC += kq;
      }
      {
        const T ps = p_i + eos_c * (rho_j - rho0);
        accPx += ps * kx;
        accPy += ps * ky;
      }
    }
    T apx = (-accPx) / rho_i, apy = (-accPy) / rho_i;
    T vsx = 0, vsy = 0;
    if (P.viscosity == 1) {
      const T s = visc_c / rho_i;
      vsx = s * accVx;
      vsy = s * accVy;
    } else if (P.viscosity == 0) {
      const T s = (((visc_c * hi) * cs) * rho0) / rho_i;
      vsx = s * accVx;
      vsy = s * accVy;
    }
    T cont = -accC;
    if (B) {
      T bVx = 0, bVy = 0, bC = 0, bPx = 0, bPy = 0;
      for (int k = B->c.off[i]; k < B->c.off[i + 1]; ++k) {
        int32_t j = B->c.nbr[k];
        G<T> g = pg(B->c, i, j, d);
        const T ujx = B->vel[2 * j], ujy = B->vel[2 * j + 1];
        const T rho_j = B->rho[j];
        const T m_j = B->mass[j];
        const T rij = g.q * g.h;
        if (B->viscosity >= 0) {
          const T vx = uix - ujx, vy = uiy - ujy;
          const T vr = vx * g.ex + vy * g.ey;
          T pi = vr / (rij + (T)B->eps * (g.h * g.h));
          if (B->pi_switch) pi = vr < T(0) ? pi : T(0);
          const T Vb = B->viscosity == 0 ? m_j / (rho_i + rho_j) : m_j / rho_j;
          const T vs = Vb * pi;
          bVx += vs * g.gx;
          bVy += vs * g.gy;
        }
        {
          const T qx = (ujx - uix) * m_j, qy = (ujy - uiy) * m_j;
          T kq = qx * g.gx + qy * g.gy;
          kq = (kq * rho_i) / rho_j;
          bC += kq;
        }
        {
          const T Vj = m_j / rho_j;
          const T ps = p_i + eos_c * (rho_j - rho0);
          bPx += ps * (Vj * g.gx);
          bPy += ps * (Vj * g.gy);
        }
      }
      apx = apx + (-bPx) / rho_i;
      apy = apy + (-bPy) / rho_i;
      cont = cont + (-bC);
      if (B->viscosity == 1) {
        const T s = (T)(2.0 * 4.0 * B->visc_coeff * P.rho0) / rho_i;
        vsx = vsx + s * bVx;
        vsy = vsy + s * bVy;
      } else if (B->viscosity == 0) {
        const T s = ((((T)B->visc_coeff * hi) * cs) * rho0) / rho_i;
        vsx = vsx + s * bVx;
        vsy = vsy + s * bVy;
      }
    }
    const T D = P.diffusion ? ((((T)P.delta * hi) / (T)P.kscale) * cs) * accD : T(0);
    dudt[2 * i] = (apx + (T)P.gx) + vsx;
    dudt[2 * i + 1] = (apy + (T)P.gy) + vsy;
    drhodt[i] = cont + D;
    if (terms) {
      T* t = terms + 8 * i;
      t[0] = apx;
      t[1] = apy;
      t[2] = vsx;
      t[3] = vsy;
      t[4] = cont;
      t[5] = D;
      t[6] = p_i;
      t[7] = 0;
    }
  }
}

// 3x3 determinant by cofactors of the first row, as mDBC.h:200-209 writes it.
template <class T>
T det3(T m11, T m12, T m13, T m21, T m22, T m23, T m31, T m32, T m33) {
  const T a = m11 * (m22 * m33 - m23 * m32);
  const T b = m12 * (m21 * m33 - m23 * m31);
  const T c = m13 * (m21 * m32 - m22 * m31);
  return a - b + c;
}

// The reference's C++ mDBC op (mDBCDensityKernel, v2/cppSrc/mDBC.h:129-244):
// the same sums as mDBCDensity with x_gj = -(q h) xhat and b = sum m_j
// [W, gradW]; Cramer's rule with det + 1e-7; Shepard density for <= 5
// neighbors; the result clamped from below at rho0.
template <class T>
T mdbc_op_row(const Csr<T>& c, const DomT<T>& d, int64_t i, const T* mass, const T* rho, const T* bpos, T rho0,
              T* shep_out) {
  T nom = 0, den = 0;
  T m00 = 0, m01 = 0, m02 = 0, m10 = 0, m11 = 0, m12 = 0, m20 = 0, m21 = 0, m22 = 0, bx = 0, by = 0, bz = 0;
  const int cnt = c.off[i + 1] - c.off[i];
  for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
    int32_t j = c.nbr[k];
    G<T> g = pg(c, i, j, d);
    const T m_j = mass[j], rho_j = rho[j];
    const T V = m_j / rho_j;
    const T xx = ((-g.q) * g.h) * g.ex, xy = ((-g.q) * g.h) * g.ey;
    nom += m_j * g.w;
    den += V * g.w;
    const T vw = V * g.w, vgx = V * g.gx, vgy = V * g.gy;
    m00 += vw;
    m10 += vgx;
    m20 += vgy;
    m01 += vw * xx;
    m11 += vgx * xx;
    m21 += vgx * xy;
    m02 += vw * xy;
    m12 += vgy * xx;
    m22 += vgy * xy;
    bx += m_j * g.w;
    by += m_j * g.gx;
    bz += m_j * g.gy;
  }
  const T shep = nom / (den + (T)1e-7);
  if (shep_out) shep_out[i] = shep;
  if (cnt <= 5) return shep < rho0 ? rho0 : shep;
  const T d0 = det3(m00, m01, m02, m10, m11, m12, m20, m21, m22) + (T)1e-7;
  const T rx = det3(bx, m01, m02, by, m11, m12, bz, m21, m22) / d0;
  const T ry = det3(m00, bx, m02, m10, by, m12, m20, bz, m22) / d0;
  const T rz = det3(m00, m01, bx, m10, m11, by, m20, m21, bz) / d0;
  const T px = bpos[2 * i] - c.qpos[2 * i], py = bpos[2 * i + 1] - c.qpos[2 * i + 1];
  const T out = rx + (px * ry + py * rz);
  return out < rho0 ? rho0 : out;
}

// mDBCDensity (v2/modules/mDBC.py:50-136); rows of `c` are ghosts.
// variant 1: the C++ op (mdbc_op_row).
template <class T>
void mdbc(const Csr<T>& c, const DomT<T>& d, const T* mass, const T* rho, const T* bpos, T rho0, T support,
          T* rho_b, T* shep_out, int variant) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    if (variant == 1) {
      rho_b[i] = mdbc_op_row(c, d, i, mass, rho, bpos, rho0, shep_out);
      continue;
    }
    T nom = 0, vol = 0, gsx = 0, gsy = 0, vgx = 0, vgy = 0, psx = 0, psy = 0, p00 = 0, p01 = 0, p10 = 0, p11 = 0;
    const int cnt = c.off[i + 1] - c.off[i];
    for (int k = c.off[i]; k < c.off[i + 1]; ++k) {
      int32_t j = c.nbr[k];
      G<T> g = pg(c, i, j, d);
      const T m_j = mass[j], rho_j = rho[j];
      const T Vj = m_j / rho_j;
      nom += m_j * g.w;
      const T kw = Vj * g.w;
      vol += T(1) * kw;
      const T kx = Vj * g.gx, ky = Vj * g.gy;
      gsx += rho_j * kx;
      gsy += rho_j * ky;
      vgx += T(1) * kx;
      vgy += T(1) * ky;
      const T xx = ((-g.ex) * g.q) * support, xy = ((-g.ey) * g.q) * support;
      psx += xx * kw;
      psy += xy * kw;
      p00 += xx * kx;
      p01 += xx * ky;
      p10 += xy * kx;
      p11 += xy * ky;
    }
    const T shep = nom / (vol + (T)1e-7);
    T out = rho0;
    if (cnt > 0) out = shep;
    if (cnt > 5) {
      const double A[9] = {(double)vol, (double)psx, (double)psy, (double)vgx, (double)p00,
                           (double)p01, (double)vgy, (double)p10, (double)p11};
      double Ai[9];
      dsph_pm::pinv3x3(A, 3.0 * 1.1920928955078125e-07, Ai);
      const double b0 = (double)nom, b1 = (double)gsx, b2 = (double)gsy;
      const T c0 = (T)(Ai[0] * b0 + Ai[1] * b1 + Ai[2] * b2);
      const T c1 = (T)(Ai[3] * b0 + Ai[4] * b1 + Ai[5] * b2);
      const T c2 = (T)(Ai[6] * b0 + Ai[7] * b1 + Ai[8] * b2);
      const T rx = bpos[2 * i] - c.qpos[2 * i], ry = bpos[2 * i + 1] - c.qpos[2 * i + 1];
      out = c0 + (rx * c1 + ry * c2);
    }
    rho_b[i] = out;
    if (shep_out) shep_out[i] = shep;
  }
}

}  // namespace

extern "C" {

#define ORC_CSR(T) make_csr<T>(nq, off, nbr, qpos, rpos, qh, rh, h0, mode)

int orc_density_f32(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr, const float* qpos,
                    const float* rpos, const float* qh, const float* rh, float h0, int32_t mode, const float* rmass,
                    const float* qmass, float* rho, float* vol) {
  density<float>(ORC_CSR(float), DomT<float>(*dom), rmass, qmass, rho, vol);
  return 0;
}
int orc_density_f64(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr, const double* qpos,
                    const double* rpos, const double* qh, const double* rh, double h0, int32_t mode,
                    const double* rmass, const double* qmass, double* rho, double* vol) {
  density<double>(ORC_CSR(double), DomT<double>(*dom), rmass, qmass, rho, vol);
  return 0;
}

int orc_sph_operation_f32(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr,
                          const float* qpos, const float* rpos, const float* qh, const float* rh, float h0,
                          int32_t mode, int32_t op, int32_t gm, const float* qm, const float* rm, const float* qr,
                          const float* rr, const float* qq, const float* rq, int32_t width, float* out) {
  sph_operation<float>(ORC_CSR(float), DomT<float>(*dom), op, gm, qm, rm, qr, rr, qq, rq, width, out);
  return 0;
}

int orc_dfsph_prepare_f32(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr,
                          const float* qpos, const float* qh, float h0, int32_t mode, const float* vel,
                          const float* areas, const float* vol, const float* rho, float dt, float rho0,
                          int32_t visc_on, int32_t visc_scheme, float v_alpha, float v_cs, float v_rho0, float v_eps,
                          int32_t pi_switch, float* alpha_div, float* alpha_den, float* div_src, float* visc) {
  ViscP vp{visc_on, visc_scheme, pi_switch, v_alpha, v_cs, v_rho0, v_eps};
  const float* rpos = qpos;
  const float* rh = qh;
  dfsph_prepare<float>(ORC_CSR(float), DomT<float>(*dom), vel, areas, vol, rho, dt, rho0, vp, alpha_div, alpha_den,
                       div_src, visc);
  return 0;
}
int orc_dfsph_prepare_f64(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr,
                          const double* qpos, const double* qh, double h0, int32_t mode, const double* vel,
                          const double* areas, const double* vol, const double* rho, double dt, double rho0,
                          int32_t visc_on, int32_t visc_scheme, double v_alpha, double v_cs, double v_rho0,
                          double v_eps, int32_t pi_switch, double* alpha_div, double* alpha_den, double* div_src,
                          double* visc) {
  ViscP vp{visc_on, visc_scheme, pi_switch, v_alpha, v_cs, v_rho0, v_eps};
  const double* rpos = qpos;
  const double* rh = qh;
  dfsph_prepare<double>(ORC_CSR(double), DomT<double>(*dom), vel, areas, vol, rho, dt, rho0, vp, alpha_div,
                        alpha_den, div_src, visc);
  return 0;
}

int orc_dfsph_solve_f32(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr, const float* qpos,
                        const float* qh, float h0, int32_t mode, float dt, float rho0, float omega, float thr,
                        int32_t min_iters, int32_t max_iters, int32_t clampp, int32_t density_src, const float* vel,
                        const float* adv, const float* rho, const float* vol, const float* areas, const float* alpha,
                        const float* src0, float* p, float* acc, float* resid, int32_t* iters, double* err,
                        double* err0, int32_t isph) {
  const float* rpos = qpos;
  const float* rh = qh;
  return dfsph_solve<float>(ORC_CSR(float), DomT<float>(*dom), dt, rho0, omega, thr, min_iters, max_iters, clampp,
                            density_src, vel, adv, rho, vol, areas, alpha, src0, p, acc, resid, iters, err, err0,
                            isph);
}
int orc_dfsph_solve_f64(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr,
                        const double* qpos, const double* qh, double h0, int32_t mode, double dt, double rho0,
                        double omega, double thr, int32_t min_iters, int32_t max_iters, int32_t clampp,
                        int32_t density_src, const double* vel, const double* adv, const double* rho,
                        const double* vol, const double* areas, const double* alpha, const double* src0, double* p,
                        double* acc, double* resid, int32_t* iters, double* err, double* err0, int32_t isph) {
  const double* rpos = qpos;
  const double* rh = qh;
  return dfsph_solve<double>(ORC_CSR(double), DomT<double>(*dom), dt, rho0, omega, thr, min_iters, max_iters,
                             clampp, density_src, vel, adv, rho, vol, areas, alpha, src0, p, acc, resid, iters, err,
                             err0, isph);
}

// Wendland C2 values for a list of (q, h): W and dW/dr (kernelFunctions/
// Wendland2.py:30-36), in float64.
int orc_wendland(const double* q, const double* h, int64_t n, double* w, double* dw) {
  for (int64_t i = 0; i < n; ++i) {
    w[i] = kernelW<double>(q[i], h[i]);
    dw[i] = kernelDW<double>(q[i], h[i]);
  }
  return 0;
}

// Semi-implicit Euler, v2/modules/integration.py:111-138; dudt summed as
// v2/simulationSchemes/dfsph.py:209 (a_p + g + visc + a_div).
int orc_integrate_sie_f32(float* pos, float* vel, const float* a0, const float* a1, const float* a2, const float* a3,
                          const float* g, int64_t n, float dt) {
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 2; ++k) {
      float d = a0 ? a0[2 * i + k] : 0.0f;
      if (g) d = d + g[k];
      if (a1) d = d + a1[2 * i + k];
      if (a2) d = d + a2[2 * i + k];
      if (a3) d = d + a3[2 * i + k];
      float v = vel[2 * i + k] + d * dt;
      vel[2 * i + k] = v;
      pos[2 * i + k] = pos[2 * i + k] + v * dt;
    }
  }
  return 0;
}

// delta-SPH exports (same-set neighbor lists: rpos = qpos, rh = qh); the
// optional wall set: CSR (b_off over the fluid rows, b_nbr wall indices),
// wall positions, mass, density, velocity, viscosity (coefficient, eps,
// scheme, pi-switch) and the wall support.
struct orc_boundary {
  const int32_t* off;
  const int32_t* nbr;
  const void* pos;
  const void* mass;
  const void* rho;
  const void* vel;
  double visc_coeff, eps, h0;
  int32_t viscosity, pi_switch;
};

}  // extern "C"

namespace {
template <class T>
BSet<T> make_bset(const orc_boundary* b, int64_t nq, const T* qpos, const T* qh, T h0q) {
  BSet<T> B;
  B.c = make_csr<T>(nq, b->off, b->nbr, qpos, static_cast<const T*>(b->pos), qh, nullptr, (T)b->h0, 0);
  (void)h0q;
  B.mass = static_cast<const T*>(b->mass);
  B.rho = static_cast<const T*>(b->rho);
  B.vel = static_cast<const T*>(b->vel);
  B.visc_coeff = b->visc_coeff;
  B.eps = b->eps;
  B.viscosity = b->viscosity;
  B.pi_switch = b->pi_switch;
  return B;
}
}  // namespace

extern "C" {

#define ORC_NORM(T, SUF)                                                                                         \
  int orc_deltasph_normalization_##SUF(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr,  \
                                       const T* qpos, const T* qh, T h0, int32_t mode, const T* mass, const T* rho, \
                                       T* L, T* ev, T* grad_rho, const orc_boundary* bnd) {                     \
    const T* rpos = qpos;                                                                                        \
    const T* rh = qh;                                                                                            \
    BSet<T> B;                                                                                                   \
    if (bnd) B = make_bset<T>(bnd, nq, qpos, qh, h0);                                                            \
    dsph_normalization<T>(ORC_CSR(T), DomT<T>(*dom), mass, rho, L, ev, grad_rho, bnd ? &B : nullptr);           \
    return 0;                                                                                                    \
  }
ORC_NORM(float, f32)
ORC_NORM(double, f64)

// prm: rho0, cs, delta, kernel_scale, visc_coeff, eps, kappa, gx, gy;
// iprm: viscosity, pi_switch, eos, diffusion.
#define ORC_RHS(T, SUF)                                                                                          \
  int orc_deltasph_rhs_##SUF(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr,          \
                             const T* qpos, const T* qh, T h0, int32_t mode, const double* prm, const int32_t* iprm, \
                             const T* vel, const T* mass, const T* rho, const T* grad_rho, T* dudt, T* drhodt,    \
                             T* terms, const orc_boundary* bnd) {                                                \
    const T* rpos = qpos;                                                                                        \
    const T* rh = qh;                                                                                            \
    DeltaP P{prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], prm[7], prm[8], iprm[0], iprm[1], iprm[2],   \
             iprm[3]};                                                                                           \
    BSet<T> B;                                                                                                   \
    if (bnd) B = make_bset<T>(bnd, nq, qpos, qh, h0);                                                            \
    dsph_rhs<T>(ORC_CSR(T), DomT<T>(*dom), P, vel, mass, rho, grad_rho, dudt, drhodt, terms, bnd ? &B : nullptr); \
    return 0;                                                                                                    \
  }
ORC_RHS(float, f32)
ORC_RHS(double, f64)

// mDBC: rows of (off, nbr) are the ghosts at qpos; entries fluid indices at rpos.
#define ORC_MDBC(T, SUF)                                                                                         \
  int orc_mdbc_##SUF(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr, const T* qpos,   \
                     const T* rpos, T h0, const T* mass, const T* rho, const T* bpos, T rho0, T support, T* rho_b, \
                     T* shep, int32_t variant) {                                                                 \
    const T* qh = nullptr;                                                                                       \
    const T* rh = nullptr;                                                                                       \
    int32_t mode = 0;                                                                                            \
    mdbc<T>(ORC_CSR(T), DomT<T>(*dom), mass, rho, bpos, rho0, support, rho_b, shep, variant);                     \
    return 0;                                                                                                    \
  }
ORC_MDBC(float, f32)
ORC_MDBC(double, f64)

// ISPH: rows of (off, nbr) over one particle set (qpos = rpos).
int orc_isph_laplacian_f32(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr,
                           const float* qpos, float h0, const float* mass, const float* rho, float c2dt, float* val,
                           float* aii, float* pre) {
  const float* rpos = qpos;
  const float* qh = nullptr;
  const float* rh = nullptr;
  int32_t mode = 0;
  isph_laplacian<float>(ORC_CSR(float), DomT<float>(*dom), mass, rho, c2dt, val, aii, pre);
  return 0;
}

int orc_spmv_f32(int64_t nq, const int32_t* off, const int32_t* nbr, const float* val, const float* x, float* y) {
  Csr<float> c = make_csr<float>(nq, off, nbr, nullptr, nullptr, nullptr, nullptr, 0.f, 0);
  spmv_csr<float>(c, val, x, y);
  return 0;
}

int orc_bicgstab_f32(int64_t nq, const int32_t* off, const int32_t* nbr, const float* val, const float* pre,
                     const float* b, float* x, float x0_scale, double rtol, double atol, int32_t maxiter,
                     int32_t* iters, double* rnorm, double* bnorm, int32_t fp64) {
  Csr<float> c = make_csr<float>(nq, off, nbr, nullptr, nullptr, nullptr, nullptr, 0.f, 0);
  if (fp64) return bicgstab<double>(c, val, pre, b, x, x0_scale, rtol, atol, maxiter, iters, rnorm, bnorm);
  return bicgstab<float>(c, val, pre, b, x, x0_scale, rtol, atol, maxiter, iters, rnorm, bnorm);
}

int orc_isph_source_f32(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr, const float* qpos,
                        float h0, const float* vel, const float* adv, const float* mass, const float* rho, float dt,
                        float rho0, float* B) {
  const float* rpos = qpos;
  const float* qh = nullptr;
  const float* rh = nullptr;
  int32_t mode = 0;
  isph_source<float>(ORC_CSR(float), DomT<float>(*dom), vel, adv, mass, rho, dt, rho0, B);
  return 0;
}

int orc_isph_close_f32(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr, const float* qpos,
                       float h0, const float* mass, const float* rho, const float* vel_in, const float* adv,
                       const float* pstar, const float* pss, float dt, float rho0, float* vprime, float* shift,
                       float* pos, float* vel) {
  const float* rpos = qpos;
  const float* qh = nullptr;
  const float* rh = nullptr;
  int32_t mode = 0;
  isph_close<float>(ORC_CSR(float), DomT<float>(*dom), mass, rho, vel_in, adv, pstar, pss, dt, rho0, vprime, shift,
                    pos, vel);
  return 0;
}

// deltaPlusShifting (v2/modules/shifting.py:119-147) applied as
// solveShifting (shifting.py:880-1018, scheme 'deltaSPH', one iteration, no
// free surface) and runner.py:40-42.  prm: dx, support, kernel_scale, R, n,
// CFL, cs, threshold, mach (< 0: max |vel| / cs).  pos is updated in place
// (row order), shift_out (may be NULL) receives the added displacement.
}  // extern "C"

namespace {
template <class T>
T shift_pow(T k, double n) {
  if (n >= 0.0 && n <= 16.0 && n == std::floor(n)) {
    double p = 1.0;
    for (int t = 0; t < (int)n; ++t) p *= (double)k;
    return (T)p;
  }
  return (T)std::pow(k, (T)n);
}

template <class T>
void deltaplus_shift(const Csr<T>& c, const DomT<T>& d, const double* prm, const T* mass, const T* rho,
                     const T* vel, const BSet<T>* B, T* pos, T* shift_out) {
  const T q0 = (T)(prm[0] / prm[1]);
  const T h0 = (T)prm[1];
  const T w0 = kernelW_exact<T>(q0, h0);
  const T kscale = (T)prm[2], R = (T)prm[3], cfl = (T)prm[5], cs = (T)prm[6];
  const T thr = (T)(prm[7] * prm[0]);
  T ma = (T)prm[8];
  if (prm[8] < 0.0) {
    T m = T(0);
    for (int64_t i = 0; i < c.nq; ++i) {
      T a = vel[2 * i] * vel[2 * i];
      T b = vel[2 * i + 1] * vel[2 * i + 1];
      T s = std::sqrt(a + b);
      m = s > m ? s : m;
    }
    ma = m / cs;
  }
  set_threads();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nq; ++i) {
    T sx = T(0), sy = T(0);
    for (int32_t k = c.off[i]; k < c.off[i + 1]; ++k) {
      const int32_t j = c.nbr[k];
      const G<T> g = pg<T>(c, i, j, d);
      const T term = T(1) + R * shift_pow<T>(g.w / w0, prm[4]);
      const T sc = term * (mass[j] / (rho[i] + rho[j]));
      sx += sc * g.gx;
      sy += sc * g.gy;
    }
    if (B) {
      for (int32_t k = B->c.off[i]; k < B->c.off[i + 1]; ++k) {
        const int32_t b = B->c.nbr[k];
        const G<T> g = pg<T>(B->c, i, b, d);
        const T term = T(1) + R * shift_pow<T>(g.w / w0, prm[4]);
        const T sc = term * (B->mass[b] / (rho[i] + B->rho[b]));
        sx += sc * g.gx;
        sy += sc * g.gy;
      }
    }
    const T hi = c.qh ? c.qh[i] : c.h0;
    const T hk = (hi / kscale) * T(2);
    const T scale = (-cfl * ma) * (hk * hk);
    T u[2] = {-(scale * sx), -(scale * sy)};
    for (int k = 0; k < 2; ++k) {
      u[k] = u[k] < -thr ? -thr : (u[k] > thr ? thr : u[k]);
      const T x = pos[2 * i + k];
      const T dxk = (x - u[k]) - x;
      pos[2 * i + k] = x + dxk;
      if (shift_out) shift_out[2 * i + k] = dxk;
    }
  }
}
}  // namespace

extern "C" {

#define ORC_SHIFT(T, SUF)                                                                                         \
  int orc_deltaplus_shift_##SUF(const orc_domain* dom, int64_t nq, const int32_t* off, const int32_t* nbr,         \
                                const T* qpos, T h0, const double* prm, const T* mass, const T* rho, const T* vel, \
                                const orc_boundary* bnd, T* pos, T* shift_out) {                                  \
    const T* rpos = qpos;                                                                                         \
    const T* qh = nullptr;                                                                                        \
    const T* rh = nullptr;                                                                                        \
    int32_t mode = 0;                                                                                             \
    BSet<T> B;                                                                                                    \
    if (bnd) B = make_bset<T>(bnd, nq, qpos, qh, h0);                                                             \
    deltaplus_shift<T>(ORC_CSR(T), DomT<T>(*dom), prm, mass, rho, vel, bnd ? &B : nullptr, pos, shift_out);      \
    return 0;                                                                                                     \
  }
ORC_SHIFT(float, f32)
ORC_SHIFT(double, f64)

// The 2x2 pseudo-inverse on its own ([n,4] row-major in, L [n,4], ev [n,2]).
int orc_pinv2x2_f32(const float* M, int64_t n, float* L, float* ev) {
  for (int64_t i = 0; i < n; ++i) pinv2<float>(M[4 * i], M[4 * i + 1], M[4 * i + 2], M[4 * i + 3], L + 4 * i, ev + 2 * i);
  return 0;
}
int orc_pinv2x2_f64(const double* M, int64_t n, double* L, double* ev) {
  for (int64_t i = 0; i < n; ++i)
    pinv2<double>(M[4 * i], M[4 * i + 1], M[4 * i + 2], M[4 * i + 3], L + 4 * i, ev + 2 * i);
  return 0;
}
// ---------------------------------------------------------------------------
// Drop-in ops on the reference's per-pair arrays, as the reference's CPU
// branches run them: one loop over rows, each row's CSR segment in order.
// ---------------------------------------------------------------------------

// viscosityFn, partiBench/viscosityKernel.h:98-145 (CPU branch
// partiBench/viscosityKernel.cpp:71-77).
int orc_op_viscosity(const float* vel_i, const float* vel_j, const float* m_j, const float* rho_i, const float* rho_j,
                     const float* h_i, const int64_t* idx_j, const float* supports, const float* grads,
                     const float* dists, const float* vecs, int32_t n, const int32_t* nn, const int32_t* off, float cs,
                     float rho0, float alpha, float eps, int32_t pi_switch, float* out) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    float res_x = 0.f, res_y = 0.f;
    const float vel_x = vel_i[2 * i], vel_y = vel_i[2 * i + 1];
    const float dens_i = rho_i[i];
    for (int32_t t = 0; t < nn[i]; ++t) {
      const int64_t k = (int64_t)off[i] + t;
      const int32_t j = (int32_t)idx_j[k];
      const float hh = supports[k];
      const float v_x = vel_x - vel_j[2 * j];
      const float v_y = vel_y - vel_j[2 * j + 1];
      const float x_x = vecs[2 * k];
      const float x_y = vecs[2 * k + 1];
      const float r = dists[k] * hh;
      const float vr = v_x * x_x + v_y * x_y;
      float pij = vr / (r + eps + hh * hh);
      if (pi_switch) pij = vr < 0 ? pij : 0;
      const float Vj = m_j[j] / (dens_i + rho_j[j]);
      res_x += Vj * pij * grads[2 * k];
      res_y += Vj * pij * grads[2 * k + 1];
    }
    const float factor = alpha * h_i[i] * cs * rho0 / dens_i;
    out[2 * i] = res_x * factor;
    out[2 * i + 1] = res_y * factor;
  }
  return 0;
}

// mDBCDensityKernel, v2/cppSrc/mDBC.h:129-244 (CPU branch v2/cppSrc/mDBC.cpp:
// 74-80): moment sums in the reference's fill order, Cramer's rule.
static float det3_ref(float m11, float m12, float m13, float m21, float m22, float m23, float m31, float m32,
                      float m33) {
  float a = m11 * (m22 * m33 - m23 * m32);
  float b = m12 * (m21 * m33 - m23 * m31);
  float c = m13 * (m21 * m32 - m22 * m31);
  return a - b + c;
}
int orc_op_mdbc(const float* m_j, const float* rho_j, const float* pos_b, const float* pos_g, int32_t nb,
                const int64_t* idx_j, const float* W, const float* grads, const float* dists, const float* vecs,
                const float* supports, const int32_t* nn, const int32_t* off, float rho0, float* out_mdbc,
                float* out_shepard) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nb; ++i) {
    float sn = 0.f, sd = 0.f;
    float A[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    float b[3] = {0.f, 0.f, 0.f};
    const int32_t cnt = nn[i];
    for (int32_t t = 0; t < cnt; ++t) {
      const int64_t k = (int64_t)off[i] + t;
      const int32_t j = (int32_t)idx_j[k];
      const float mj = m_j[j], rj = rho_j[j], w = W[k];
      const float gx = grads[2 * k], gy = grads[2 * k + 1];
      const float r = dists[k], hh = supports[k];
      const float xx = -r * hh * vecs[2 * k], xy = -r * hh * vecs[2 * k + 1];
      sn += mj * w;
      sd += mj / rj * w;
      A[0][0] += mj / rj * w;
      A[1][0] += mj / rj * gx;
      A[2][0] += mj / rj * gy;
      A[0][1] += mj / rj * w * xx;
      A[1][1] += mj / rj * gx * xx;
      A[2][1] += mj / rj * gx * xy;
      A[0][2] += mj / rj * w * xy;
      A[1][2] += mj / rj * gy * xx;
      A[2][2] += mj / rj * gy * xy;
      b[0] += mj * w;
      b[1] += mj * gx;
      b[2] += mj * gy;
    }
    const float shep = sn / (sd + 1e-7f);
    out_shepard[i] = shep;
    if (cnt <= 5) {
      out_mdbc[i] = std::max(shep, rho0);
      continue;
    }
    const float d0 = det3_ref(A[0][0], A[0][1], A[0][2], A[1][0], A[1][1], A[1][2], A[2][0], A[2][1], A[2][2]);
    const float c0 = det3_ref(b[0], A[0][1], A[0][2], b[1], A[1][1], A[1][2], b[2], A[2][1], A[2][2]) / (d0 + 1e-7f);
    const float c1 = det3_ref(A[0][0], b[0], A[0][2], A[1][0], b[1], A[1][2], A[2][0], b[2], A[2][2]) / (d0 + 1e-7f);
    const float c2 = det3_ref(A[0][0], A[0][1], b[0], A[1][0], A[1][1], b[1], A[2][0], A[2][1], b[2]) / (d0 + 1e-7f);
    const float rx = pos_b[2 * i] - pos_g[2 * i];
    const float ry = pos_b[2 * i + 1] - pos_g[2 * i + 1];
    out_mdbc[i] = std::max(c0 + (rx * c1 + ry * c2), rho0);
  }
  return 0;
}

// moduloOp, v2/cppSrc/computeNeighborhood.h:110-113: float p - q, the rest in
// double.
static double modulo_ref(float p, float q, float h) {
  return ((p - q + h / 2.0) - std::floor((p - q + h / 2.0) / h) * h) - h / 2.0;
}

// countingKernel / updateKernel, v2/cppSrc/computeNeighborhood.h:115-187.
// With out_i == NULL only counts[i] is written; otherwise entries go to
// new_off[i] + n.
int orc_op_neighborhood(const float* pos_i, const float* h_i, int32_t n, const float* pos_j, const float* h_j,
                        const int64_t* idx_j, const int32_t* nn, const int32_t* off, const float* dmin,
                        const float* dmax, const uint8_t* per, int32_t* counts, const int32_t* new_off,
                        int64_t* out_i, int64_t* out_j, float* out_rij, float* out_xij, float* out_hij) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float xi = pos_i[2 * i], yi = pos_i[2 * i + 1], hi = h_i[i];
    int32_t counter = 0;
    for (int32_t t = 0; t < nn[i]; ++t) {
      const int32_t j = (int32_t)idx_j[(int64_t)off[i] + t];
      const float xj = pos_j[2 * j], yj = pos_j[2 * j + 1], hj = h_j[j];
      auto dx = per[0] ? modulo_ref(xi, xj, dmax[0] - dmin[0]) : xi - xj;
      auto dy = per[1] ? modulo_ref(yi, yj, dmax[1] - dmin[1]) : yi - yj;
      auto sum = dx * dx + dy * dy;
      if (!(sum <= hj * hj)) continue;
      if (out_i) {
        const int64_t k = (int64_t)new_off[i] + counter;
        out_i[k] = i;
        out_j[k] = j;
        auto dist = std::sqrt(sum);
        out_rij[k] = dist / hj;
        out_xij[2 * k] = dx / (dist + 1e-6f * hi);
        out_xij[2 * k + 1] = dy / (dist + 1e-6f * hi);
        out_hij[k] = hj;
      }
      counter++;
    }
    if (counts) counts[i] = counter;
  }
  return 0;
}

// computeNeighborhood (v2/modules/neighborhood.py:433-460) on a candidate
// CSR, float32, with the kernel values of neighborSearch (:585-589).  Same
// count / fill convention as orc_op_neighborhood; positions are wrapped.
int orc_pairs_filter(const orc_domain* dom, int32_t mode, const float* pos_i, const float* h_i, int32_t n,
                     const float* pos_j, const float* h_j, const int64_t* idx_j, const int32_t* nn,
                     const int32_t* off, int32_t* counts, const int32_t* new_off, int64_t* out_i, int64_t* out_j,
                     float* out_rij, float* out_xij, float* out_hij, float* out_w, float* out_gw) {
  const DomT<float> d(*dom);
  set_threads();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int32_t counter = 0;
    const float hi = h_i ? h_i[i] : 0.f;
    for (int32_t t = 0; t < nn[i]; ++t) {
      const int64_t j = idx_j[(int64_t)off[i] + t];
      const float hj = h_j ? h_j[j] : 0.f;
      const G<float> g = geom_exact<float>(pos_i + 2 * i, pos_j + 2 * j, pick_h<float>(mode, hi, hj), d);
      if (!(g.q <= 1.0f)) continue;
      if (out_i) {
        const int64_t k = (int64_t)new_off[i] + counter;
        out_i[k] = i;
        out_j[k] = j;
        out_rij[k] = g.q;
        out_xij[2 * k] = g.ex;
        out_xij[2 * k + 1] = g.ey;
        out_hij[k] = g.h;
        out_w[k] = g.w;
        out_gw[2 * k] = g.gx;
        out_gw[2 * k + 1] = g.gy;
      }
      counter++;
    }
    if (counts) counts[i] = counter;
  }
  return 0;
}

// countNeighborsForParticle / buildNeighborhood, neighborhood.h:100-325, with
// the cell offsets of neighborhood.cpp:150-163 passed in (noff rows of 2).
// The hash bucket is the non-negative remainder (see the HIP twin).
int orc_hash_neighbors(const float* qpos, const float* qh, int32_t nq, const int32_t* offsets, int32_t noff,
                       const float* spos, const float* sh, const int32_t* table, int32_t hlen, const int32_t* ncell,
                       const int64_t* cells, float hcell, const float* dmin, const float* dmax, const uint8_t* per,
                       int32_t mode, int32_t* counts, const int32_t* list_off, int32_t* out_i, int32_t* out_j) {
  set_threads();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nq; ++i) {
    int32_t cell[2];
    for (int dd = 0; dd < 2; ++dd)
      cell[dd] = std::fmin(std::fmax(std::floor((qpos[2 * i + dd] - dmin[dd]) / hcell), -1e9f), 1e9f);
    int32_t counter = 0;
    for (int32_t c = 0; c < noff; ++c) {
      int32_t oc[2];
      for (int dd = 0; dd < 2; ++dd) {
        const int32_t v = cell[dd] + offsets[2 * c + dd];
        oc[dd] = per[dd] ? (v >= 0 ? v % ncell[dd] : ((v % ncell[dd]) + ncell[dd]) % ncell[dd]) : v;
      }
      const int32_t lin = (int32_t)((uint32_t)oc[0] + (uint32_t)oc[1] * (uint32_t)ncell[0]);
      const int32_t hsh = (int32_t)((uint32_t)oc[0] * 73856093u + (uint32_t)oc[1] * 19349663u);
      const int32_t bucket = ((hsh % hlen) + hlen) % hlen;
      const int32_t hb = table[2 * bucket], hl = table[2 * bucket + 1];
      if (hb == -1) continue;
      for (int32_t t = hb; t < hb + hl; ++t) {
        if (cells[3 * t] != lin) continue;
        const int64_t cb = cells[3 * t + 1], ce = cb + cells[3 * t + 2];
        for (int64_t j = cb; j < ce; ++j) {
          float sum = 0.f;
          for (int dd = 0; dd < 2; ++dd) {
            const float a = qpos[2 * i + dd], b = spos[2 * j + dd];
            auto diff = per[dd] ? modulo_ref(a, b, dmax[dd] - dmin[dd]) : a - b;
            sum += diff * diff;
          }
          const float dist = std::sqrt(sum);
          bool keep = mode == 0 ? dist < sh[j] : (mode == 1 ? dist < qh[i] : dist < (qh[i] + sh[j]) / 2.f);
          if (!keep) continue;
          if (out_i) {
            out_i[list_off[i] + counter] = i;
            out_j[list_off[i] + counter] = (int32_t)j;
          }
          counter++;
        }
        break;
      }
    }
    if (counts) counts[i] = counter;
  }
  return 0;
}

// The portable float32 atan2 and sin/cos.
int orc_atan2_pm(const float* y, const float* x, int64_t n, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = dsph_pm::atan2f_pm(y[i], x[i]);
  return 0;
}
int orc_sincos_pm(const float* x, int64_t n, float* s, float* c) {
  for (int64_t i = 0; i < n; ++i) dsph_pm::sincosf_pm(x[i], s + i, c + i);
  return 0;
}

}  // extern "C"
