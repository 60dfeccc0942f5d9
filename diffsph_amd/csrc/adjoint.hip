// Vector-Jacobian product of the delta-SPH right-hand side (fluid only,
// uniform support), the building block of the hand-written adjoint of the
// RK4 rollout.
//
// Forward (deltasph.hip): with d = x_i - x_j (min-image), r = |d|,
// xhat = d/(r+1e-7), q = r/h, rr = q h, w1 = dW/dr, gradW = xhat w1,
// V_j = m_j / rho_j:
//   a_i    = -(1/rho_i) sum_j V_j (p_i + p_j) gradW             p = c (rho - rho0)
//   visc_i = (cV_i/rho_i) sum_j V_j pi_ij gradW,  pi = (u_i-u_j).xhat / (rr + eps h^2)
//   cont_i = -sum_j m_j ((u_j-u_i).gradW) rho_i/rho_j
//   D_i    = cD_i sum_j V_j psi_ij.gradW,  psi = -2(rho_j-rho_i)/(rr+1e-6 h) xhat + g_i + g_j
//   g_i    = L_i s_i,  s_i = sum_j V_j (rho_i-rho_j) gradW,  L_i = pinv(M_i),
//   M_i    = sum_j V_j (-rr xhat) (x) gradW   (M_i = I, L_i constant, for rows < 4)
//   dudt = a + visc + g0, drho/dt = cont + D.
// Given cotangents (la = dl/d dudt, lc = dl/d drhodt) the VJP returns
// dl/dx, dl/du, dl/drho.  dL = -L dM L (the regularisation of the reference's
// closed-form pseudo-inverse changes it by ~1e-7 relative and is not
// differentiated), dt and the neighbor sets are constants.
//
// Uniform h makes the neighbor sets symmetric and the pair geometry exactly
// antisymmetric (fl(x_j - x_i) = -fl(x_i - x_j)), so every kernel gathers:
// particle k sums, over its own row, the pair terms where it is the row
// owner and those where it is the neighbor, using the neighbor's
// cotangents.  No atomics; results do not depend on scheduling.
#include "common.h"
#include "pmath.h"

namespace dsph {

struct VjpConst {
  float h;
  float rho0, eos_c, cs;
  float cV;        // viscid: 2(d+2) nu rho0; inviscid: alpha h cs rho0
  float eps;       // viscosity eps
  float cD;        // delta h cs / kernel_scale
  int visc;        // -1 none, 0 inviscid, 1 viscid
  int pi_switch;
  int diffusion;
};

// Pair geometry and derivatives for d = x_i - x_j (d != 0).
struct PGeo {
  float dx, dy, r, e, ex, ey, q, rr, w1, w1p, gx, gy;
};

__device__ __forceinline__ PGeo pgeo(float dx, float dy, float h) {
  PGeo g;
  g.dx = dx;
  g.dy = dy;
  g.r = sqrt_rn(dist2_exact(dx, dy));
  g.e = g.r + 1e-7f;
  // the forward kernels' reciprocal form (pair_geom)
  const KConst k = kconst(h);
  const float ri = 1.0f / g.e;
  g.ex = dx * ri;
  g.ey = dy * ri;
  g.q = g.r * k.ih;
  g.rr = g.q * h;
  const float t = clamp01(1.0f - g.q);
  g.w1 = wendland_dw(g.q, k);
  // d w1 / dr = C/h^4 (-20) (1-q)^2 (1-4q)
  g.w1p = g.q < 1.0f ? (((-20.0f * (t * t)) * (1.0f - 4.0f * g.q)) * kC2) / (((h * h) * h) * h) : 0.0f;
  g.gx = g.ex * g.w1;
  g.gy = g.ey * g.w1;
  return g;
}

// The geometry of the mirrored pair (d -> -d): every quantity of pgeo is
// either even in d (r, e, q, rr, w1, w1p: the squares, sums and kernel
// values are the same floats) or odd and formed as d times an even one
// (dx, dy, ex, ey, gx, gy: IEEE negation commutes with the products), so
// this equals pgeo(-dx, -dy, h) bit for bit without its sqrt and division.
__device__ __forceinline__ PGeo mirrored(const PGeo& g) {
  PGeo m = g;
  m.dx = -g.dx;
  m.dy = -g.dy;
  m.ex = -g.ex;
  m.ey = -g.ey;
  m.gx = -g.gx;
  m.gy = -g.gy;
  return m;
}

// J_xhat^T v (J_xhat is symmetric): v/e - d (d.v) / (r e^2)
__device__ __forceinline__ float2 jxhat(const PGeo& g, float vx, float vy) {
  const float dv = g.dx * vx + g.dy * vy;
  const float c = dv / (g.r * g.e * g.e);
  return make_float2(vx / g.e - g.dx * c, vy / g.e - g.dy * c);
}

// J_gradW^T v = w1 J_xhat^T v + (w1'/r)(xhat.v) d
__device__ __forceinline__ float2 jgradw(const PGeo& g, float vx, float vy) {
  const float2 a = jxhat(g, vx, vy);
  const float c = (g.w1p / g.r) * (g.ex * vx + g.ey * vy);
  return make_float2(g.w1 * a.x + c * g.dx, g.w1 * a.y + c * g.dy);
}

struct PairCot {
  float lax, lay, lc;        // cotangents of dudt_i, drhodt_i
  float lsx, lsy;            // of s_i
  float m00, m01, m10, m11;  // of M_i
};

struct PairOut {
  float gdx, gdy;      // dl/dd
  float gui_x, gui_y;  // dl/du_i
  float guj_x, guj_y;  // dl/du_j
  float gri, grj;      // dl/drho_i, dl/drho_j
};

// All pair terms of row i with neighbor j (see the header).
__device__ __forceinline__ PairOut pair_vjp(const PGeo& g, const VjpConst& K, float rho_i, float rho_j, float2 ui,
                                            float2 uj, float m_j, float2 gi, float2 gj, const PairCot& c) {
  PairOut o = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const float Vj = m_j / rho_j;
  const float lw = c.lax * g.gx + c.lay * g.gy;
  const float r_inv = 1.0f / g.r;
  // pressure
  {
    const float P = K.eos_c * (rho_i - K.rho0) + K.eos_c * (rho_j - K.rho0);
    const float coef = -Vj * P / rho_i;
    const float2 jl = jgradw(g, c.lax, c.lay);
    o.gdx += coef * jl.x;
    o.gdy += coef * jl.y;
    o.gri += (Vj * P / (rho_i * rho_i)) * lw - (Vj * K.eos_c / rho_i) * lw;
    o.grj += (Vj * P / (rho_i * rho_j)) * lw - (Vj * K.eos_c / rho_i) * lw;
  }
  // viscosity
  if (K.visc >= 0) {
    const float vx = ui.x - uj.x, vy = ui.y - uj.y;
    const float den = g.rr + K.eps * (K.h * K.h);
    const float vr = vx * g.ex + vy * g.ey;
    if (!(K.pi_switch && !(vr < 0.0f))) {
      const float pi = vr / den;
      const float Kv = K.cV * Vj / rho_i;
      const float t = Kv * pi * lw;
      const float f = Kv * lw / den;
      o.gui_x += f * g.ex;
      o.gui_y += f * g.ey;
      o.guj_x -= f * g.ex;
      o.guj_y -= f * g.ey;
      o.gri -= t / rho_i;
      o.grj -= t / rho_j;
      const float2 jv = jxhat(g, vx, vy);
      const float2 jl = jgradw(g, c.lax, c.lay);
      const float dpx = jv.x / den - (pi / den) * (g.dx * r_inv);
      const float dpy = jv.y / den - (pi / den) * (g.dy * r_inv);
      o.gdx += Kv * (lw * dpx + pi * jl.x);
      o.gdy += Kv * (lw * dpy + pi * jl.y);
    }
  }
  // continuity
  {
    const float dux = uj.x - ui.x, duy = uj.y - ui.y;
    const float f = m_j * rho_i / rho_j;
    const float t = -c.lc * f * (dux * g.gx + duy * g.gy);
    o.guj_x += -c.lc * f * g.gx;
    o.guj_y += -c.lc * f * g.gy;
    o.gui_x += c.lc * f * g.gx;
    o.gui_y += c.lc * f * g.gy;
    o.gri += t / rho_i;
    o.grj += -t / rho_j;
    const float2 jd = jgradw(g, dux, duy);
    o.gdx += -c.lc * f * jd.x;
    o.gdy += -c.lc * f * jd.y;
  }
  // density diffusion (g_i, g_j held fixed here; their cotangent is gathered separately)
  if (K.diffusion) {
    const float den2 = g.rr + 1e-6f * K.h;
    const float rt = 2.0f * (rho_j - rho_i) / den2;
    const float xx = g.ex * g.ex + g.ey * g.ey;
    const float xw = xx * g.w1;
    const float Gx = gi.x + gj.x, Gy = gi.y + gj.y;
    const float psiw = -rt * xw + (Gx * g.gx + Gy * g.gy);
    const float a = c.lc * K.cD * Vj;
    o.gri += a * (2.0f / den2) * xw;
    o.grj += -a * (2.0f / den2) * xw - a * psiw / rho_j;
    const float2 jx = jxhat(g, g.ex, g.ey);
    const float dxw_x = 2.0f * g.w1 * jx.x + xx * g.w1p * (g.dx * r_inv);
    const float dxw_y = 2.0f * g.w1 * jx.y + xx * g.w1p * (g.dy * r_inv);
    const float drt = -rt / den2;  // times d/r
    const float2 jG = jgradw(g, Gx, Gy);
    o.gdx += a * (-(drt * (g.dx * r_inv)) * xw - rt * dxw_x + jG.x);
    o.gdy += a * (-(drt * (g.dy * r_inv)) * xw - rt * dxw_y + jG.y);
  }
  // s_i = sum V_j (rho_i - rho_j) gradW
  {
    const float ls = c.lsx * g.gx + c.lsy * g.gy;
    const float ts = Vj * (rho_i - rho_j) * ls;
    o.gri += Vj * ls;
    o.grj += -Vj * ls - ts / rho_j;
    const float2 js = jgradw(g, c.lsx, c.lsy);
    o.gdx += Vj * (rho_i - rho_j) * js.x;
    o.gdy += Vj * (rho_i - rho_j) * js.y;
  }
  // M_i = sum V_j (-rr xhat) (x) gradW = -sum V_j rr w1 xhat xhat^T
  {
    const float a = g.ex * (c.m00 * g.ex + c.m01 * g.ey) + g.ey * (c.m10 * g.ex + c.m11 * g.ey);
    const float tM = -Vj * g.rr * g.w1 * a;
    o.grj += -tM / rho_j;
    const float sx = (c.m00 + c.m00) * g.ex + (c.m01 + c.m10) * g.ey;
    const float sy = (c.m10 + c.m01) * g.ex + (c.m11 + c.m11) * g.ey;
    const float2 jm = jxhat(g, sx, sy);
    const float b = a * (g.w1 + g.rr * g.w1p);
    o.gdx += -Vj * (b * (g.dx * r_inv) + g.rr * g.w1 * jm.x);
    o.gdy += -Vj * (b * (g.dy * r_inv) + g.rr * g.w1 * jm.y);
  }
  return o;
}

// lambda_g_k = sum_j (cD lc_k V_j - cD lc_j V_k) gradW_kj
__global__ void __launch_bounds__(kBlock) k_vjp_lg(PairsDev P, VjpConst K, const float* __restrict__ mass,
                                                   const float* __restrict__ rho, const float* __restrict__ lc,
                                                   float2* __restrict__ lg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nq) return;
  const float2 xk = P.qpos[k];
  const float Vk = mass[k] / rho[k];
  const float lck = lc[k];
  float ax = 0.f, ay = 0.f;
  const RowRef rr = row_ref(P, k);
  for (int n = 0, e = rr.base; n < rr.cnt; ++n, e += rr.step) {
    const int j = P.nbr[e];
    if (j == k) continue;
    const float2 xj = P.rpos[j];
    const PGeo g = pgeo(mod_offset(xk.x - xj.x, P.dom, 0), mod_offset(xk.y - xj.y, P.dom, 1), K.h);
    const float Vj = mass[j] / rho[j];
    const float f = K.cD * lck * Vj - K.cD * lc[j] * Vk;
    ax += f * g.gx;
    ay += f * g.gy;
  }
  lg[k] = make_float2(ax, ay);
}

// lambda_s_i = L_i^T lambda_g_i, lambda_M_i = -L_i^T (lambda_g_i s_i^T) L_i^T
// (zero when the row has < 4 entries: M_i = I).
__global__ void __launch_bounds__(kBlock) k_vjp_lsm(PairsDev P, VjpConst K, const float* __restrict__ mass,
                                                    const float* __restrict__ rho, const float2* __restrict__ lg,
                                                    float2* __restrict__ ls, float4* __restrict__ lM) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nq) return;
  const float2 xi = P.qpos[i];
  const float rho_i = rho[i];
  float m00 = 0.f, m01 = 0.f, m10 = 0.f, m11 = 0.f, sx = 0.f, sy = 0.f;
  const RowRef rr = row_ref(P, i);
  for (int n = 0, e = rr.base; n < rr.cnt; ++n, e += rr.step) {
    const int j = P.nbr[e];
    const Geom g = pair_geom(xi, P.rpos[j], K.h, P.dom);
    const float Vj = mass[j] / rho[j];
    const float s = -(g.q * g.h);
    const float qx = s * g.ex, qy = s * g.ey;
    const float kx = Vj * g.gx, ky = Vj * g.gy;
    m00 += qx * kx;
    m01 += qx * ky;
    m10 += qy * kx;
    m11 += qy * ky;
    const float d = rho_i - rho[j];
    sx += d * kx;
    sy += d * ky;
  }
  const float2 l = lg[i];
  if (rr.cnt < 4) {
    ls[i] = l;  // L = I
    lM[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const dsph_pm::Pinv2 L = dsph_pm::pinv2x2(m00, m01, m10, m11);
  ls[i] = make_float2(L.l00 * l.x + L.l10 * l.y, L.l01 * l.x + L.l11 * l.y);
  // A = lambda_g s^T ; lM = -L^T A L^T
  const float a00 = l.x * sx, a01 = l.x * sy, a10 = l.y * sx, a11 = l.y * sy;
  // B = L^T A
  const float b00 = L.l00 * a00 + L.l10 * a10, b01 = L.l00 * a01 + L.l10 * a11;
  const float b10 = L.l01 * a00 + L.l11 * a10, b11 = L.l01 * a01 + L.l11 * a11;
  // C = B L^T
  lM[i] = make_float4(-(b00 * L.l00 + b01 * L.l01), -(b00 * L.l10 + b01 * L.l11), -(b10 * L.l00 + b11 * L.l01),
                      -(b10 * L.l10 + b11 * L.l11));
}

struct VjpArgs {
  const float2* vel;
  const float* mass;
  const float* rho;
  const float2* grad_rho;
  const float2* la;
  const float* lc;
  const float2* ls;
  const float4* lM;
  const float2* lx_in;  // cotangent of dx/dt (= u), added to lambda_u; may be NULL
  float2* out_x;
  float2* out_u;
  float* out_rho;
};

__global__ void __launch_bounds__(kBlock) k_vjp_main(PairsDev P, VjpConst K, VjpArgs A) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nq) return;
  const float2 xk = P.qpos[k];
  const float2 uk = A.vel[k];
  const float rk = A.rho[k];
  const float mk = A.mass[k];
  const float2 gk = K.diffusion ? A.grad_rho[k] : make_float2(0.f, 0.f);
  PairCot ck;
  {
    const float2 a = A.la[k];
    const float2 s = A.ls[k];
    const float4 m = A.lM[k];
    ck = {a.x, a.y, A.lc[k], s.x, s.y, m.x, m.y, m.z, m.w};
  }
  float lxx = 0.f, lxy = 0.f, lux = 0.f, luy = 0.f, lr = 0.f;
  const RowRef rr = row_ref(P, k);
  for (int n = 0, e = rr.base; n < rr.cnt; ++n, e += rr.step) {
    const int j = P.nbr[e];
    if (j == k) continue;
    const float2 xj = P.rpos[j];
    const float dx = mod_offset(xk.x - xj.x, P.dom, 0), dy = mod_offset(xk.y - xj.y, P.dom, 1);
    const float2 uj = A.vel[j];
    const float rj = A.rho[j];
    const float mj = A.mass[j];
    const float2 gj = K.diffusion ? A.grad_rho[j] : make_float2(0.f, 0.f);
    // k as the row owner of pair (k, j)
    const PGeo g1 = pgeo(dx, dy, K.h);
    const PairOut o1 = pair_vjp(g1, K, rk, rj, uk, uj, mj, gk, gj, ck);
    lxx += o1.gdx;
    lxy += o1.gdy;
    lux += o1.gui_x;
    luy += o1.gui_y;
    lr += o1.gri;
    // k as the neighbor in pair (j, k): d_jk = -d_kj
    PairCot cj;
    {
      const float2 a = A.la[j];
      const float2 s = A.ls[j];
      const float4 m = A.lM[j];
      cj = {a.x, a.y, A.lc[j], s.x, s.y, m.x, m.y, m.z, m.w};
    }
    const PGeo g2 = mirrored(g1);
    const PairOut o2 = pair_vjp(g2, K, rj, rk, uj, uk, mk, gj, gk, cj);
    lxx -= o2.gdx;
    lxy -= o2.gdy;
    lux += o2.guj_x;
    luy += o2.guj_y;
    lr += o2.grj;
  }
  if (A.lx_in) {
    const float2 t = A.lx_in[k];
    lux += t.x;
    luy += t.y;
  }
  A.out_x[k] = make_float2(lxx, lxy);
  A.out_u[k] = make_float2(lux, luy);
  A.out_rho[k] = lr;
}

}  // namespace dsph

using namespace dsph;

extern "C" size_t dsph_deltasph_vjp_workspace(int64_t n) {
  return ((size_t)n * (2 + 2 + 4) * sizeof(float) + 255) / 256 * 256 + 256;
}

extern "C" int dsph_deltasph_rhs_vjp(const dsph_pairs* pairs, const dsph_deltasph_params* prm, const float* vel,
                                     const float* mass, const float* rho, const float* grad_rho,
                                     const float* cot_dxdt, const float* cot_dudt, const float* cot_drhodt,
                                     float* grad_pos, float* grad_vel, float* grad_rho_out, void* ws,
                                     size_t ws_bytes, void* stream) {
  int s = check_pairs(pairs);
  if (s != DSPH_OK) return s;
  if (!prm || pairs->nq != pairs->nr) return DSPH_ERR_ARG;
  // symmetric lists: one uniform support, scatter mode
  if (pairs->qh || pairs->rh || !(pairs->h_uniform > 0.0f) || pairs->mode != DSPH_MODE_SCATTER) return DSPH_ERR_ARG;
  if (prm->viscosity < -1 || prm->viscosity > 1 || prm->eos < 0 || prm->eos > 1) return DSPH_ERR_ARG;
  const int64_t n = pairs->nq;
  if (n == 0) return DSPH_OK;
  if (!vel || !mass || !rho || !cot_dudt || !cot_drhodt || !grad_pos || !grad_vel || !grad_rho_out)
    return DSPH_ERR_ARG;
  if (prm->diffusion && (!grad_rho || !(prm->kernel_scale > 0.0f))) return DSPH_ERR_ARG;
  if (!ws || ws_bytes < dsph_deltasph_vjp_workspace(n)) return DSPH_ERR_CAPACITY;
  hipStream_t st = as_stream(stream);
  PairsDev P = make_pairs(*pairs);
  VjpConst K;
  K.h = pairs->h_uniform;
  K.rho0 = prm->rho0;
  K.eos_c = prm->eos == 0 ? prm->cs * prm->cs : prm->eos_kappa;
  K.cs = prm->cs;
  K.cV = prm->viscosity == 1 ? (float)(2.0 * 4.0 * (double)prm->visc_coeff * (double)prm->rho0)
                             : (((prm->visc_coeff * K.h) * prm->cs) * prm->rho0);
  K.eps = prm->visc_eps;
  K.cD = prm->diffusion ? ((prm->delta * K.h) / prm->kernel_scale) * prm->cs : 0.0f;
  K.visc = prm->viscosity;
  K.pi_switch = prm->pi_switch;
  K.diffusion = prm->diffusion;
  char* w = static_cast<char*>(ws);
  float2* lg = reinterpret_cast<float2*>(w);
  float2* ls = lg + n;
  float4* lM = reinterpret_cast<float4*>(ls + n);
  const int nb = grid_for(n);
  if (prm->diffusion) {
    k_vjp_lg<<<nb, kBlock, 0, st>>>(P, K, mass, rho, cot_drhodt, lg);
    k_vjp_lsm<<<nb, kBlock, 0, st>>>(P, K, mass, rho, lg, ls, lM);
  } else {
    DSPH_CHECK_HIP(hipMemsetAsync(ls, 0, sizeof(float) * 6 * n, st));
  }
  VjpArgs A;
  A.vel = reinterpret_cast<const float2*>(vel);
  A.mass = mass;
  A.rho = rho;
  A.grad_rho = reinterpret_cast<const float2*>(grad_rho);
  A.la = reinterpret_cast<const float2*>(cot_dudt);
  A.lc = cot_drhodt;
  A.ls = ls;
  A.lM = lM;
  A.lx_in = reinterpret_cast<const float2*>(cot_dxdt);
  A.out_x = reinterpret_cast<float2*>(grad_pos);
  A.out_u = reinterpret_cast<float2*>(grad_vel);
  A.out_rho = grad_rho_out;
  k_vjp_main<<<nb, kBlock, 0, st>>>(P, K, A);
  return launch_status();
}
