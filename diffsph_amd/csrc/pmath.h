// Portable float32 atan2 / sin / cos and the 2x2 pseudo-inverse built on them.
//
// The functions evaluate their series in float64 with IEEE +, -, *, / only
// (no library calls; the code is compiled without FP contraction) and round
// once to float32.  The same header is compiled by hipcc for the device and
// by g++ for the CPU, so both produce identical bits; the float64 series are
// accurate to well below one float32 ulp.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define DSPH_HD __host__ __device__
#else
#define DSPH_HD
#endif

namespace dsph_pm {

constexpr double kPi = 3.14159265358979323846;
constexpr double kHalfPi = 1.57079632679489661923;
constexpr double kQuarterPi = 0.78539816339744830962;
constexpr double kTanEighthPi = 0.41421356237309504880;

// atan(z), |z| <= tan(pi/8): Maclaurin series sum_k (-z^2)^k z / (2k+1),
// 24 terms (the first omitted term is below 0.1716^24 / 49 < 1e-19).
DSPH_HD inline double atan_small(double z) {
  const double z2 = z * z;
  double s = 0.0;
  for (int k = 23; k >= 0; --k) s = 1.0 / (double)(2 * k + 1) - z2 * s;
  return z * s;
}

// atan(t) for t in [0, 1]; above tan(pi/8) use atan t = pi/4 + atan((t-1)/(t+1)).
DSPH_HD inline double atan_unit(double t) {
  if (t > kTanEighthPi) return kQuarterPi + atan_small((t - 1.0) / (t + 1.0));
  return atan_small(t);
}

// atan2 with the C99 special cases for signed zeros; NaN in -> NaN out.
DSPH_HD inline float atan2f_pm(float y, float x) {
  if (y != y || x != x) return y + x;
  const double dy = (double)y, dx = (double)x;
  const double ay = dy < 0.0 ? -dy : dy;
  const double ax = dx < 0.0 ? -dx : dx;
  double r;
  if (ax == 0.0 && ay == 0.0)
    r = 0.0;
  else if (ay <= ax)
    r = atan_unit(ay / ax);
  else
    r = kHalfPi - atan_unit(ax / ay);
  if (std::signbit(x)) r = kPi - r;
  return (float)(std::signbit(y) ? -r : r);
}

// sin and cos for |x| <= 2 (the angles here lie in [-pi/2, pi/2]), nested
// Horner form of the Maclaurin series, 14 terms each (remainder < 1e-24).
DSPH_HD inline void sincosf_pm(float xf, float* s, float* c) {
  const double x = (double)xf, x2 = x * x;
  double ss = 1.0, cc = 1.0;
  for (int k = 14; k >= 1; --k) {
    ss = 1.0 - x2 / (double)((2 * k) * (2 * k + 1)) * ss;
    cc = 1.0 - x2 / (double)((2 * k - 1) * (2 * k)) * cc;
  }
  *s = (float)(x * ss);
  *c = (float)cc;
}

DSPH_HD inline float signf_torch(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }

struct Pinv2 {
  float l00, l01, l10, l11;  // pseudo-inverse, row-major
  float ev0, ev1;            // singular values, larger magnitude first
};

// Pseudo-inverse of M = [[a, b], [c, d]] through the closed-form 2x2 SVD
// M = U S V^T, pinv = V S^-1 U^T (v2/math.py:10-57): U from
// theta = atan2(2ac + 2bd, a^2 + b^2 - c^2 - d^2)/2, V from
// phi = atan2(2ab + 2cd, a^2 - b^2 + c^2 - d^2)/2, singular values from
// S1 = |M|_F^2 and S2, the smaller one regularised by +1e-7 and floored at
// 1e-7, reciprocals of values below 1e-5 set to 0, V's columns signed so
// that U^T M V has a non-negative diagonal (sign(0) = 0).
DSPH_HD inline Pinv2 pinv2x2(float a, float b, float c, float d) {
  const float a2 = a * a, b2 = b * b, c2 = c * c, d2 = d * d;
  const float theta = 0.5f * atan2f_pm(2.0f * a * c + 2.0f * b * d, a2 + b2 - c2 - d2);
  float st, ct;
  sincosf_pm(theta, &st, &ct);
  const float S1 = a2 + b2 + c2 + d2;
  const float e = a2 + b2 - c2 - d2;
  const float f = a * c + b * d;
  const float S2 = sqrtf(e * e + 4.0f * (f * f));
  const float o1 = sqrtf((S1 + S2) / 2.0f);
  float t = S1 - S2 + 1e-7f;
  t = t < 1e-7f ? 1e-7f : t;
  const float o2 = sqrtf(t / 2.0f);
  const float phi = 0.5f * atan2f_pm(2.0f * a * b + 2.0f * c * d, a2 - b2 + c2 - d2);
  float sp, cp;
  sincosf_pm(phi, &sp, &cp);
  const float s11 = signf_torch((a * ct + c * st) * cp + (b * ct + d * st) * sp);
  const float s22 = signf_torch((a * st - c * ct) * sp + (-b * st + d * ct) * cp);
  const float v00 = cp * s11, v01 = -sp * s22, v10 = sp * s11, v11 = cp * s22;
  const float i1 = fabsf(o1) > 1e-5f ? 1.0f / o1 : 0.0f;
  const float i2 = fabsf(o2) > 1e-5f ? 1.0f / o2 : 0.0f;
  const float w00 = v00 * i1, w01 = v01 * i2, w10 = v10 * i1, w11 = v11 * i2;  // V S^-1
  Pinv2 r;
  r.l00 = w00 * ct + w01 * (-st);
  r.l01 = w00 * st + w01 * ct;
  r.l10 = w10 * ct + w11 * (-st);
  r.l11 = w10 * st + w11 * ct;
  const bool swap = fabsf(o2) > fabsf(o1);
  r.ev0 = swap ? o2 : o1;
  r.ev1 = swap ? o1 : o2;
  return r;
}

// Pseudo-inverse of a 3x3 matrix (row-major a[9]) as torch.linalg.pinv
// computes it: SVD, singular values <= rtol * sigma_max dropped (rtol is
// max(m, n) * eps of float32 for a float32 input).  Evaluated in float64:
// cyclic Jacobi on A^T A = V diag(l) V^T (16 sweeps, far past convergence),
// then pinv = sum over kept i of v_i (A v_i)^T / l_i.  Returns the number of
// singular values kept.
DSPH_HD inline int pinv3x3(const double* a, double rtol, double* out) {
  double S[3][3], V[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += a[3 * k + r] * a[3 * k + c];
      S[r][c] = s;
      V[r][c] = r == c ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 16; ++sweep) {
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      const double apq = S[p][q];
      if (apq == 0.0) continue;
      const double theta = (S[q][q] - S[p][p]) / (2.0 * apq);
      const double at = theta < 0.0 ? -theta : theta;
      const double t = (theta < 0.0 ? -1.0 : 1.0) / (at + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; ++k) {  // S <- S J (columns p, q)
        const double skp = S[k][p], skq = S[k][q];
        S[k][p] = c * skp - s * skq;
        S[k][q] = s * skp + c * skq;
      }
      for (int k = 0; k < 3; ++k) {  // S <- J^T S (rows p, q)
        const double spk = S[p][k], sqk = S[q][k];
        S[p][k] = c * spk - s * sqk;
        S[q][k] = s * spk + c * sqk;
      }
      for (int k = 0; k < 3; ++k) {  // V <- V J
        const double vkp = V[k][p], vkq = V[k][q];
        V[k][p] = c * vkp - s * vkq;
        V[k][q] = s * vkp + c * vkq;
      }
    }
  }
  double sig[3], smax = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double l = S[i][i] > 0.0 ? S[i][i] : 0.0;
    sig[i] = sqrt(l);
    smax = sig[i] > smax ? sig[i] : smax;
  }
  for (int k = 0; k < 9; ++k) out[k] = 0.0;
  int kept = 0;
  for (int i = 0; i < 3; ++i) {
    if (!(sig[i] > rtol * smax) || sig[i] == 0.0) continue;
    ++kept;
    double av[3];
    for (int r = 0; r < 3; ++r) av[r] = a[3 * r] * V[0][i] + a[3 * r + 1] * V[1][i] + a[3 * r + 2] * V[2][i];
    const double inv = 1.0 / (sig[i] * sig[i]);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) out[3 * r + c] += V[r][i] * av[c] * inv;
  }
  return kept;
}

}  // namespace dsph_pm
