// Per-ghost accumulation and solve of the reference's C++ mDBC op
// (mDBCDensityKernel, v2/cppSrc/mDBC.h:129-244), shared by the wall pass of
// the delta-SPH step (geometry evaluated on the fly) and the per-pair-array
// drop-in (geometry read from the caller's neighbor arrays).
#pragma once

#include <hip/hip_runtime.h>

namespace dsph {

__device__ __forceinline__ float det3(float m11, float m12, float m13, float m21, float m22, float m23, float m31,
                                      float m32, float m33) {
  const float a = m11 * (m22 * m33 - m23 * m32);
  const float b = m12 * (m21 * m33 - m23 * m31);
  const float c = m13 * (m21 * m32 - m22 * m31);
  return a - b + c;
}

// Sums over the fluid neighbors j of one ghost (V_j = m_j / rho_j, x_gj the
// ghost-to-neighbor offset): Shepard numerator / denominator, the moment
// matrix A in the reference's fill order and b = sum m_j [W, gradW].
struct MdbcOpAcc {
  float nom = 0.f, den = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
  float m00 = 0.f, m01 = 0.f, m02 = 0.f, m10 = 0.f, m11 = 0.f, m12 = 0.f, m20 = 0.f, m21 = 0.f, m22 = 0.f;

  __device__ __forceinline__ void add(float m_j, float rho_j, float w, float gx, float gy, float xx, float xy) {
    const float V = m_j / rho_j;
    nom += m_j * w;
    den += V * w;
    const float vw = V * w, vgx = V * gx, vgy = V * gy;
    m00 += vw;
    m10 += vgx;
    m20 += vgy;
    m01 += vw * xx;
    m11 += vgx * xx;
    m21 += vgx * xy;
    m02 += vw * xy;
    m12 += vgy * xx;
    m22 += vgy * xy;
    bx += m_j * w;
    by += m_j * gx;
    bz += m_j * gy;
  }

  // Shepard density nom / (den + 1e-7); with <= 5 neighbors the wall density
  // is the Shepard density, else Cramer's rule on A c = b (det + 1e-7) and
  // c0 + (x_b - x_g) . c[1:]; both clamped below at rho0.
  __device__ __forceinline__ float finish(int cnt, float px, float py, float rho0, float* shep_out) const {
    const float shep = nom / (den + 1e-7f);
    *shep_out = shep;
    if (cnt <= 5) return shep < rho0 ? rho0 : shep;
    const float d0 = det3(m00, m01, m02, m10, m11, m12, m20, m21, m22) + 1e-7f;
    const float rx = det3(bx, m01, m02, by, m11, m12, bz, m21, m22) / d0;
    const float ry = det3(m00, bx, m02, m10, by, m12, m20, bz, m22) / d0;
    const float rz = det3(m00, m01, bx, m10, m11, by, m20, m21, bz) / d0;
    const float out = rx + (px * ry + py * rz);
    return out < rho0 ? rho0 : out;
  }
};

}  // namespace dsph
