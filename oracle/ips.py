"""Implicit particle shifting (IPS) restated in numpy: test infrastructure,
the checker of the GPU path, never part of it.

Follows src/diffSPH/v2/modules/shifting.py: computeShifting (:361-460) with
getShiftingMatrices (:192-284, no free surface, rest areas as volumes,
'zero' initialization), bicgstab_shifting (:318-~460, its stopping rules:
atol = rtol |B|, rho and omega breakdowns, the solver threshold on the
per-particle |x|) and the kernel Hessian of src/diffSPH/kernels.py:263-303
(Wendland C2, src/diffSPH/kernelFunctions/Wendland2.py).  Vectors are float32,
dot products are summed in float64 (the GPU path does the same).
"""
from __future__ import annotations

import numpy as np

C2D = 7.0 / np.pi


def _dkdq(q):
    t = np.clip(1.0 - q, 0.0, 1.0)
    return -20.0 * q * t ** 3


def _d2kdq2(q):
    t = np.clip(1.0 - q, 0.0, 1.0)
    return 20.0 * (4.0 * q - 1.0) * t ** 2


def hessian(rij, xij, hij):
    """Hessian2D (kernels.py:263-303): [P, 2, 2] from the neighborDict's
    r_ij / h_ij, unit vectors and supports (float64 arithmetic)."""
    rij = rij.astype(np.float64)
    hij = hij.astype(np.float64)
    factor = C2D / hij ** 2
    r = rij * hij
    x = xij.astype(np.float64) * r[:, None]
    q = rij
    k2 = _d2kdq2(q)
    k1 = _dkdq(q)
    tA_x = x[:, 0] ** 2 / (hij ** 2 * r ** 2 + 1e-5 * hij) * k2
    tA_y = x[:, 1] ** 2 / (hij ** 2 * r ** 2 + 1e-5 * hij) * k2
    tB_x = x[:, 1] ** 2 / (hij * r ** 3 + 1e-10 * hij) * k1
    tB_y = x[:, 0] ** 2 / (hij * r ** 3 + 1e-10 * hij) * k1
    self_term = 1.0 / hij ** 2 * np.where(q < 1e-5, _d2kdq2(np.zeros_like(q)), 0.0)
    dxx = factor * (tA_x + tB_x + self_term)
    dyy = factor * (tA_y + tB_y + self_term)
    dxy = C2D / hij ** 2 * np.where(q > -1e-7, (x[:, 0] * x[:, 1]) / (hij * r ** 2 + 1e-5 * hij)
                                    * (1.0 / hij * k2 - 1.0 / (r + 1e-3 * hij) * k1), 0.0)
    H = np.zeros((rij.shape[0], 2, 2))
    inside = q <= 1
    H[:, 0, 0] = np.where(inside, dxx, 0.0)
    H[:, 1, 1] = np.where(inside, dyy, 0.0)
    H[:, 0, 1] = np.where(inside, dxy, 0.0)
    H[:, 1, 0] = H[:, 0, 1]
    return H


def _matvec(Hs, i, j, x, n):
    y = np.zeros(2 * n, np.float64)
    xj0 = x[2 * j].astype(np.float64)
    xj1 = x[2 * j + 1].astype(np.float64)
    np.add.at(y, 2 * i, Hs[:, 0, 0] * xj0 + Hs[:, 0, 1] * xj1)
    np.add.at(y, 2 * i + 1, Hs[:, 1, 0] * xj0 + Hs[:, 1, 1] * xj1)
    return y.astype(np.float32)


def _dot(a, b):
    return float(np.dot(a.astype(np.float64), b.astype(np.float64)))


def bicgstab_shifting(Hs, i, j, n, b, rtol=1e-4, atol=0.0, maxiter=64, threshold=-0.5):
    """bicgstab_shifting (shifting.py:318-...) with x0 = 0; returns (x, code)
    with code the iteration count, 0 (converged at s), -10 / -11 (rho /
    omega breakdown) or -12 (the threshold on |x_i| tripped)."""
    b = b.astype(np.float32)
    bnrm2 = np.sqrt(_dot(b, b))
    atol = max(float(atol), float(rtol) * bnrm2)
    if bnrm2 == 0:
        return b.copy(), 0
    eps = float(np.finfo(np.float32).eps)
    rhotol = omegatol = eps * eps
    x = np.zeros_like(b)
    r = b.copy()
    r0 = r.copy()
    p = r.copy()
    for it in range(maxiter):
        if np.sqrt(_dot(r, r)) < atol:
            return x, it
        rho = _dot(r, r0)
        if abs(rho) < rhotol:
            return x, -10
        ap = _matvec(Hs, i, j, p, n)
        rv = _dot(ap, r0)
        if rv == 0:
            return x, -11
        alpha = np.float32(rho / rv)
        s = (r - alpha * ap).astype(np.float32)
        if np.sqrt(_dot(s, s)) < atol:
            return (x + alpha * p).astype(np.float32), 0
        as_ = _matvec(Hs, i, j, s, n)
        omega = np.float32(_dot(as_, s) / _dot(as_, as_))
        x = ((x + alpha * p) + omega * s).astype(np.float32)
        rho_prev = _dot(r, r0)
        r = (s - omega * as_).astype(np.float32)
        beta = np.float32((_dot(r, r0) / rho_prev) * (float(alpha) / float(omega)))
        p = (r + beta * (p - omega * ap)).astype(np.float32)
        if it > 0 and abs(float(omega)) < omegatol:
            return x, -11
        if np.any(np.sqrt((x.reshape(-1, 2).astype(np.float64) ** 2).sum(1)) > threshold):
            return x, -12
    return x, maxiter


def compute_shifting(i, j, rij, xij, hij, grad, areas, n, rtol=1e-4, maxiter=64, threshold=-0.5):
    """computeShifting (scheme IPS, no free surface, no boundary, rest
    areas): returns (update [n, 2] = -x, B, code, |A x - B|)."""
    A = areas.astype(np.float64)
    H = hessian(rij, xij, hij)
    Jsum = np.zeros((n, 2))
    np.add.at(Jsum, i, grad.astype(np.float64) * A[j][:, None])
    Hs = H * A[j][:, None, None]
    B = np.zeros(2 * n, np.float32)
    B[0::2] = Jsum[:, 0]
    B[1::2] = Jsum[:, 1]
    x, code = bicgstab_shifting(Hs, i, j, n, B, rtol=rtol, maxiter=maxiter, threshold=threshold)
    res = np.sqrt(_dot(_matvec(Hs, i, j, x, n) - B, _matvec(Hs, i, j, x, n) - B))
    upd = np.stack([-x[0::2], -x[1::2]], 1)
    return upd, B, code, res
