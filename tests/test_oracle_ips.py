"""The numpy IPS oracle (implicit particle shifting, computeShifting of
v2/modules/shifting.py) against the reference's own outputs
(tests/golden/reference_ips_16.npz, written by
tests/golden/make_reference_golden.py): the kernel Hessians, the assembled
right-hand side, the shift and the BiCGStab exit for the default solver
threshold (one iteration) and for a solve that runs to rtol."""
import os

import numpy as np
import pytest

from oracle import ips

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_ips_16.npz")


def _pairs(d):
    i = d["i"].astype(np.int64)
    j = d["j"].astype(np.int64)
    p = d["pos"].astype(np.float32)
    x = p[i] - p[j]
    # v2/math.py mod on [-1, 1]
    x = np.where(np.abs(x) > 1.0, np.sign(x) * ((np.abs(x) - 1.0) % 2.0 - 1.0), x).astype(np.float32)
    r = np.sqrt((x * x).sum(1)).astype(np.float32)
    xhat = (x / (r + np.float32(1e-7))[:, None]).astype(np.float32)
    h = np.float32(d["h"])
    rij = (r / h).astype(np.float32)
    q = rij.astype(np.float64)
    dw = (-20.0 * q * np.clip(1 - q, 0, 1) ** 3) * ips.C2D / float(h) ** 3
    return i, j, rij, xhat, np.full_like(rij, h), (xhat * dw[:, None]).astype(np.float32)


@pytest.fixture(scope="module")
def g():
    return dict(np.load(GOLD))


def test_hessian_matches_reference(g):
    i, j, rij, xhat, hij, grad = _pairs(g)
    H = ips.hessian(rij, xhat, hij)
    ref = g["hessian"].astype(np.float64)
    assert np.abs(H - ref).max() <= 1e-6 * np.abs(ref).max()
    assert np.array_equal(H[:, 0, 1], H[:, 1, 0])


@pytest.mark.parametrize("label,thr,code", [("default", -0.5, -12), ("solve", 1.0, 11)])
def test_shift_matches_reference(g, label, thr, code):
    i, j, rij, xhat, hij, grad = _pairs(g)
    n = g["pos"].shape[0]
    areas = np.full(n, float(g["dx"]) ** 2, np.float32)
    upd, B, got_code, res = ips.compute_shifting(i, j, rij, xhat, hij, grad, areas, n, threshold=thr)
    ref_b = g[f"{label}_B"].astype(np.float64)
    assert np.abs(B - ref_b).max() <= 1e-5 * np.abs(ref_b).max()
    ref_u = g[f"{label}_update"].astype(np.float64)
    assert np.abs(upd - ref_u).max() <= 5e-5 * np.abs(ref_u).max()
    assert got_code == code == int(g[f"{label}_iters"])
    assert res == pytest.approx(float(g[f"{label}_residual"]), rel=1e-3)


def test_bicgstab_solves_the_system(g):
    """With a positive threshold the returned shift satisfies |H x - B| <=
    rtol |B| (the rule it stopped on), and a uniform lattice (B = 0) gives
    no shift."""
    i, j, rij, xhat, hij, grad = _pairs(g)
    n = g["pos"].shape[0]
    areas = np.full(n, float(g["dx"]) ** 2, np.float32)
    upd, B, code, res = ips.compute_shifting(i, j, rij, xhat, hij, grad, areas, n, threshold=1.0)
    assert code > 0 and res <= 1.01e-4 * np.linalg.norm(B.astype(np.float64))
    x, c = ips.bicgstab_shifting(np.zeros((i.shape[0], 2, 2)), i, j, n, np.zeros(2 * n, np.float32))
    assert c == 0 and not x.any()
