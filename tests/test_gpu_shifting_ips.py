"""Implicit particle shifting on the GPU (computeShifting / solveShifting
with scheme 'IPS', v2/modules/shifting.py) against the reference's outputs
(tests/golden/reference_ips_16.npz) and against the numpy IPS oracle on a
larger jittered set.  The Hessian blocks and right-hand side are summed in
the neighbor list's order and the dot products in float64, so values agree
to float32 rounding (amplified by the BiCGStab iterations) and the solver's
exit code is the same."""
import os

import numpy as np
import pytest
import torch

from oracle import ips

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_ips_16.npz")


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def _run(pos, dx, h, thr):
    from diffsph_amd import _lib
    from diffsph_amd.shifting import ips_shift
    dev = torch.device("cuda:0")
    n = pos.shape[0]
    areas = torch.full((n,), float(dx) ** 2, device=dev)
    dom = _lib.make_domain((-1, -1), (1, 1), (True, True))
    return ips_shift(torch.from_numpy(pos).to(dev), areas, float(h), dom, threshold=thr)


@pytest.mark.parametrize("label,thr", [("default", -0.5), ("solve", 1.0)])
def test_ips_matches_reference(label, thr):
    g = dict(np.load(GOLD))
    print(f"case {label}", flush=True)
    r = _run(g["pos"], g["dx"], g["h"], thr)
    assert r.num_pairs == g["i"].shape[0]
    assert _rel(r.rhs.cpu().numpy(), g[f"{label}_B"]) <= 1e-5
    assert r.code == int(g[f"{label}_iters"])
    e = _rel(r.update.cpu().numpy(), g[f"{label}_update"])
    print(f"update rel err {e:.2e}, residual {r.residual:.3e} (reference {float(g[f'{label}_residual']):.3e})",
          flush=True)
    assert e <= 5e-5
    assert r.residual == pytest.approx(float(g[f"{label}_residual"]), rel=1e-3)


def test_ips_matches_oracle_64(orc):
    from diffsph_amd import sampling
    d = sampling.dfsph_random_flow(64, seed=9, jitter_width=0.4)
    pos, h, dx = d["pos"], np.float32(d["h"]), np.float32(d["dx"])
    n = pos.shape[0]
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    _, off, nbr = orc.neighbors_brute(dom, pos, pos, h0=float(h))
    cnt = np.diff(off)
    i = np.repeat(np.arange(n), cnt).astype(np.int64)
    j = nbr.astype(np.int64)
    x = pos[i] - pos[j]
    x = np.where(np.abs(x) > 1.0, np.sign(x) * ((np.abs(x) - 1.0) % 2.0 - 1.0), x).astype(np.float32)
    r = np.sqrt((x * x).sum(1)).astype(np.float32)
    xhat = (x / (r + np.float32(1e-7))[:, None]).astype(np.float32)
    rij = (r / h).astype(np.float32)
    q = rij.astype(np.float64)
    dw = (-20.0 * q * np.clip(1 - q, 0, 1) ** 3) * ips.C2D / float(h) ** 3
    grad = (xhat * dw[:, None]).astype(np.float32)
    areas = np.full(n, float(dx) ** 2, np.float32)
    for thr in (-0.5, 1.0):
        upd, B, code, res = ips.compute_shifting(i, j, rij, xhat, np.full_like(rij, h), grad, areas, n,
                                                 threshold=thr)
        got = _run(pos, dx, h, thr)
        print(f"threshold {thr}: code {got.code} / {code}, residual {got.residual:.3e} / {res:.3e}", flush=True)
        assert got.code == code
        # B sums ~50 terms that cancel to a few percent of their size (float32
        # on the GPU, float64 in the oracle)
        assert _rel(got.rhs.cpu().numpy(), B) <= 5e-5
        assert _rel(got.update.cpu().numpy(), upd) <= 2e-4


def test_solve_shifting_ips_moves_toward_the_lattice():
    """solveShifting with scheme 'IPS': one clamped shift applied and
    returned as the displacement; a jittered lattice gets more uniform (the
    summation-density spread shrinks)."""
    from diffsph_amd import sampling
    from diffsph_amd.neighborhood import neighborSearch
    from diffsph_amd.shifting import solveShifting
    from diffsph_amd.sphops import sphOperationStates
    dev = torch.device("cuda:0")
    d = sampling.dfsph_random_flow(64, seed=4, jitter_width=0.4)
    n = d["pos"].shape[0]
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    f = {"positions": T(d["pos"]), "velocities": T(d["vel"]), "masses": T(d["mass"]), "areas": T(d["area"]),
         "densities": torch.ones(n, device=dev), "supports": torch.full((n,), float(d["h"]), device=dev),
         "numParticles": n}
    cfg = {"shifting": {"scheme": "IPS", "maxIterations": 1, "threshold": 0.5, "solverThreshold": 1.0,
                        "rtol": 1e-4, "maxSolveIter": 64},
           "particle": {"dx": float(d["dx"])}, "boundary": {"active": False},
           "domain": {"minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0], "periodicity": [True, True], "dim": 2},
           "neighborhood": {"verletScale": 1.0}, "kernel": {"kernelScale": 1.0}}
    before = f["positions"].clone()
    dx, states = solveShifting({"fluid": f}, cfg)
    assert torch.equal(f["positions"], before)  # solveShifting restores the positions
    assert float(dx.abs().max()) <= 0.5 * float(d["dx"]) * (1 + 1e-6)
    # one solve; code = iterations or 0 (converged at the half step), never a
    # breakdown (-10 / -11) or the solver threshold (-12)
    conv, code, res = states[0]
    assert len(states) == 1 and code >= 0 and len(conv) > 0 and res < conv[0]

    def spread(p):
        st = dict(f, positions=p)
        _, nd = neighborSearch(st, st, cfg)
        rho = sphOperationStates(st, st, None, nd, operation="density")
        return float(rho.std() / rho.mean())

    s0, s1 = spread(f["positions"]), spread(f["positions"] + dx)
    print(f"density spread {s0:.3e} -> {s1:.3e}", flush=True)
    assert s1 < 0.7 * s0
