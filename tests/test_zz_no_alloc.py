"""The allocation-free DFSPH solve (dsph_dfsph_solve_noalloc, scratch in a
static device array) against the workspace entry point dsph_dfsph_solve on
the same inputs: pressures, accelerations, residuals and solve statistics
bit for bit, on two calls in a row with the second on a smaller input (the
scratch keeps the first call's contents), both solve kinds; sizes beyond the
scratch are rejected before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _setup(nside, seed):
    from diffsph_amd import sampling
    from diffsph_amd.dfsph import DFSPHConfig, DFSPHSolver
    dev = torch.device("cuda:0")
    d = sampling.dfsph_random_flow(nside, seed=seed, jitter_width=0.1)
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    s = DFSPHSolver(T(d["pos"]), T(d["vel"]), T(d["mass"]), T(d["area"]),
                    DFSPHConfig(dt=3.0 * d["dx"] / 10.0, support=d["h"]))
    s.step()
    s.step()  # the lists, coefficients, densities and alphas of this step feed the solves below
    return s


def _solve(s, density, noalloc):
    from diffsph_amd import _lib
    from diffsph_amd.neighborhood import WORKSPACE
    L = _lib.lib()
    cfg = s.cfg
    prm = _lib.DfsphParams()
    prm.dt, prm.rho0 = cfg.dt, cfg.rho0
    prm.omega = cfg.omega_den if density else cfg.omega_div
    prm.error_threshold = (cfg.threshold_den if density else cfg.threshold_div) / cfg.rho0
    prm.min_iters, prm.max_iters = cfg.min_iters, cfg.max_iters
    prm.clamp_pressure, prm.source_density, prm.check_every = 1, 1 if density else 0, 8
    pairs = s.nl.pairs()
    p = (0.75 * (s.p_den if density else s.p_div)).contiguous()
    acc = torch.full((s.n, 2), float("nan"), device=p.device)
    res = torch.full((s.n,), float("nan"), device=p.device)
    stats = _lib.SolveStats()
    alpha = s.alpha_den if density else s.alpha_div
    adv = s.adv if density else None
    P = _lib.ptr
    args = [C.byref(pairs), C.byref(prm), P(s.kgrad), P(s.vel), P(adv), P(s.rho), P(s.vol), P(s.areas), P(alpha),
            P(s.div_src), P(p), P(acc), P(res), C.byref(stats)]
    st = _lib.stream_handle(p.device)
    if noalloc:
        _lib.check(L.dsph_dfsph_solve_noalloc(*args, st), "dfsph_solve_noalloc")
    else:
        ws = WORKSPACE.get(L.dsph_dfsph_workspace(s.n), p.device)
        _lib.check(L.dsph_dfsph_solve(*args, P(ws), ws.numel(), st), "dfsph_solve")
    torch.cuda.synchronize()
    return p.cpu().numpy(), acc.cpu().numpy(), res.cpu().numpy(), (stats.iterations, stats.converged,
                                                                    stats.final_error, stats.initial_error)


@pytest.mark.parametrize("density", [False, True])
def test_noalloc_solve_matches_workspace_solve(density):
    for nside, seed in [(128, 3), (80, 5)]:  # second call smaller: stale scratch contents must not matter
        print(f"case nside={nside} seed={seed} density={density}", flush=True)
        s = _setup(nside, seed)
        ref = _solve(s, density, noalloc=False)
        got = _solve(s, density, noalloc=True)
        assert ref[3] == got[3], (ref[3], got[3])
        assert ref[3][0] >= 1
        for a, b in zip(ref[:3], got[:3]):
            np.testing.assert_array_equal(a, b)


def test_noalloc_rejects_sizes_beyond_the_scratch():
    from diffsph_amd import _lib
    L = _lib.lib()
    nmax = L.dsph_dfsph_solve_noalloc_max_particles()
    assert nmax >= 1_048_576  # the 1M workload fits
    assert L.dsph_dfsph_workspace(nmax) <= L.dsph_noalloc_scratch_bytes() < L.dsph_dfsph_workspace(nmax + 1)
    pairs = _lib.Pairs()
    pairs.nq = pairs.nr = nmax + 1
    prm = _lib.DfsphParams()
    # rejected on the host before any pointer is used
    assert L.dsph_dfsph_solve_noalloc(C.byref(pairs), C.byref(prm), *([None] * 11), None, None) == 5
