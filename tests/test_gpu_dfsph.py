"""GPU parity of the DFSPH step (v2/simulationSchemes/dfsph.py) vs the oracle.

Both keep the state in cell order, so every neighbor list is walked in the
same order; per-step fields must agree within 1e-5 relative (float32), and
the iteration counts of both solves must be identical.  The float64 oracle
gives the error of the float32 GPU fields, which is reported.
"""
import numpy as np
import pytest
import torch

from diffsph_amd import sampling
from diffsph_amd.dfsph import DFSPHConfig, DFSPHSolver

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _rel(gpu, ref):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(gpu - ref).max() / max(np.abs(ref).max(), 1e-30)


def _setup(nside, seed, jit=0.1):
    d = sampling.dfsph_random_flow(nside, seed=seed, jitter_width=jit)
    dt = 3.0 * d["dx"] / 10.0  # acoustic CFL 1.5 with c_s = 10 |u|max
    return d, dt


def _solver(d, dt, **kw):
    dev = torch.device("cuda:0")
    cfg = DFSPHConfig(dt=dt, support=d["h"], **kw)
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    return DFSPHSolver(T(d["pos"]), T(d["vel"]), T(d["mass"]), T(d["area"]), cfg)


@pytest.mark.parametrize("nside,seed,steps", [(64, 0, 3), (128, 1, 2)])
def test_dfsph_steps_match_oracle(orc, nside, seed, steps):
    d, dt = _setup(nside, seed)
    print(f"case nside={nside} seed={seed} steps={steps}", flush=True)
    s = _solver(d, dt)
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    n = d["pos"].shape[0]
    ost = orc.DfsphState(d["pos"].copy(), d["vel"].copy(), d["mass"].copy(), d["area"].copy(),
                         np.arange(n, dtype=np.int32))
    for k in range(steps):
        stats = s.step()
        ost, info = orc.dfsph_step(ost, dom, d["h"], dt)
        # same cell order on both sides
        np.testing.assert_array_equal(s.uid.cpu().numpy(), info["uid"])
        assert stats.num_pairs == info["num_pairs"]
        assert stats.div_iters == info["it_div"], (stats.div_iters, info["it_div"])
        assert stats.den_iters == info["it_den"], (stats.den_iters, info["it_den"])
        for name, gpu, ref in [("rho", s.rho, info["rho"]), ("alpha_den", s.alpha_den, info["alpha_den"]),
                               ("visc", s.visc, info["visc"]), ("p_div", s.p_div, info["p_div"]),
                               ("p_den", s.p_den, info["p_den"]), ("a_den", s.a_den, info["a_den"]),
                               ("vel", s.vel, ost.vel), ("pos", s.pos, ost.pos)]:
            e = _rel(gpu.cpu().numpy(), ref)
            print(f"step {k} {name}: rel err vs oracle {e:.2e}", flush=True)
            assert e <= RTOL, (k, name, e)
        assert abs(stats.den_error - info["err_den"]) <= 1e-6 * max(1e-12, abs(info["err_den"])) + 1e-9


def test_dfsph_1m_steps_match_oracle(orc):
    """The 1M-particle headline workload (1024^2, seed 0) through the HIP
    path against the oracle: the first step (p = 0 warm start) and the
    second (0.75 p_prev warm start; the divergence solve runs to maxIters
    there), with identical pair counts, iteration counts and cell order and
    every field within 1e-5 relative."""
    d, dt = _setup(1024, 0)
    print("case 1024^2 seed 0, 2 steps", flush=True)
    s = _solver(d, dt)
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    n = d["pos"].shape[0]
    ost = orc.DfsphState(d["pos"].copy(), d["vel"].copy(), d["mass"].copy(), d["area"].copy(),
                         np.arange(n, dtype=np.int32))
    for k in range(2):
        stats = s.step()
        ost, info = orc.dfsph_step(ost, dom, d["h"], dt)
        print(f"step {k}: pairs {stats.num_pairs}/{info['num_pairs']} div it {stats.div_iters}/{info['it_div']} "
              f"den it {stats.den_iters}/{info['it_den']}", flush=True)
        np.testing.assert_array_equal(s.uid.cpu().numpy(), info["uid"])
        assert stats.num_pairs == info["num_pairs"]
        assert stats.div_iters == info["it_div"] and stats.den_iters == info["it_den"]
        for name, gpu, ref in [("rho", s.rho, info["rho"]), ("alpha_den", s.alpha_den, info["alpha_den"]),
                               ("visc", s.visc, info["visc"]), ("p_div", s.p_div, info["p_div"]),
                               ("p_den", s.p_den, info["p_den"]), ("a_den", s.a_den, info["a_den"]),
                               ("vel", s.vel, ost.vel), ("pos", s.pos, ost.pos)]:
            e = _rel(gpu.cpu().numpy(), ref)
            print(f"  {name}: rel err vs oracle {e:.2e}", flush=True)
            assert e <= RTOL, (k, name, e)


def test_dfsph_float64_error_reported(orc):
    """Error of the float32 GPU step against a float64 run of the same ops."""
    d, dt = _setup(64, 3)
    s = _solver(d, dt)
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    n = d["pos"].shape[0]
    ost = orc.DfsphState(d["pos"].copy(), d["vel"].copy(), d["mass"].copy(), d["area"].copy(),
                         np.arange(n, dtype=np.int32))
    s.step()
    _, info = orc.dfsph_step(ost, dom, d["h"], dt, dtype=np.float64)
    errs = {k: _rel(v.cpu().numpy(), info[r]) for k, v, r in [("density", s.rho, "rho"), ("pressure", s.p_den, "p_den"),
                                                           ("accel", s.a_den, "a_den")]}
    print("float32 GPU vs float64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert errs["density"] < 1e-5
    # pressure from an iterative solve: iteration counts may differ by one
    assert errs["pressure"] < 5e-2


def test_dfsph_1m_step_properties():
    """The 1M-particle workload over 12 steps: the density solve converges
    every step; the divergence solve (relaxed Jacobi, as the reference) may
    stop at maxIters on the first steps but never diverges.  After the
    start-up transient the mean |rho/rho0 - 1| of the summation density is
    within the headline's 0.1 %.  The largest deviation is an expansion
    (rho < rho0) left by the jittered lattice, which the clamped pressure
    (p >= 0) cannot correct; compression stays within 0.5 %."""
    d, dt = _setup(1024, 0)
    s = _solver(d, dt)
    for k in range(12):
        st = s.step()
        print(f"1M step {k}: div it {st.div_iters} err {st.div_error:.3e}; den it {st.den_iters} "
              f"err {st.den_error:.3e}; pairs {st.num_pairs}", flush=True)
        assert st.den_iters < 256 and st.den_error <= 5e-4
        assert st.div_iters <= 256 and np.isfinite(st.div_error) and st.div_error < 1.0
    mean_e, max_e = s.density_error()
    e = (s.rho - 1.0)
    print(f"density error mean {mean_e:.3e} max {max_e:.3e}; compression max {float(e.max()):.3e}, "
          f"expansion max {float(-e.min()):.3e}", flush=True)
    assert mean_e <= 1e-3
    assert float(e.max()) <= 5e-3 and max_e <= 5e-2
    # UID permutation is a bijection
    assert torch.equal(torch.sort(s.uid.long()).values, torch.arange(s.n, device=s.uid.device))


def test_dfsph_solver_limits_and_nan_guard():
    """maxIters caps the loop (reference dfsph.py:88); minIters=0 with error 0
    runs no iteration, as the reference's while-condition does."""
    d, dt = _setup(48, 5)
    s = _solver(d, dt, max_iters=2, threshold_den=0.0, threshold_div=0.0)
    st = s.step()
    assert st.den_iters == 2 and st.div_iters == 2
    s2 = _solver(d, dt, min_iters=0)
    st2 = s2.step()
    assert st2.den_iters == 0 and st2.div_iters == 0


def test_compact16_deltas_decode_and_overflow():
    """The 16-bit modular deltas written by the fill pass decode to the
    int32 list entry for entry and equal the separate compaction pass; a
    limit below the largest delta raises the device flag (a list without
    neighbor tiles: with them the fill writes tile slots instead)."""
    d, dt = _setup(64, 2)
    s = _solver(d, dt, tile_cap=0)
    s.neighbor_step()
    nl = s.nl
    assert nl.nbr16 is not None and int(nl.nbr16_overflow.item()) == 0
    cnt = nl.counts.cpu().numpy()
    so = nl.slice_off.cpu().numpy()
    nbr = nl.nbr.cpu().numpy()
    d16 = nl.nbr16.cpu().numpy().astype(np.int64)
    buf = torch.empty_like(nl.nbr16)
    flag = torch.zeros(1, dtype=torch.int32, device=buf.device)
    fused, fused_flag = nl.nbr16, nl.nbr16_overflow
    nl.attach_compact16(buf, flag)
    assert int(flag.item()) == 0
    d16b = buf.cpu().numpy().astype(np.int64)
    n = s.n
    dmax = 0
    for i in range(n):
        k = so[i // 64] + 64 * np.arange(cnt[i]) + i % 64
        j = (i + d16[k]) % n
        np.testing.assert_array_equal(j, nbr[k])
        np.testing.assert_array_equal(d16b[k], d16[k])
        dmax = max(dmax, int(np.abs(d16[k]).max()))
    assert 0 < dmax <= n // 2
    nl.attach_compact16(buf, flag, limit=dmax - 1)
    assert int(flag.item()) == 1
    assert int(fused_flag.item()) == 0 and fused is not buf


def test_record_sweeps_match_separate_arrays():
    """The solve loop's sweeps on (x, y, p) / (x, y, a) records give the
    accelerations, pressures and residuals of the separate-array entry
    points bit for bit, and the records carry the positions."""
    import ctypes as C
    from diffsph_amd import _lib
    d, dt = _setup(64, 2)
    s = _solver(d, dt)
    s.step()
    L = _lib.lib()
    pr = s.nl.pairs()
    dev = s.pos.device
    st = _lib.stream_handle(dev)
    n = s.n
    p0 = torch.rand(n, device=dev)
    acc = torch.empty(n, 2, device=dev)
    _lib.check(L.dsph_pressure_accel(C.byref(pr), _lib.ptr(s.kgrad), _lib.ptr(s.vol), _lib.ptr(s.rho),
                                     _lib.ptr(p0), _lib.ptr(acc), st))
    xp = torch.empty(n, 4, device=dev)
    xa = torch.empty(n, 4, device=dev)
    _lib.check(L.dsph_pack_pressure(C.byref(pr), _lib.ptr(p0), _lib.ptr(xp), st))
    _lib.check(L.dsph_pressure_accel_rec(C.byref(pr), _lib.ptr(s.kgrad), _lib.ptr(s.rho), _lib.ptr(xp),
                                         _lib.ptr(xa), st))
    assert torch.equal(xa[:, 2:], acc) and torch.equal(xa[:, :2], s.grid.sorted_pos)
    prm = _lib.DfsphParams()
    prm.dt, prm.rho0, prm.omega, prm.error_threshold = dt, 1.0, 0.3, 5e-4
    prm.min_iters, prm.max_iters, prm.clamp_pressure, prm.source_density = 1, 1, 1, 0
    nb = (n + 255) // 256
    p1, p2 = p0.clone(), p0.clone()
    r1, r2 = torch.empty(n, device=dev), torch.empty(n, device=dev)
    b1 = torch.empty(nb, dtype=torch.float64, device=dev)
    b2 = torch.empty(nb, dtype=torch.float64, device=dev)
    _lib.check(L.dsph_pressure_update(C.byref(pr), C.byref(prm), _lib.ptr(s.kgrad), _lib.ptr(s.vol),
                                      _lib.ptr(s.div_src), _lib.ptr(s.alpha_div), _lib.ptr(acc), _lib.ptr(p1),
                                      _lib.ptr(r1), _lib.ptr(b1), st))
    _lib.check(L.dsph_pressure_update_rec(C.byref(pr), C.byref(prm), _lib.ptr(s.kgrad), _lib.ptr(s.div_src),
                                          _lib.ptr(s.alpha_div), _lib.ptr(xa), _lib.ptr(p2), _lib.ptr(xp),
                                          _lib.ptr(r2), _lib.ptr(b2), st))
    assert torch.equal(p1, p2) and torch.equal(r1, r2) and torch.equal(b1, b2)
    assert torch.equal(xp[:, 2], p2) and torch.equal(xp[:, :2], s.grid.sorted_pos)


def test_dfsph_compact_index_paths_identical():
    """16-bit deltas, the int32 list, the device-side fallback after an
    overflow, the two-pass neighbor list and the one-pass list's fallback
    (then widened) give the same trajectory bit for bit."""
    d, dt = _setup(96, 4)
    runs = [_solver(d, dt), _solver(d, dt, compact_index=False), _solver(d, dt, compact_limit=2),
            _solver(d, dt, neighbor_width=0), _solver(d, dt, neighbor_width=8)]
    for k in range(4):
        its = [r.step() for r in runs]
        assert len({(st.div_iters, st.den_iters) for st in its}) == 1, k
        assert len({st.num_pairs for st in its}) == 1, k
    assert int(runs[2].nl.nbr16_overflow.item()) == 1
    assert runs[0].nl.width == 67 and runs[3].nl.width == 0
    assert runs[4]._nbr_width > 8 and runs[4].nl.width == runs[4]._nbr_width
    for r in runs[1:]:
        assert torch.equal(runs[0].pos, r.pos) and torch.equal(runs[0].vel, r.vel)
        assert torch.equal(runs[0].p_div, r.p_div) and torch.equal(runs[0].p_den, r.p_den)


@pytest.mark.skip(reason="the persistent (cooperative-launch) solve is disabled: it matched the chunked solve "
                         "bit for bit at 96x96 but did not return at 1024x1024; cause not found")
@pytest.mark.parametrize("nside,steps", [(96, 3), (1024, 2)])
def test_persistent_solve_matches_chunked(nside, steps):
    """The solves as one cooperative launch each (loop state, residual and
    stopping rule on the device) give the same iterations and bit-identical
    fields as the chunked host-checked loop."""
    d, dt = _setup(nside, 5)
    print(f"case nside={nside}", flush=True)
    a = _solver(d, dt)
    b = _solver(d, dt, persistent=True)
    for k in range(steps):
        sa, sb = a.step(), b.step()
        print(f"step {k}: div {sa.div_iters}/{sb.div_iters} den {sa.den_iters}/{sb.den_iters}", flush=True)
        assert (sa.div_iters, sa.den_iters) == (sb.div_iters, sb.den_iters)
        assert sa.den_error == sb.den_error and sa.div_error == sb.div_error
        for name in ("p_div", "p_den", "a_div", "a_den", "vel", "pos", "rho"):
            assert torch.equal(getattr(a, name), getattr(b, name)), name
