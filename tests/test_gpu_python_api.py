"""The reference's Python-level operator interface on the GPU, driven with
reference-shaped state dicts and neighborDicts:

* sphOperation / sphOperationStates in every operation and mode against the
  reference's own outputs (tests/golden/reference_sphops_16.npz, written by
  v2/sphOps.py) on the neighborDict our neighborSearch drop-in builds;
* dfsphSolve (v2/modules/dfsph.py:52-109) against the reference's alpha,
  source, one Jacobi update and pressure acceleration
  (tests/golden/reference_modules_16.npz);
* the DFSPH scheme's simulationStep against the oracle (bit-exact: it runs
  the fused kernels) and the delta-SPH scheme's simulationStep against the
  reference's module outputs and the oracle;
* searchNeighbors and callModule.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 2e-5


def _close(a, b, tol=TOL):
    a = np.asarray(a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a, np.float64)
    b = np.asarray(b, np.float64)
    a = a.reshape(b.shape)
    err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
    assert err <= tol, f"norm-relative error {err:.3e} > {tol}"


def _config(h, dt=1e-3, verlet=1.5):
    # the candidate grid needs 3 cells of h * verletScale per periodic axis:
    # the 16x16 golden sets (h = 0.5 on a side of 2) search with verletScale 1
    return {"domain": {"minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0], "periodicity": [True, True],
                       "dim": 2},
            "neighborhood": {"verletScale": verlet}, "boundary": {"active": False},
            "fluid": {"rho0": 1.0, "cs": 10.0}, "timestep": {"dt": dt},
            "dfsph": {"minIters": 2, "maxIters": 256, "clampPressure": True, "omega": 0.3,
                      "sourceTerm": "density", "errorThreshold": 5e-4},
            "diffusion": {"velocityScheme": "deltaSPH_inviscid", "alpha": 0.01, "nu": 0.01, "eps": 1e-6,
                          "pi-switch": False, "delta": 0.1, "densityScheme": "deltaSPH"},
            "gravity": {"active": False}, "EOS": {"type": "isoThermal"},
            "kernel": {"kernelScale": 1.0}, "particle": {"support": h}}


def _golden_state(name):
    d = dict(np.load(os.path.join(GOLD, name)))
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    n = d["pos"].shape[0]
    st = {"positions": T(d["pos"]), "masses": T(d["mass"]), "densities": T(d["rho"]),
          "supports": torch.full((n,), float(d["h"]), device=dev), "numParticles": n}
    if "vel" in d:
        st["velocities"] = T(d["vel"])
    return d, st, T


def _neighborhood(st, d):
    from diffsph_amd.neighborhood import neighborSearch
    _, nd = neighborSearch(st, st, _config(float(d["h"]), verlet=1.0), computeKernels=True)
    # the pair set is the reference's (brute force r/h <= 1), row-sorted
    i, j = (t.cpu().numpy() for t in nd["indices"])
    order = np.lexsort((j, i))
    assert np.array_equal(i[order], d["i"]) and np.array_equal(j[order], d["j"])
    assert np.all(np.diff(i) >= 0)
    return nd


@pytest.fixture(scope="module")
def ops():
    d, st, T = _golden_state("reference_sphops_16.npz")
    nd = _neighborhood(st, d)
    return d, st, T, nd


def _op(ops_, q, operation, gm="symmetric", dm="div", lap=None):
    from diffsph_amd.sphops import sphOperation
    d, st, T, nd = ops_
    m, r = st["masses"], st["densities"]
    return sphOperation((m, m), (r, r), q, nd["indices"], nd["kernels"], nd["gradients"], nd["distances"],
                        nd["vectors"], nd["supports"], st["numParticles"], operation=operation, gradientMode=gm,
                        divergenceMode=dm, kernelLaplacians=lap)


def _pair_lap(ops_):
    """The reference's per-pair kernel Laplacians (fixture order: i, then j)
    re-ordered into our neighborDict's pair order."""
    d, st, T, nd = ops_
    i, j = (t.cpu().numpy() for t in nd["indices"])
    n = st["numParticles"]
    key_g = d["i"].astype(np.int64) * n + d["j"]
    pos = np.searchsorted(key_g, i.astype(np.int64) * n + j)
    return T(d["laplacian_kernels"][pos].astype(np.float32))


def test_sph_operation_all_modes_match_reference(ops):
    d, st, T, nd = ops
    S, V = (T(d["qs"]), T(d["qs"])), (T(d["qv"]), T(d["qv"]))
    _close(_op(ops, S, "interpolate"), d["interp_scalar"])
    _close(_op(ops, V, "interpolate"), d["interp_vector"])
    for gm in ("naive", "difference", "summation", "symmetric"):
        print("mode", gm, flush=True)
        _close(_op(ops, S, "gradient", gm), d[f"grad_{gm}_scalar"])
        _close(_op(ops, V, "gradient", gm), d[f"grad_{gm}_vector"])
        _close(_op(ops, V, "divergence", gm), d[f"div_{gm}"])
        _close(_op(ops, V, "divergenceConsistent", gm), d[f"divc_{gm}"])
        _close(_op(ops, V, "curl", gm), d[f"curl_{gm}"])
    TT = (T(d["qt"]), T(d["qt"]))
    _close(_op(ops, TT, "gradient", "difference"), d["grad_difference_tensor"])
    _close(_op(ops, TT, "divergence", "difference", "div"), d["div_difference_tensor_div"])
    _close(_op(ops, TT, "divergence", "difference", "dot"), d["div_difference_tensor_dot"])
    lap = _pair_lap(ops)
    _close(_op(ops, S, "laplacian", "naive", lap=lap), d["lap_naive_scalar"])
    _close(_op(ops, V, "laplacian", "naive", lap=lap), d["lap_naive_vector"])
    for gm in ("difference", "conserving", "divergenceFree", "dot"):
        print("laplacian", gm, flush=True)
        _close(_op(ops, S, "laplacian", gm), d[f"lap_{gm}_scalar"])
        _close(_op(ops, V, "laplacian", gm), d[f"lap_{gm}_vector"])
    _close(_op(ops, S, "directLaplacian", "difference"), d["direct_laplacian_difference"])


def test_sph_operation_states_and_per_pair_quantity(ops):
    from diffsph_amd.sphops import sphOperation, sphOperationStates
    d, st, T, nd = ops
    S = (T(d["qs"]), T(d["qs"]))
    _close(sphOperationStates(st, st, None, nd, operation="density"), d["states_density"])
    _close(sphOperationStates(st, st, S, nd, operation="gradient", gradientMode="difference"),
           d["states_grad_difference"])
    # a per-pair quantity (the reference's single-tensor form) equal to q_i - q_j
    i, j = nd["indices"]
    qij = S[0][i] - S[1][j]
    m, r = st["masses"], st["densities"]
    out = sphOperation((m, m), (r, r), qij, nd["indices"], nd["kernels"], nd["gradients"], nd["distances"],
                       nd["vectors"], nd["supports"], st["numParticles"], "gradient", "difference")
    _close(out, d["grad_difference_scalar"])


def test_sph_operation_rejects_unsorted_rows(ops):
    from diffsph_amd._lib import DsphError
    from diffsph_amd.sphops import sphOperation
    d, st, T, nd = ops
    i, j = nd["indices"]
    perm = torch.randperm(i.numel(), device=i.device)
    m, r = st["masses"], st["densities"]
    with pytest.raises(DsphError, match="sorted"):
        sphOperation((m, m), (r, r), None, (i[perm].contiguous(), j[perm].contiguous()), nd["kernels"][perm], None,
                     None, None, None, st["numParticles"], "density")


@pytest.mark.parametrize("solve", ["div", "den"])
def test_dfsph_solve_matches_reference(solve):
    """One Jacobi iteration of the reference's dfsphSolve from the golden
    warm start: alpha, source, updated pressure, residual and the final
    pressure acceleration."""
    from diffsph_amd.dfsph import dfsphSolve
    d, st, T = _golden_state("reference_modules_16.npz")
    nd = _neighborhood(st, d)
    dt = float(d["dfsph_dt"])
    st["areas"] = T(d["mass"])
    dens = solve == "den"
    st["advection"] = T(d["dfsph_adv"]) if dens else torch.zeros_like(st["velocities"])
    p0 = T(d["dfsph_p0"])
    st["pressureIncompressible" if dens else "pressureDivergence"] = p0 / 0.75
    cfg = _config(float(d["h"]), dt, verlet=1.0)
    cfg["dfsph"].update(sourceTerm="density" if dens else "divergence", minIters=1, maxIters=1,
                        errorThreshold=5e-4 if dens else 1e-2)
    a_p = dfsphSolve(st, st, nd, cfg)
    _close(st["alpha"], d[f"dfsph_{solve}_alpha"])
    _close(st["sourceTerm"], d[f"dfsph_{solve}_source"])
    _close(st["pressureB"], d[f"dfsph_{solve}_p1"])
    _close(st["residual"], d[f"dfsph_{solve}_residual"])
    _close(a_p, d[f"dfsph_{solve}_accel"])
    assert len(st[f"convergence_{'density' if dens else 'divergence'}"]) == 1
    # the convergence metric of the reference: mean(clamp(residual / rho0, min=-thr))
    thr = 5e-4 if dens else 1e-2
    ref_err = float(np.mean(np.maximum(d[f"dfsph_{solve}_residual"].astype(np.float64), -thr)))
    got = st[f"convergence_{'density' if dens else 'divergence'}"][0]
    assert abs(got - ref_err) <= 1e-5 * max(abs(ref_err), 1e-3)


def test_dfsph_solve_runs_to_the_stopping_rule():
    """Full dfsphSolve loops stop by the reference's rule (minIters, then
    error <= threshold or maxIters) and agree with the fused solver's
    iteration count and pressure."""
    from diffsph_amd import sampling
    from diffsph_amd.dfsph import DFSPHConfig, DFSPHSolver, dfsphSolve
    from diffsph_amd.neighborhood import neighborSearch
    dev = torch.device("cuda:0")
    dd = sampling.dfsph_random_flow(48, seed=2, jitter_width=0.1)
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    n = dd["pos"].shape[0]
    dt = 3.0 * dd["dx"] / 10.0
    st = {"positions": T(dd["pos"]), "velocities": T(dd["vel"]), "masses": T(dd["mass"]), "areas": T(dd["area"]),
          "supports": torch.full((n,), float(dd["h"]), device=dev), "numParticles": n}
    cfg = _config(float(dd["h"]), dt)
    _, nd = neighborSearch(st, st, cfg)
    from diffsph_amd.sphops import sphOperationStates
    st["densities"] = sphOperationStates(st, st, None, nd, operation="density")
    st["advection"] = torch.zeros_like(st["velocities"])
    cfg["dfsph"].update(sourceTerm="divergence", errorThreshold=1e-2, minIters=2, maxIters=64)
    dfsphSolve(st, st, nd, cfg)
    errs = st["convergence_divergence"]
    assert 2 <= len(errs) <= 64 and (errs[-1] <= 1e-2 or len(errs) == 64)
    assert all(e > 1e-2 for e in errs[1:-1])
    s = DFSPHSolver(T(dd["pos"]), T(dd["vel"]), T(dd["mass"]), T(dd["area"]),
                    DFSPHConfig(dt=dt, support=float(dd["h"]), max_iters=64))
    stats = s.step()
    assert abs(stats.div_iters - len(errs)) <= 1
    if stats.div_iters == len(errs):
        order = s.uid.long()
        _close(st["pressureB"][order], s.p_div.cpu().numpy(), tol=1e-4)


def test_dfsph_simulation_step_matches_oracle(orc):
    """The DFSPH scheme's simulationStep on a reference-shaped state: dudt,
    densities, both pressures and the warm start of a second step equal the
    oracle's DFSPH step (bit for bit: the fused kernels run)."""
    from diffsph_amd import sampling
    from diffsph_amd.schemes import dfsphSimulationStep
    dev = torch.device("cuda:0")
    dd = sampling.dfsph_random_flow(64, seed=4, jitter_width=0.1)
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    n = dd["pos"].shape[0]
    dt = 3.0 * dd["dx"] / 10.0
    state = {"fluid": {"positions": T(dd["pos"]), "velocities": T(dd["vel"]), "masses": T(dd["mass"]),
                       "areas": T(dd["area"]), "densities": T(dd["mass"] / dd["area"]),
                       "supports": torch.full((n,), float(dd["h"]), device=dev), "numParticles": n}}
    cfg = _config(float(dd["h"]), dt)
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    ost = orc.DfsphState(dd["pos"].copy(), dd["vel"].copy(), dd["mass"].copy(), dd["area"].copy(),
                         np.arange(n, dtype=np.int32))
    for k in range(2):
        print("step", k, flush=True)
        (dxdt, dudt, drhodt), _ = dfsphSimulationStep(state, cfg)
        assert drhodt is None
        ost_new, info = orc.dfsph_step(ost, dom, dd["h"], dt)
        uid = info["uid"].astype(np.int64)
        f = state["fluid"]
        ref_dudt = (info["a_den"] + info["visc"]) + info["a_div"]
        np.testing.assert_array_equal(dudt.cpu().numpy()[uid], ref_dudt)
        np.testing.assert_array_equal(f["densities"].cpu().numpy()[uid], info["rho"])
        np.testing.assert_array_equal(f["pressureIncompressible"].cpu().numpy()[uid], info["p_den"])
        np.testing.assert_array_equal(f["pressureDivergence"].cpu().numpy()[uid], info["p_div"])
        assert f["convergence_iterations"] == {"divergence": info["it_div"], "density": info["it_den"]}
        # integrate as the reference's integrate does (semi-implicit Euler) and
        # carry the oracle's state in input order
        v = f["velocities"] + dt * dudt
        f["velocities"] = v
        f["positions"] = f["positions"] + dt * v
        inv = np.argsort(uid)
        ost = orc.DfsphState(ost_new.pos[inv], ost_new.vel[inv], ost_new.mass[inv], ost_new.areas[inv],
                             np.arange(n, dtype=np.int32), ost_new.p_div[inv], ost_new.p_den[inv])
        np.testing.assert_array_equal(f["velocities"].cpu().numpy(), ost.vel)
    # the neighborDict is materialised on request with the reference's keys
    nd = state["fluid"]["neighborhood"]
    assert {"indices", "kernels", "gradients", "numNeighbors", "fullIndices"} <= set(nd.keys())
    assert int(nd["numNeighbors"].sum()) == nd["indices"][0].numel()


def test_deltaplus_simulation_step_matches_reference_modules():
    """The delta-SPH scheme's simulationStep (fluid only, continuity density)
    on the golden state: dudt = pressure accel + viscosity and drhodt =
    continuity + diffusion as the reference's modules computed them."""
    from diffsph_amd.schemes import deltaPlusSimulationStep
    d, st, T = _golden_state("reference_modules_16.npz")
    cfg = _config(float(d["h"]), verlet=1.0)
    cfg["diffusion"]["velocityScheme"] = "deltaSPH_viscid"
    cfg["kernel"]["kernelScale"] = float(d["kernel_scale"])
    cfg["particle"]["dx"] = float(d["dx"])
    state = {"fluid": st}
    (dxdt, dudt, drhodt), bnd = deltaPlusSimulationStep(state, cfg)
    assert bnd == (None, None, None)
    f = state["fluid"]
    _close(f["pressureAccel"], d["pressure_accel"])
    _close(f["velocityDiffusion"], d["visc_viscid"])
    _close(f["momentumEquation"], d["continuity"])
    _close(f["densityDiffusion"], d["diffusion"])
    _close(f["pressures"], d["pressure"])
    _close(f["gradRho^L"], d["grad_rho"])
    _close(f["L"], d["L"])
    _close(dudt, d["pressure_accel"] + d["visc_viscid"])
    _close(drhodt, d["continuity"] + d["diffusion"])
    _close(dxdt, d["vel"], tol=0.0)


def test_search_neighbors_and_call_module():
    from diffsph_amd.neighborhood import neighborSearch, searchNeighbors
    from diffsph_amd.schemes import callModule, computeDensity
    d, st, T = _golden_state("reference_sphops_16.npz")
    cfg = _config(float(d["h"]), verlet=1.0)
    state = {"fluid": st}
    searchNeighbors(state, cfg)
    f = state["fluid"]
    assert {"datastructure", "neighborhood", "numNeighbors"} <= set(f)
    _, nd = neighborSearch(st, st, cfg)
    assert torch.equal(f["neighborhood"]["indices"][1], nd["indices"][1])
    rho, none = callModule(state, computeDensity, cfg, "fluid")
    assert none is None
    _close(rho, d["states_density"])
    # a second search reuses the candidate list (Verlet), same pairs
    searchNeighbors(state, cfg)
    assert f["neighborhood"]["fullIndices"][0] is nd["fullIndices"][0] or \
        torch.equal(f["neighborhood"]["fullIndices"][0], nd["fullIndices"][0])
