"""The oracle against golden vectors written by the reference's own module
files (tests/golden/make_reference_golden.py): delta-SPH normalization,
density gradient, density diffusion, viscosity (viscid and inviscid),
continuity, EOS, pressure acceleration, the time step, and the DFSPH alpha,
sources and one Jacobi iteration of each solve, on a 16x16 jittered
periodic set.  The reference sums with torch.scatter_add (another order
than the oracle's row loop), so values agree to float32 rounding."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_modules_16.npz")
TOL = 2e-5


@pytest.fixture(scope="module")
def g():
    d = dict(np.load(GOLD))
    n = d["pos"].shape[0]
    cnt = np.bincount(d["i"], minlength=n)
    off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    d["off"], d["nbr"], d["n"] = off, d["j"].astype(np.int32), n
    d["dom"] = orc.make_domain((-1, -1), (1, 1), (True, True))
    return d


def _close(a, b, tol=TOL):
    a = np.asarray(a, np.float64).reshape(np.shape(b))
    b = np.asarray(b, np.float64)
    err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
    assert err <= tol, f"norm-relative error {err:.3e} > {tol}"


def test_golden_pairs_match_oracle_search(g):
    """The golden pair list (brute force r/h <= 1 with torch.linalg.norm) is
    the oracle's pair set."""
    _, off, nbr = orc.neighbors_brute(g["dom"], g["pos"], g["pos"], h0=float(g["h"]))
    assert np.array_equal(off, g["off"]) and np.array_equal(nbr, g["nbr"])


def test_golden_normalization_and_density_gradient(g):
    L, ev, grad = orc.deltasph_normalization(g["dom"], g["off"], g["nbr"], g["pos"], g["mass"], g["rho"],
                                             float(g["h"]))
    _close(L.reshape(-1, 2, 2), g["L"])
    _close(ev, g["L_ev"])
    _close(grad, g["grad_rho"])


@pytest.mark.parametrize("scheme,key", [(1, "visc_viscid"), (0, "visc_inviscid")])
def test_golden_rhs_terms(g, scheme, key):
    _, _, grad = orc.deltasph_normalization(g["dom"], g["off"], g["nbr"], g["pos"], g["mass"], g["rho"],
                                            float(g["h"]))
    prm = orc.DeltaParams(rho0=1.0, cs=10.0, delta=0.1, kernel_scale=float(g["kernel_scale"]), viscosity=scheme,
                          visc_coeff=0.01, eps=1e-6)
    dudt, drhodt, t = orc.deltasph_rhs(g["dom"], g["off"], g["nbr"], g["pos"], g["vel"], g["mass"], g["rho"], grad,
                                       prm, float(g["h"]))
    _close(t[:, 2:4], g[key])
    _close(t[:, 0:2], g["pressure_accel"])
    _close(t[:, 4], g["continuity"])
    _close(t[:, 5], g["diffusion"])
    _close(t[:, 6], g["pressure"])
    _close(dudt, g["pressure_accel"] + g[key])
    _close(drhodt, g["continuity"] + g["diffusion"])


def test_golden_timestep(g):
    tp = orc.TimestepParams(support=float(g["h"]), kernel_scale=float(g["kernel_scale"]), cs=10.0, nu=0.01)
    dudt = (g["pressure_accel"] + g["visc_viscid"]).astype(np.float64)
    # computeTimestep's "max" is over a scalar: torch.linalg.norm of the
    # flattened non-NaN components (timestep.py:22-25)
    amax = float(np.sqrt((dudt ** 2).sum()))
    assert orc.compute_timestep(tp, 0.001) == pytest.approx(float(g["dt_first"]), rel=1e-6)
    assert orc.compute_timestep(tp, 0.001, amax) == pytest.approx(float(g["dt_next"]), rel=1e-6)


@pytest.mark.parametrize("solve", ["div", "den"])
def test_golden_dfsph_alpha_source_and_iteration(g, solve):
    dt = float(g["dfsph_dt"])
    vol = (g["mass"] / g["rho"]).astype(np.float32)
    areas = g["mass"]
    ad, an, ds, _ = orc.dfsph_prepare(g["dom"], g["off"], g["nbr"], g["pos"], g["vel"], areas, vol, g["rho"], dt,
                                      1.0, None, float(g["h"]))
    dens = solve == "den"
    alpha = an if dens else ad
    _close(alpha, g[f"dfsph_{solve}_alpha"])
    if not dens:
        _close(ds, g["dfsph_div_source"])
    p1, acc, res, it, err, _ = orc.dfsph_solve(g["dom"], g["off"], g["nbr"], g["pos"], float(g["h"]), dt, 1.0, 0.3,
                                                1.0, 1, 1, True, dens, g["vel"], g["dfsph_adv"] if dens else None,
                                                g["rho"], vol, areas, alpha, ds, g["dfsph_p0"])
    assert it == 1
    _close(p1, g[f"dfsph_{solve}_p1"])
    _close(res, g[f"dfsph_{solve}_residual"], tol=5e-5)
    _close(acc, g[f"dfsph_{solve}_accel"])


def test_golden_rollout_forward_matches_oracle():
    """The oracle's float64 RK4 delta-SPH steps reproduce the reference
    composition's final state and kinetic energy (same fixed time steps)."""
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_rollout_grad_16.npz")))
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    n = g["pos"].shape[0]
    prm = orc.DeltaParams(rho0=1.0, cs=10.0, delta=0.1, kernel_scale=float(g["h"]) / (2 * float(g["dx"])),
                          viscosity=1, visc_coeff=0.01, eps=1e-6)
    st = orc.DeltaState(g["pos"].copy(), g["vel"].copy(), np.ones(n), g["mass"].copy(), np.arange(n, dtype=np.int32))
    for dt in g["dts"]:
        st, _ = orc.rk4_step(st, dom, float(g["h"]), prm, float(dt), dtype=np.float64)
    order = np.argsort(st.uid)
    u = st.vel[order]
    E = 0.5 * (g["mass"] * (u ** 2).sum(1)).sum()
    assert E == pytest.approx(float(g["E"]), rel=1e-10)
    np.testing.assert_allclose(u, g["u_T"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(st.rho[order], g["rho_T"], rtol=0, atol=1e-10)


def test_golden_rhs_float64_initial_state():
    """float64 right-hand side at the rollout's initial state (rho = rho0):
    the oracle's normalization, viscosity and continuity equal the
    reference modules' to float64 rounding."""
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_rollout_grad_16.npz")))
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    n = g["pos"].shape[0]
    prm = orc.DeltaParams(rho0=1.0, cs=10.0, delta=0.1, kernel_scale=float(g["h"]) / (2 * float(g["dx"])),
                          viscosity=1, visc_coeff=0.01, eps=1e-6)
    _, _, ex = orc.deltaplus_rhs(dom, g["pos"], g["vel"], np.ones(n), g["mass"], float(g["h"]), prm, np.float64)
    _close(ex["L"].reshape(-1, 2, 2), g["rhs0_L"], 1e-12)
    _close(ex["terms"][:, 2:4], g["rhs0_visc"], 1e-12)
    _close(ex["terms"][:, 4], g["rhs0_cont"], 1e-12)
    assert np.all(g["rhs0_ap"] == 0) and np.all(ex["terms"][:, 0:2] == 0)
