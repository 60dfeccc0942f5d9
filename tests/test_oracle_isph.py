"""The ISPH oracle (v2/modules/incompressible.py, v2/sparse.py) pinned by
mathematics: the pressure Laplacian against the exact Laplacian of a
quadratic field, its zero row sums and symmetry; BiCGStab against scipy's
bicgstab (the algorithm the reference restates) and an exact dense solve;
the relaxed-Jacobi operator's linearity and its diagonal derived by hand;
the source term on a field of known divergence; the closing passes
against the generic SPH operations."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from diffsph_amd import sampling
from oracle import oracle as orc


def _lattice(ns, periodic, jit=0.0, seed=0):
    pos, dx = sampling.sample_regular(ns, ns, (0.0, 0.0), (1.0, 1.0))
    if jit:
        pos = sampling.jitter(pos, dx, jit, seed)
    h = sampling.volume_to_support(dx * dx, 50.0)
    dom = orc.make_domain((0.0, 0.0), (1.0, 1.0), (periodic, periodic))
    _, off, nbr = orc.neighbors_brute(dom, pos, pos, h0=h)
    return pos, dx, h, dom, off, nbr


def _csr(off, nbr, val, n):
    rows = np.repeat(np.arange(n), np.diff(off))
    return sp.csr_matrix((val.astype(np.float64), (rows, nbr)), shape=(n, n))


def test_laplacian_of_quadratic_field_and_row_sums():
    """A p = sum_j t_ij (p_j - p_i) approximates dt * lap(p): for
    p = x^2 + y^2 the interior rows give 4 dt; constant fields are in the
    null space; with uniform volumes the off-diagonal part is symmetric."""
    pos, dx, h, dom, off, nbr = _lattice(40, periodic=False)
    n = pos.shape[0]
    mass = np.full(n, dx * dx, np.float32)
    rho = np.ones(n, np.float32)
    dt = 2e-3
    val, aii, pre = orc.isph_laplacian(dom, off, nbr, pos, h, mass, rho, dt)
    x = pos.astype(np.float64)
    p = (x ** 2).sum(1)
    A = _csr(off, nbr, val, n)
    lap = A @ p / dt
    inner = np.all((x > 1.2 * h) & (x < 1.0 - 1.2 * h), axis=1)
    assert inner.sum() > 400
    np.testing.assert_allclose(lap[inner], 4.0, rtol=0.03)
    ones = A @ np.ones(n)
    assert np.abs(ones).max() < 1e-5 * aii.max()
    off_diag = A - sp.diags(A.diagonal())
    assert (off_diag - off_diag.T).count_nonzero() == 0
    np.testing.assert_allclose(A.diagonal(), -aii)
    np.testing.assert_array_equal(pre, np.float32(-1.0) / aii)
    # the operator is negative semi-definite: 2 dt grad p . grad p >= 0
    rng = np.random.default_rng(1)
    q = rng.standard_normal(n)
    assert q @ (A @ q) < 0


def _system(ns=32, seed=1, noise=0.0, jit=0.1, uniform_rho=False):
    """Periodic jittered lattice with a divergence-free spectral velocity
    (plus `noise` times white noise, which is not), its Laplacian and
    source at dt = 0.3 dx / 10.  With `uniform_rho` the volumes are equal,
    so A is symmetric and its left null space is the constants too."""
    pos, dx, h, dom, off, nbr = _lattice(ns, periodic=True, jit=jit, seed=seed)
    n = pos.shape[0]
    mass = np.full(n, dx * dx, np.float32)
    rho, vol = orc.density(dom, off, nbr, pos, pos, mass, mass, h0=h)
    if uniform_rho:
        rho = np.ones(n, np.float32)
    vel = sampling.spectral_velocity(pos, kmax=4, seed=seed, period=1.0)
    rng = np.random.default_rng(seed)
    vel = (vel + noise * rng.standard_normal(vel.shape)).astype(np.float32)
    dt = 0.3 * dx / 10.0
    val, aii, pre = orc.isph_laplacian(dom, off, nbr, pos, h, mass, rho, dt)
    B = orc.isph_source(dom, off, nbr, pos, h, vel, None, mass, rho, dt, 1.0)
    return dict(pos=pos, h=h, dom=dom, off=off, nbr=nbr, mass=mass, rho=rho, vel=vel, dt=dt, val=val, aii=aii,
                pre=pre, B=B, n=n)


@pytest.mark.parametrize("fp64", [False, True])
def test_bicgstab_matches_scipy(fp64):
    """Same system, same x0 = 0, same Jacobi preconditioner, rtol 1e-3: the
    iteration counts agree with scipy's bicgstab (float64) and both
    solutions reach the tolerance, with float32 or float64 vectors."""
    s = _system()
    n = s["n"]
    A = _csr(s["off"], s["nbr"], s["val"], n)
    b = s["B"].astype(np.float64)
    x, it, status, rn, bn = orc.bicgstab(s["off"], s["nbr"], s["val"], s["pre"], s["B"], np.zeros(n, np.float32),
                                         rtol=1e-3, maxiter=512, fp64=fp64)
    assert status in (0, 1)
    assert bn == pytest.approx(np.linalg.norm(b), rel=1e-6)
    res = np.linalg.norm(A @ x.astype(np.float64) - b) / np.linalg.norm(b)
    count = [0]

    def cb(_):
        count[0] += 1

    M = spla.LinearOperator((n, n), matvec=lambda v: s["pre"].astype(np.float64) * v)
    xs, info = spla.bicgstab(A, b, x0=np.zeros(n), rtol=1e-3, atol=0.0, maxiter=512, M=M, callback=cb)
    assert info == 0
    res_s = np.linalg.norm(A @ xs - b) / np.linalg.norm(b)
    print(f"oracle {it} iterations (status {status}, residual {res:.2e}), scipy {count[0]} ({res_s:.2e})")
    # scipy's callback runs once per completed iteration (not on the s-exit),
    # as `it` counts; float32 rounding may shift the exit by a few iterations
    assert abs(it - count[0]) <= (1 if fp64 else 3)
    assert res < 1.5e-3
    rel = np.linalg.norm(x - xs) / np.linalg.norm(xs)
    assert rel < 1e-2


def test_bicgstab_on_a_nearly_singular_system():
    """The Laplacian's rows sum to zero only up to float32 rounding of the
    stored values (~1e-7 relative), so constants are a near-null mode.  On
    a source with a non-solenoidal part the reference's default float32
    BiCGStab diverges, and the float64 solve reaches the tolerance only by
    a constant offset far larger than the solution's variation.  The first
    iterations of both precisions agree."""
    s = _system(ns=24, seed=3, noise=0.2, jit=0.2)
    n = s["n"]
    args = (s["off"], s["nbr"], s["val"], s["pre"], s["B"], np.zeros(n, np.float32))
    _, _, _, rn32_5, bn = orc.bicgstab(*args, rtol=1e-3, maxiter=5)
    _, _, _, rn64_5, _ = orc.bicgstab(*args, rtol=1e-3, maxiter=5, fp64=True)
    assert rn32_5 == pytest.approx(rn64_5, rel=1e-4)
    A = _csr(s["off"], s["nbr"], s["val"], n)
    assert np.abs(A @ np.ones(n)).max() > 0
    x64, it64, st64, rn64, _ = orc.bicgstab(*args, rtol=1e-3, maxiter=512, fp64=True)
    x32, it32, st32, rn32, _ = orc.bicgstab(*args, rtol=1e-3, maxiter=512)
    print(f"float64: {it64} iterations, |r|/|b| {rn64 / bn:.1e}, mean {x64.mean():.3g} std {x64.std():.3g}; "
          f"float32: status {st32}, |r|/|b| {rn32 / bn:.1e}")
    assert st64 in (0, 1) and abs(x64.mean()) > 1e3 * x64.std()
    assert st32 == 2 and not rn32 < 1e-2 * bn


def test_bicgstab_exact_solution_of_a_small_system():
    """A tight tolerance reproduces the least-squares solution of the small
    periodic system (symmetric, so the mean-free source is consistent) up
    to a constant.  (With varying volumes A is not symmetric, its left
    null vector is not constant, the mean-free source keeps a component
    along it, and the solution picks up a large constant offset.)"""
    s = _system(ns=16, seed=5, uniform_rho=True)
    n = s["n"]
    A = _csr(s["off"], s["nbr"], s["val"], n).toarray()
    b = s["B"].astype(np.float64)
    x, it, status, rn, bn = orc.bicgstab(s["off"], s["nbr"], s["val"], s["pre"], s["B"], np.zeros(n, np.float32),
                                         rtol=1e-6, maxiter=2000, fp64=True)
    assert status in (0, 1)
    ref = np.linalg.lstsq(A, b, rcond=1e-6)[0]
    d = (x - x.mean()) - (ref - ref.mean())
    assert np.linalg.norm(d) / np.linalg.norm(ref - ref.mean()) < 1e-3


def test_bicgstab_zero_rhs_and_iteration_limit():
    s = _system(ns=12)
    n = s["n"]
    x, it, status, rn, bn = orc.bicgstab(s["off"], s["nbr"], s["val"], s["pre"], np.zeros(n, np.float32),
                                         np.ones(n, np.float32))
    assert status == 0 and it == 0 and bn == 0.0 and np.all(x == 0)
    x, it, status, rn, bn = orc.bicgstab(s["off"], s["nbr"], s["val"], s["pre"], s["B"], np.zeros(n, np.float32),
                                         rtol=1e-12, maxiter=3)
    assert status == 2 and it == 3


def _pair_gradients(dom, pos, off, nbr, h):
    """Per-entry grad W_ij (float64) with the reference's geometry."""
    x = pos.astype(np.float64)
    i = np.repeat(np.arange(pos.shape[0]), np.diff(off))
    d = x[i] - x[nbr]
    L = np.array([dom.max[0] - dom.min[0], dom.max[1] - dom.min[1]])
    for k in range(2):
        if dom.periodic[k]:
            d[:, k] = np.where(np.abs(d[:, k]) > L[k] / 2, d[:, k] - np.sign(d[:, k]) * L[k], d[:, k])
    r = np.sqrt((d ** 2).sum(1))
    _, dw = orc.wendland(r / h, h)
    return i, d / (r + 1e-7)[:, None] * dw[:, None]


def test_isph_jacobi_operator_diagonal_and_update():
    """solveIncompressible_relaxedJacobi: with the 'difference' pressure
    gradient the operator K(p) = -dt^2 div(a(p)) is linear with diagonal
    -dt^2 [ -V_i sum_j V_j |gradW_ij|^2 / rho_j + |sum_j V_j gradW_ij|^2 / rho_i ]
    (derived by hand; the DFSPH 'summation' form has + on the first term);
    one update is p + 0.5 (s - K)/alpha without clamping and the error is
    mean |r|."""
    pos, dx, h, dom, off, nbr = _lattice(12, periodic=True, jit=0.3, seed=4)
    n = pos.shape[0]
    m = np.full(n, dx * dx)
    area = m.copy()
    rho, vol = orc.density(dom, off, nbr, pos, pos, m, m, h0=h, dtype=np.float64)
    dt = 1e-3
    vel = np.zeros((n, 2))
    ad, an, ds, vs = orc.dfsph_prepare(dom, off, nbr, pos, vel, area, vol, rho, dt, 1.0, None, h, np.float64)

    def K(p):
        pn, acc, res, it, err, _ = orc.dfsph_solve(dom, off, nbr, pos, h, dt, 1.0, 1.0, 0.0, 1, 1, False, True, vel,
                                                   np.zeros_like(vel), rho, vol, area, an, np.zeros(n), p,
                                                   np.float64, isph=True)
        return res

    s0 = K(np.zeros(n))
    A = lambda p: K(p) - s0  # noqa: E731
    rng = np.random.default_rng(0)
    p1, p2 = rng.standard_normal(n), rng.standard_normal(n)
    np.testing.assert_allclose(A(2.0 * p1 + 3.0 * p2), 2.0 * A(p1) + 3.0 * A(p2), rtol=1e-9, atol=1e-12)
    i, gw = _pair_gradients(dom, pos, off, nbr, h)
    V = vol
    g = np.zeros((n, 2))
    np.add.at(g, i, V[nbr][:, None] * gw)
    t1 = np.zeros(n)
    np.add.at(t1, i, V[nbr] * (gw ** 2).sum(1) / rho[nbr])
    diag = -dt * dt * (-V * t1 + (g ** 2).sum(1) / rho)
    for k in range(0, n, max(1, n // 9)):
        e = np.zeros(n)
        e[k] = 1.0
        assert A(e)[k] == pytest.approx(diag[k], rel=1e-7, abs=1e-18)
    alpha = an / dt
    p0 = rng.standard_normal(n) * 1e-3
    pn, acc, res, it, err, _ = orc.dfsph_solve(dom, off, nbr, pos, h, dt, 1.0, 0.5, 0.0, 1, 1, False, True, vel,
                                               np.zeros_like(vel), rho, vol, area, alpha, np.zeros(n), p0,
                                               np.float64, isph=True)
    src = -s0
    np.testing.assert_allclose(pn, p0 + 0.5 * (src - A(p0)) / alpha, rtol=1e-9, atol=1e-15)
    assert err == pytest.approx(np.abs(res).mean(), rel=1e-12)
    assert (pn < 0).any()  # no clamp


def test_isph_source_of_a_field_with_known_divergence():
    """u = (x^2, 0): div u = 2x; B = rho0 div u - mean(B), so B_i - B_k =
    2 rho0 (x_i - x_k) between interior particles."""
    pos, dx, h, dom, off, nbr = _lattice(40, periodic=False)
    n = pos.shape[0]
    mass = np.full(n, dx * dx, np.float32)
    rho, _ = orc.density(dom, off, nbr, pos, pos, mass, mass, h0=h)
    x = pos.astype(np.float64)
    vel = np.stack([x[:, 0] ** 2, np.zeros(n)], 1).astype(np.float32)
    rho0 = 2.0
    B = orc.isph_source(dom, off, nbr, pos, h, vel, None, mass, rho, 1e-3, rho0).astype(np.float64)
    assert abs(B.mean()) < 1e-6 * np.abs(B).max()
    inner = np.all((x > 1.2 * h) & (x < 1.0 - 1.2 * h), axis=1)
    got = B[inner] - B[inner].mean()
    want = 2 * rho0 * (x[inner, 0] - x[inner, 0].mean())
    assert np.abs(got - want).max() < 0.03 * np.abs(want).max()


def test_isph_close_equals_generic_operations():
    """The fused closing passes equal the composition of the generic SPH
    operations (summation / symmetric / difference gradients) bit for bit."""
    pos, dx, h, dom, off, nbr = _lattice(16, periodic=True, jit=0.3, seed=2)
    n = pos.shape[0]
    rng = np.random.default_rng(2)
    mass = np.full(n, dx * dx, np.float32)
    rho, _ = orc.density(dom, off, nbr, pos, pos, mass, mass, h0=h)
    vel = rng.standard_normal((n, 2)).astype(np.float32)
    adv = rng.standard_normal((n, 2)).astype(np.float32)
    ps = rng.standard_normal(n).astype(np.float32)
    pss = rng.standard_normal(n).astype(np.float32)
    F = np.float32
    # dt crosses the ABI as float32; -dt^2 is formed in double from that value
    dt, rho0 = float(F(1e-3)), 1.5
    vp, sh, p_new, v_new = orc.isph_close(dom, off, nbr, pos, h, mass, rho, vel, adv, ps, pss, dt, rho0, pos)
    gs = orc.sph_operation(dom, off, nbr, pos, pos, 1, 2, mass, rho, ps, 1, qm=mass, qr=rho, qq=ps, h0=h)[:, 0, :]
    gy = orc.sph_operation(dom, off, nbr, pos, pos, 1, 3, mass, rho, pss, 1, qm=mass, qr=rho, qq=pss, h0=h)[:, 0, :]
    vstar = vel + adv * F(dt)
    a = (-gs) / F(rho0)
    vp_ref = vstar + a * F(dt)
    np.testing.assert_array_equal(vp, vp_ref)
    sh_ref = (F(-(dt * dt)) * gy) / F(rho0)
    np.testing.assert_array_equal(sh, sh_ref)
    gv = orc.sph_operation(dom, off, nbr, pos, pos, 1, 1, mass, rho, vp_ref, 2, qm=mass, qr=rho, qq=vp_ref, h0=h)
    v_ref = vp_ref + (gv[:, :, 0] * sh_ref[:, :1] + gv[:, :, 1] * sh_ref[:, 1:])
    np.testing.assert_array_equal(v_new, v_ref)
    np.testing.assert_array_equal(p_new, (pos + vp_ref * F(dt)) + sh_ref)
