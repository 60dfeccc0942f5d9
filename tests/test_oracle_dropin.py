"""The oracle's restatements of the reference's per-pair drop-in ops
(viscosityKernel, mDBC, computeNeighborhood in its C++ and Python forms) and
of the compact-hash-map search (countNeighbors / buildNeighborList), pinned
by hand-computed cases, brute force and the already pinned oracle paths."""
import numpy as np
import pytest

from diffsph_amd import sampling
from oracle import oracle as orc


def _csr(counts):
    counts = np.asarray(counts, np.int32)
    return counts, np.concatenate([[0], np.cumsum(counts)])[:-1].astype(np.int32)


# ---------------------------------------------------------------------------
# viscosityKernel (partiBench/viscosityKernel.h:98-145)
# ---------------------------------------------------------------------------
def _visc_case(vi, pi_switch):
    """Particle 0 with one neighbor (particle 1) at r = 0.5 h, h = 1,
    xhat = (1, 0), gradW = (1, 0), m = rho = 1; eps = 0.5, alpha = 1,
    h_0 = 2, cs = rho0 = 1: pi = v_x / ((0.5 + 0.5) + 1) = v_x / 2,
    V = 1 / (1 + 1), factor = 1 * 2 * 1 * 1 / 1 = 2."""
    vel = np.array([vi, [0.0, 0.0]], np.float32)
    nn, off = _csr([1, 0])
    return orc.op_viscosity(vel, vel, np.ones(2), np.ones(2), np.ones(2), np.array([2.0, 1.0]),
                            np.array([1]), np.array([1.0]), np.array([[1.0, 0.0]]), np.array([0.5]),
                            np.array([[1.0, 0.0]]), nn, off, cs=1.0, rho0=1.0, alpha=1.0, eps=0.5,
                            pi_switch=pi_switch)


def test_viscosity_closed_form():
    # 2 * 0.5 * (1/2) * 1 = 0.5: the (r + eps) + h^2 denominator and the
    # m_j / (rho_i + rho_j) volume (the Python module's m_j / rho_j and
    # r + eps h^2 would give 2.0)
    np.testing.assert_array_equal(_visc_case([1.0, 0.0], False), [[0.5, 0.0], [0.0, 0.0]])
    np.testing.assert_array_equal(_visc_case([-1.0, 0.0], False), [[-0.5, 0.0], [0.0, 0.0]])
    # pi-switch keeps approaching pairs (v_ij . x_ij < 0) only
    np.testing.assert_array_equal(_visc_case([1.0, 0.0], True), [[0.0, 0.0], [0.0, 0.0]])
    np.testing.assert_array_equal(_visc_case([-1.0, 0.0], True), [[-0.5, 0.0], [0.0, 0.0]])
    # velocity perpendicular to xhat: no contribution
    np.testing.assert_array_equal(_visc_case([0.0, 3.0], False), [[0.0, 0.0], [0.0, 0.0]])


def test_viscosity_matches_float64_scatter_form():
    rng = np.random.default_rng(3)
    n, P = 300, 6000
    i = np.sort(rng.integers(0, n, P))
    j = rng.integers(0, n, P)
    nn, off = _csr(np.bincount(i, minlength=n))
    vel = rng.standard_normal((n, 2)).astype(np.float32)
    m = rng.uniform(0.5, 1.5, n).astype(np.float32)
    rho = rng.uniform(0.9, 1.1, n).astype(np.float32)
    hp = rng.uniform(0.05, 0.1, n).astype(np.float32)
    h = rng.uniform(0.05, 0.1, P).astype(np.float32)
    q = rng.uniform(0, 1, P).astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, P)
    xh = np.stack([np.cos(ang), np.sin(ang)], 1).astype(np.float32)
    g = rng.standard_normal((P, 2)).astype(np.float32)
    cs, rho0, alpha, eps = 20.0, 1000.0, 0.02, 1e-6
    for sw in (False, True):
        got = orc.op_viscosity(vel, vel, m, rho, rho, hp, j, h, g, q, xh, nn, off, cs, rho0, alpha, eps, sw)
        d = np.float64
        vij = vel[i].astype(d) - vel[j].astype(d)
        vr = (vij * xh.astype(d)).sum(1)
        r = q.astype(d) * h
        pi = vr / (r + eps + h.astype(d) ** 2)
        if sw:
            pi = np.where(vr < 0, pi, 0.0)
        V = m[j].astype(d) / (rho[i].astype(d) + rho[j])
        acc = np.zeros((n, 2))
        np.add.at(acc, i, (V * pi)[:, None] * g)
        ref = acc * (alpha * hp.astype(d) * cs * rho0 / rho)[:, None]
        np.testing.assert_allclose(got, ref, rtol=2e-4, atol=1e-4 * np.abs(ref).max())


# ---------------------------------------------------------------------------
# mDBC on per-pair arrays (v2/cppSrc/mDBC.h:129-244)
# ---------------------------------------------------------------------------
def _walls(ns=40, layers=3):
    dx = 1.0 / ns
    xs = (np.arange(ns) + 0.5) * dx
    fx, fy = np.meshgrid(xs, xs, indexing="xy")
    fluid = np.stack([fx.ravel(), fy.ravel()], 1).astype(np.float32)
    bx, by = np.meshgrid(xs, -(np.arange(layers) + 0.5) * dx, indexing="xy")
    wall = np.stack([bx.ravel(), by.ravel()], 1).astype(np.float32)
    normals = np.tile(np.array([[0.0, -1.0]], np.float32), (wall.shape[0], 1))
    h = sampling.volume_to_support(dx * dx, 50.0)
    ghosts = orc.boundary_ghosts(wall, wall[:, 1].copy(), normals, h)
    return fluid, wall, ghosts, dx, h


def _ghost_pairs(fluid, ghosts, h):
    """Per-pair arrays of the ghost -> fluid neighbor lists, from the
    oracle's pair filter on the brute-force lists (scatter, uniform h)."""
    dom = orc.make_domain((-0.5, -0.5), (1.5, 1.5), (False, False))
    _, offs, nbr = orc.neighbors_brute(dom, ghosts, fluid, h0=h)
    cnt, off = _csr(np.diff(offs))
    hg = np.full(ghosts.shape[0], h, np.float32)
    hf = np.full(fluid.shape[0], h, np.float32)
    res = orc.pairs_filter(dom, 0, ghosts, hg, fluid, hf, nbr.astype(np.int64), cnt, off)
    np.testing.assert_array_equal(res[0], cnt)  # the filter keeps every brute-force pair
    return dom, offs, nbr, res


def test_mdbc_dropin_equals_pinned_oracle_variant():
    """On the per-pair arrays of the oracle's pair filter (the reference's
    own quotients r/h, d/(r+1e-7)), the per-pair-array restatement equals
    the on-the-fly C++-op variant of the oracle's mDBC (pinned against the
    pseudo-inverse and the clamps in test_oracle_deltasph.py, evaluated with
    the solvers' reciprocal geometry, a few ulp apart) within 1e-5,
    including rows with <= 5 neighbors and empty rows."""
    fluid, wall, ghosts, dx, h = _walls()
    # a few ghosts far from the fluid (empty rows) and near its corner (few)
    far = np.array([[0.5, -3.0], [1.0 + 0.3 * h, 1.0 + 0.3 * h]], np.float32)
    ghosts = np.concatenate([ghosts, far]).astype(np.float32)
    wall = np.concatenate([wall, far + np.float32(0.01)]).astype(np.float32)
    dom, offs, nbr, res = _ghost_pairs(fluid, ghosts, h)
    cnt, off, oi, oj, rij, xij, hij, W, gW = res
    assert cnt[-2] == 0 and 0 < cnt[-1] <= 5
    mass = np.full(fluid.shape[0], dx * dx, np.float32)
    x = fluid.astype(np.float64)
    for rho in (1.05 + x @ np.array([0.03, -0.1]), 1.0 + 0.2 * x[:, 1],
                1.1 + 0.05 * np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 1])):
        rho = rho.astype(np.float32)
        a, sa = orc.op_mdbc(mass, rho, wall, ghosts, oj, W, gW, rij, xij, hij, cnt, off, 1.0)
        b, sb = orc.mdbc(dom, offs, nbr, ghosts, fluid, h, mass, rho, wall, 1.0, h, np.float32, variant=1)
        np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(sa, sb, rtol=1e-5, atol=1e-6)
    assert a[-2] == 1.0 and b[-2] == 1.0  # no neighbors: max(0 / 1e-7, rho0)


# ---------------------------------------------------------------------------
# computeNeighborhood, C++ form (v2/cppSrc/computeNeighborhood.h:110-187)
# ---------------------------------------------------------------------------
def _all_pairs(n):
    nn, off = _csr(np.full(n, n))
    return np.tile(np.arange(n, dtype=np.int64), n), nn, off


def test_cpp_neighborhood_closed_form():
    pos = np.array([[0.9, 0.0], [-0.9, 0.0], [0.0, 0.3]], np.float32)
    h = np.full(3, 0.5, np.float32)
    jj, nn, off = _all_pairs(3)
    for per, expect in (((True, True), {(0, 0), (0, 1), (1, 0), (1, 1), (2, 2)}),
                        ((False, False), {(0, 0), (1, 1), (2, 2)})):
        counts, offs, oi, oj, rij, xij, hij = orc.op_neighborhood(pos, h, pos, h, jj, nn, off, (-1, -1), (1, 1), per)
        assert set(zip(oi.tolist(), oj.tolist())) == expect
        np.testing.assert_array_equal(offs, np.concatenate([[0], np.cumsum(counts)])[:-1])
        np.testing.assert_array_equal(hij, 0.5)
        if per[0]:
            k = int(np.where((oi == 0) & (oj == 1))[0][0])
            # d = x_0 - x_1 across the periodic boundary: 0.9 - (-0.9) -> -0.2
            np.testing.assert_allclose(rij[k], 0.4, rtol=1e-6)
            np.testing.assert_allclose(xij[k], [-1.0, 0.0], rtol=1e-5, atol=1e-7)
        self_k = np.where(oi == oj)[0]
        np.testing.assert_array_equal(rij[self_k], 0.0)
        np.testing.assert_array_equal(xij[self_k], 0.0)  # 0 / (0 + 1e-6 h)


@pytest.mark.parametrize("per", [(True, True), (False, True)])
def test_cpp_neighborhood_matches_brute_force(per):
    rng = np.random.default_rng(5)
    n = 400
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    h = rng.uniform(0.1, 0.2, n).astype(np.float32)
    jj, nn, off = _all_pairs(n)
    counts, offs, oi, oj, rij, xij, hij = orc.op_neighborhood(pos, h, pos, h, jj, nn, off, (-1, -1), (1, 1), per)
    d = pos[:, None, :].astype(np.float64) - pos[None, :, :]
    for k in range(2):
        if per[k]:
            d[..., k] -= 2.0 * np.round(d[..., k] / 2.0)
    r2 = (d ** 2).sum(-1)
    h2 = (h.astype(np.float64) ** 2)[None, :]
    sure_in = r2 <= h2 * (1 - 1e-6)
    sure_out = r2 > h2 * (1 + 1e-6)
    got = np.zeros((n, n), bool)
    got[oi, oj] = True
    assert np.all(got[sure_in]) and not np.any(got[sure_out])
    np.testing.assert_array_equal(counts, got.sum(1))
    # rows in candidate order (j ascending here), values
    for i in range(0, n, 37):
        seg = oj[offs[i]:offs[i] + counts[i]]
        assert np.all(np.diff(seg) > 0)
    r = np.sqrt(r2[oi, oj])
    # p - q is rounded to float32 before the double modulo: an absolute
    # error of ~1e-7 in d across the periodic boundary
    np.testing.assert_allclose(rij, r / h[oj], rtol=1e-6, atol=2e-6)
    np.testing.assert_allclose(hij, h[oj])
    far = r > 1e-3
    den = r + 1e-6 * h[oi].astype(np.float64)
    np.testing.assert_allclose(xij[far], d[oi, oj][far] / den[far, None], rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------------------
# computeNeighborhood, Python form (v2/modules/neighborhood.py:433-460)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_python_filter_matches_brute_force_and_wendland(mode):
    rng = np.random.default_rng(7 + mode)
    n = 350
    dom = orc.make_domain((-1, -1), (1, 1), (True, False))
    pos = orc.wrap_positions(dom, rng.uniform(-1.2, 1.2, (n, 2)), np.float32)
    h = rng.uniform(0.1, 0.2, n).astype(np.float32)
    # candidates: brute force at 1.5 h (a Verlet list)
    _, coff, cnbr = orc.neighbors_brute(dom, pos, pos, h * np.float32(1.5), h * np.float32(1.5), mode=mode)
    cnt, off = _csr(np.diff(coff))
    counts, offs, oi, oj, rij, xij, hij, W, gW = orc.pairs_filter(dom, mode, pos, h, pos, h, cnbr.astype(np.int64),
                                                                  cnt, off)
    _, boff, bnbr = orc.neighbors_brute(dom, pos, pos, h, h, mode=mode)
    np.testing.assert_array_equal(counts, np.diff(boff))
    np.testing.assert_array_equal(oj, bnbr)   # both rows in ascending j
    hh = {0: h[oj], 1: h[oi], 2: np.float32(0.5) * (h[oi] + h[oj])}[mode]
    np.testing.assert_array_equal(hij, hh)
    q = rij.astype(np.float64)
    H = hij.astype(np.float64)
    w = (1 - q) ** 4 * (1 + 4 * q) * 7 / np.pi / H ** 2
    dw = -20 * q * (1 - q) ** 3 * 7 / np.pi / H ** 3
    np.testing.assert_allclose(W, w, rtol=2e-6, atol=1e-6 * w.max())
    np.testing.assert_allclose(gW, xij * dw[:, None], rtol=2e-6, atol=1e-6 * np.abs(dw).max())
    self_k = oi == oj
    np.testing.assert_allclose(W[self_k], 7 / np.pi / H[self_k] ** 2, rtol=1e-6)
    np.testing.assert_array_equal(gW[self_k], 0.0)


# ---------------------------------------------------------------------------
# Compact hash map search (neighborhood.cpp:65-435, neighborhood.h:100-325)
# ---------------------------------------------------------------------------
def test_cell_offsets_order():
    for r in (1, 2):
        expect = [[ox, oy] for ox in range(-r, r + 1) for oy in range(-r, r + 1)]
        np.testing.assert_array_equal(orc.cell_offsets(r), expect)


def _hash_setup(n, seed, hmax, hlen=4096, lo=-1.0, hi=1.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(lo, hi, (n, 2)).astype(np.float32)
    hm = orc.build_compact_hash_map(pos, (lo, lo), (hi, hi), hmax, hlen)
    return pos, hm


def test_compact_hash_map_invariants():
    pos, hm = _hash_setup(3000, 0, 0.05)
    sp, table, cells, nc, hc = hm["sortedPositions"], hm["hashTable"], hm["cellTable"], hm["numCells"], hm["hCell"]
    np.testing.assert_array_equal(sp, pos[hm["sortIndex"]])
    assert cells[:, 2].sum() == pos.shape[0]
    cid = np.floor((sp - np.float32(-1.0)) / np.float32(hc)).astype(np.int32)
    for t in range(0, cells.shape[0], 7):
        lin, b, c = cells[t]
        idx = cid[b:b + c]
        assert np.all(idx[:, 0] + nc[0] * idx[:, 1] == lin)
        with np.errstate(over="ignore"):
            hsh = np.int32(idx[0, 0]) * np.int32(73856093) + np.int32(idx[0, 1]) * np.int32(19349663)
        bucket = int(np.mod(hsh, table.shape[0]))
        assert table[bucket, 0] <= t < table[bucket, 0] + table[bucket, 1]


@pytest.mark.parametrize("mode", ["scatter", "gather", "symmetric"])
@pytest.mark.parametrize("per", [(True, True), (False, False)])
def test_hash_search_matches_brute_force(mode, per):
    """Every pair with |d| < support (brute force in float64, min-image on
    periodic axes) is found, nothing else, each row in cell-offset order then
    sorted index.  The grid has 40 x 40 cells, so the int32 hash sums of
    most cells wrap negative."""
    n = 2500
    pos, hm = _hash_setup(n, 11, 0.05)
    rng = np.random.default_rng(12)
    hq = rng.uniform(0.03, 0.05, n).astype(np.float32)
    hs = hq[hm["sortIndex"]]
    sp = hm["sortedPositions"]
    nc = hm["numCells"]
    assert nc[0] == 40
    counts, oi, oj = orc.hash_neighbors(pos, hq, 1, sp, hs, hm["hashTable"], nc, hm["cellTable"], hm["hCell"],
                                        (-1, -1), (1, 1), per, mode)
    d = pos[:, None, :].astype(np.float64) - sp[None, :, :]
    for k in range(2):
        if per[k]:
            d[..., k] -= 2.0 * np.round(d[..., k] / 2.0)
    r = np.sqrt((d ** 2).sum(-1))
    H = {"scatter": np.broadcast_to(hs[None, :], r.shape), "gather": np.broadcast_to(hq[:, None], r.shape),
         "symmetric": 0.5 * (hq[:, None].astype(np.float64) + hs[None, :])}[mode]
    got = np.zeros((n, n), bool)
    got[oi, oj] = True
    assert np.all(got[r < H * (1 - 1e-6)]) and not np.any(got[r >= H * (1 + 1e-6)])
    np.testing.assert_array_equal(counts, got.sum(1))
    # row order: (cell offset index, sorted j) increasing
    cell_of = np.floor((sp + np.float32(1.0)) / np.float32(hm["hCell"])).astype(np.int64)
    qcell = np.floor((pos + np.float32(1.0)) / np.float32(hm["hCell"])).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(counts)])
    for i in range(0, n, 53):
        js = oj[off[i]:off[i + 1]]
        dc = cell_of[js] - qcell[i]
        dc = np.where(dc > 1, dc - nc[0], np.where(dc < -1, dc + nc[0], dc))
        key = (dc[:, 0] + 1) * 3 + (dc[:, 1] + 1)
        assert np.all(np.diff(key) >= 0)
        same = np.diff(key) == 0
        assert np.all(np.diff(js)[same] > 0)
