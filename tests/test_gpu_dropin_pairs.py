"""The per-pair-array drop-ins (viscosityKernel, mDBC, computeNeighborhood in
its C++ and Python forms), the compact-hash-map search (countNeighbors /
buildNeighborList), the Verlet reuse of neighborSearch and the
torchCompactRadius stand-in on the GPU, against the oracle: bit for bit
where the oracle restates the same arithmetic, set-exact otherwise."""
import numpy as np
import pytest
import torch

from diffsph_amd import dropin, sampling
from diffsph_amd.neighborhood import filter_pairs, max_displacement, neighborSearch
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _eq(got, ref, what):
    g = got.detach().cpu().numpy() if isinstance(got, torch.Tensor) else np.asarray(got)
    r = np.asarray(ref)
    assert g.shape == r.shape, f"{what}: shape {g.shape} vs {r.shape}"
    if not np.array_equal(g, r):
        bad = np.argwhere(g != r)[:5]
        raise AssertionError(f"{what}: {int((g != r).sum())} entries differ, first {bad.tolist()} "
                             f"got {g[tuple(bad[0])]} ref {r[tuple(bad[0])]}")


def _csr(counts):
    counts = np.asarray(counts, np.int32)
    return counts, np.concatenate([[0], np.cumsum(counts)])[:-1].astype(np.int32)


def _flow_pairs(nside, seed, per=(True, True)):
    """Positions of a jittered lattice and the oracle's per-pair arrays of its
    neighbor lists (rows in particle order, j ascending)."""
    d = sampling.dfsph_random_flow(nside, seed=seed, jitter_width=0.3)
    dom = orc.make_domain((-1, -1), (1, 1), per)
    pos = orc.wrap_positions(dom, d["pos"], np.float32)
    n = pos.shape[0]
    h = np.full(n, d["h"], np.float32)
    _, boff, bnbr = orc.neighbors_brute(dom, pos, pos, h, h)
    cnt, off = _csr(np.diff(boff))
    res = orc.pairs_filter(dom, 0, pos, h, pos, h, bnbr.astype(np.int64), cnt, off)
    return d, dom, pos, h, res


def test_viscosity_dropin_bitexact():
    d, dom, pos, h, res = _flow_pairs(48, 1)
    cnt, off, oi, oj, rij, xij, hij, W, gW = res
    n = pos.shape[0]
    rng = np.random.default_rng(2)
    vel = d["vel"].astype(np.float32)
    m = d["mass"].astype(np.float32)
    rho = rng.uniform(0.95, 1.05, n).astype(np.float32)
    for sw in (False, True):
        ref = orc.op_viscosity(vel, vel, m, rho, rho, h, oj, hij, gW, rij, xij, cnt, off, 10.0, 1.0, 0.01, 1e-6, sw)
        tv, tm, tr, th = T(vel), T(m), T(rho), T(h)
        got = dropin.viscosityKernel((T(oi), T(oj)), T(hij), T(W), T(gW), T(rij), T(xij), n, T(cnt), T(off),
                                     (tv, tv), (tr, tr), (tm, tm), (th, th), 10.0, 1.0, 0.01, 1e-6, sw)
        _eq(got, ref, f"viscosity pi_switch={sw}")
    with pytest.raises(RuntimeError, match="not of the correct type"):
        dropin.viscosityKernel((T(oi), T(oj).int()), T(hij), T(W), T(gW), T(rij), T(xij), n, T(cnt), T(off),
                               (tv, tv), (tr, tr), (tm, tm), (th, th), 10.0, 1.0, 0.01, 1e-6, False)


def test_mdbc_dropin_bitexact():
    ns = 40
    dx = 1.0 / ns
    xs = (np.arange(ns) + 0.5) * dx
    fx, fy = np.meshgrid(xs, xs, indexing="xy")
    fluid = np.stack([fx.ravel(), fy.ravel()], 1).astype(np.float32)
    bx, by = np.meshgrid(xs, -(np.arange(3) + 0.5) * dx, indexing="xy")
    wall = np.stack([bx.ravel(), by.ravel()], 1).astype(np.float32)
    hh = sampling.volume_to_support(dx * dx, 50.0)
    normals = np.tile(np.array([[0.0, -1.0]], np.float32), (wall.shape[0], 1))
    ghosts = orc.boundary_ghosts(wall, wall[:, 1].copy(), normals, hh)
    far = np.array([[0.5, -3.0], [1.0 + 0.3 * hh, 1.0 + 0.3 * hh]], np.float32)
    ghosts = np.concatenate([ghosts, far]).astype(np.float32)
    wall = np.concatenate([wall, far + np.float32(0.01)]).astype(np.float32)
    dom = orc.make_domain((-0.5, -0.5), (1.5, 1.5), (False, False))
    _, boff, bnbr = orc.neighbors_brute(dom, ghosts, fluid, h0=hh)
    cnt, off = _csr(np.diff(boff))
    hg = np.full(ghosts.shape[0], hh, np.float32)
    hf = np.full(fluid.shape[0], hh, np.float32)
    _, _, oi, oj, rij, xij, hij, W, gW = orc.pairs_filter(dom, 0, ghosts, hg, fluid, hf, bnbr.astype(np.int64),
                                                          cnt, off)
    mass = np.full(fluid.shape[0], dx * dx, np.float32)
    x = fluid.astype(np.float64)
    for k, rho in enumerate((1.05 + x @ np.array([0.03, -0.1]), 1.0 + 0.2 * x[:, 1])):
        rho = rho.astype(np.float32)
        ref, rshep = orc.op_mdbc(mass, rho, wall, ghosts, oj, W, gW, rij, xij, hij, cnt, off, 1.0)
        got, gshep = dropin.mDBC((T(mass), T(mass)), (T(rho), T(rho)), (T(wall), T(ghosts)), wall.shape[0],
                                 fluid.shape[0], (T(oi), T(oj)), T(cnt), T(off), T(W), T(gW), T(rij), T(xij), T(hij),
                                 1.0)
        _eq(got, ref, f"mDBC density {k}")
        _eq(gshep, rshep, f"Shepard density {k}")


@pytest.mark.parametrize("per", [(True, True), (True, False)])
def test_compute_neighborhood_dropin_bitexact(per):
    """Candidates: the brute-force lists at 1.5 h (a Verlet list); all seven
    outputs equal the oracle's (double-precision offsets, sqrt and
    divisions included)."""
    rng = np.random.default_rng(3)
    n = 3000
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    h = rng.uniform(0.04, 0.06, n).astype(np.float32)
    dom = orc.make_domain((-1, -1), (1, 1), per)
    h15 = h * np.float32(1.5)
    _, coff, cnbr = orc.neighbors_brute(dom, pos, pos, h15, h15)
    cnt, off = _csr(np.diff(coff))
    jj = cnbr.astype(np.int64)
    ii = np.repeat(np.arange(n), cnt).astype(np.int64)
    ref = orc.op_neighborhood(pos, h, pos, h, jj, cnt, off, (-1, -1), (1, 1), per)
    got = dropin.computeNeighborhood((T(ii), T(jj)), n, T(cnt), T(off), (T(pos), T(pos)), (T(h), T(h)),
                                     T(np.array([-1.0, -1.0], np.float32)), T(np.array([1.0, 1.0], np.float32)),
                                     T(np.array(per)))
    for name, g, r in zip(("counts", "offsets", "i", "j", "rij", "xij", "hij"), got, ref):
        _eq(g, r, name)
    # inconsistent input is clipped and skipped, not read out of bounds
    bad = jj.copy()
    bad[::11] = n + 7
    bad[5::11] = -1
    big = cnt.copy()
    big[-1] += 500
    got = dropin.computeNeighborhood((T(ii), T(bad)), n, T(big), T(off), (T(pos), T(pos)), (T(h), T(h)),
                                     [-1.0, -1.0], [1.0, 1.0], list(per))
    ok = (bad >= 0) & (bad < n)
    cnt_ok = np.bincount(ii[ok], minlength=n).astype(np.int32)
    ref = orc.op_neighborhood(pos, h, pos, h, jj[ok], *_csr(cnt_ok), (-1, -1), (1, 1), per)
    for name, g, r in zip(("counts", "offsets", "i", "j", "rij", "xij", "hij"), got, ref):
        _eq(g, r, "clipped " + name)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_pairs_filter_bitexact(mode):
    from diffsph_amd import _lib
    rng = np.random.default_rng(4 + mode)
    n = 4000
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    pos = orc.wrap_positions(dom, rng.uniform(-1.5, 1.5, (n, 2)), np.float32)
    h = rng.uniform(0.03, 0.05, n).astype(np.float32)
    h15 = h * np.float32(1.5)
    _, coff, cnbr = orc.neighbors_brute(dom, pos, pos, h15, h15, mode=mode)
    cnt, off = _csr(np.diff(coff))
    ref = orc.pairs_filter(dom, mode, pos, h, pos, h, cnbr.astype(np.int64), cnt, off)
    got = filter_pairs(_lib.make_domain((-1, -1), (1, 1), (True, True)), mode, T(pos), T(h), T(pos), T(h),
                       T(cnbr.astype(np.int64)), T(cnt), T(off))
    names = ("numNeighbors", "neighborOffsets", "i", "j", "distances", "vectors", "supports", "kernels", "gradients")
    for name, r in zip(names, ref):
        _eq(got[name], r, name)


def _hash_case(n, seed, hmax, per):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    hm = orc.build_compact_hash_map(pos, (-1, -1), (1, 1), hmax, 4096)
    hq = rng.uniform(0.6, 1.0, n).astype(np.float32) * np.float32(hmax)
    return pos, hq, hm


@pytest.mark.parametrize("mode", ["scatter", "gather", "symmetric"])
@pytest.mark.parametrize("per", [(True, True), (False, True)])
def test_hash_search_dropins_bitexact(mode, per):
    """countNeighbors / buildNeighborList over the notebook's compact hash
    map (40 x 40 cells: most int32 cell hashes wrap negative) equal the
    oracle's CPU order entry for entry."""
    n = 6000
    pos, hq, hm = _hash_case(n, 21, 0.05, per)
    hs = hq[hm["sortIndex"]]
    ref_cnt, ref_i, ref_j = orc.hash_neighbors(pos, hq, 1, hm["sortedPositions"], hs, hm["hashTable"],
                                               hm["numCells"], hm["cellTable"], hm["hCell"], (-1, -1), (1, 1), per,
                                               mode)
    args = (T(pos), T(hq), 1, T(hm["sortedPositions"]), T(hs), T(hm["hashTable"]), 4096, T(hm["numCells"]),
            T(hm["cellTable"]), T(hm["qMin"]), hm["hCell"], T(np.array([1.0, 1.0], np.float32)),
            T(np.array([-1.0, -1.0], np.float32)), T(np.array(per)), mode)
    cnt = dropin.countNeighbors(*args)
    _eq(cnt, ref_cnt, "counts")
    off = T(np.concatenate([[0], np.cumsum(ref_cnt)])[:-1].astype(np.int32))
    gi, gj = dropin.buildNeighborList(cnt, off, int(ref_cnt.sum()), *args)
    _eq(gi, ref_i, "i")
    _eq(gj, ref_j, "j")
    with pytest.raises(RuntimeError, match="Invalid support mode"):
        dropin.countNeighbors(*args[:-1], "nearest")


def test_verlet_check_and_neighbor_search_reuse():
    """neighborSearch keeps its candidate list while the particles move
    less than (verletScale - 1) min(h) / 2 and re-filters it; the result is
    the brute-force neighbor set at the current positions either way."""
    d = sampling.dfsph_random_flow(64, seed=8, jitter_width=0.3)
    n = d["pos"].shape[0]
    h = d["h"]
    dom = orc.make_domain((-1, -1), (1, 1), (True, True))
    config = {"domain": {"minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0], "periodicity": [True, True]},
              "neighborhood": {"verletScale": 1.5}}
    sup = torch.full((n,), h, device=DEV)
    p0 = T(d["pos"])
    rng = np.random.default_rng(9)
    step = rng.standard_normal((n, 2)).astype(np.float32)
    step /= np.linalg.norm(step, axis=1, keepdims=True)

    def check_set(nd, pos_np):
        wp = orc.wrap_positions(dom, pos_np, np.float32)
        _, off_b, nbr_b = orc.neighbors_brute(dom, wp, wp, h0=h)
        _eq(nd["numNeighbors"], np.diff(off_b), "numNeighbors")
        i = nd["indices"][0].cpu().numpy()
        j = nd["indices"][1].cpu().numpy()
        got = set(zip(i.tolist(), j.tolist()))
        assert got == {(a, int(b)) for a in range(n) for b in nbr_b[off_b[a]:off_b[a + 1]]}

    state = {"positions": p0, "supports": sup}
    ds, nd = neighborSearch(state, state, config)
    check_set(nd, d["pos"])
    full0 = nd["fullIndices"]
    # move every particle by 0.2 h (< 0.25 h): the candidates are reused
    p1 = (d["pos"] + np.float32(0.2 * h) * step).astype(np.float32)
    dmax, hmin = max_displacement(p0, T(p1), sup)
    np.testing.assert_allclose(dmax, np.linalg.norm(p1 - d["pos"], axis=1).max(), rtol=1e-6)
    assert hmin == np.float32(h)
    state1 = {"positions": T(p1), "supports": sup}
    ds, nd1 = neighborSearch(state1, state1, config, priorState=nd, neighborDatastructure=ds)
    assert nd1["fullIndices"] is full0
    check_set(nd1, p1)
    # move by 0.3 h from the initial positions: rebuilt
    p2 = (d["pos"] + np.float32(0.3 * h) * step).astype(np.float32)
    state2 = {"positions": T(p2), "supports": sup}
    ds, nd2 = neighborSearch(state2, state2, config, priorState=nd1, neighborDatastructure=ds)
    assert nd2["fullIndices"] is not full0
    check_set(nd2, p2)


def test_radius_search_stand_in():
    from diffsph_amd import compact_radius
    rng = np.random.default_rng(10)
    x = rng.uniform(0, 1, (2000, 2)).astype(np.float32)
    y = rng.uniform(0, 1, (2500, 2)).astype(np.float32)
    hq = rng.uniform(0.02, 0.04, 2000).astype(np.float32)
    dom = orc.make_domain((0, 0), (1, 1), (True, True))
    for kw, mode, qh, h0 in ((dict(fixedSupport=0.03), "scatter", None, 0.03),
                             (dict(support=T(hq)), "gather", hq, 0.0)):
        i, j = compact_radius.radiusSearch(T(x), T(y), mode=mode, domainMin=torch.zeros(2), domainMax=torch.ones(2),
                                           periodicity=True, **kw)
        _, off_b, nbr_b = orc.neighbors_brute(dom, orc.wrap_positions(dom, x, np.float32),
                                              orc.wrap_positions(dom, y, np.float32), qh=qh, h0=h0,
                                              mode=1 if mode == "gather" else 0)
        assert np.all(np.diff(i.cpu().numpy()) >= 0)
        got = set(zip(i.cpu().numpy().tolist(), j.cpu().numpy().tolist()))
        assert got == {(a, int(b)) for a in range(len(off_b) - 1) for b in nbr_b[off_b[a]:off_b[a + 1]]}
