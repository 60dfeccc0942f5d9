"""GPU parity: HIP neighbor search and pair operators vs the CPU oracle.

Neighbor lists must match bit for bit (same rows, same order); fields
within 1e-5 relative to the field's max magnitude (float32, both sides sum
in the same neighbor order).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from diffsph_amd import _lib, sampling
from diffsph_amd.neighborhood import CellGrid, neighborSearch, search

pytestmark = pytest.mark.gpu

RTOL = 1e-5  # float32 fields: |gpu - oracle| <= RTOL * max|oracle|


def _close(gpu, ref, rtol=RTOL):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = max(np.abs(ref).max(), 1e-30)
    err = np.abs(gpu - ref).max() / scale
    assert err <= rtol, f"max rel error {err:.3e} > {rtol}"


def _case(nside, seed, jit=0.1, periodic=(True, True)):
    d = sampling.dfsph_random_flow(nside, seed=seed, jitter_width=jit)
    return d, periodic


@pytest.mark.parametrize("nside,seed,jit", [(64, 0, 0.0), (64, 1, 0.1), (96, 2, 0.5), (200, 3, 0.1)])
def test_search_matches_oracle_bitexact(orc, nside, seed, jit):
    print(f"case nside={nside} seed={seed} jit={jit}", flush=True)
    d, per = _case(nside, seed, jit)
    dom_o = orc.make_domain((-1, -1), (1, 1), per)
    g_o = orc.grid_build(dom_o, d["pos"], d["h"])
    c_o, off_o, nbr_o = orc.neighbors_cell(dom_o, g_o, g_o.sorted_pos, h0=d["h"])
    dev = torch.device("cuda:0")
    dom = _lib.make_domain((-1, -1), (1, 1), per)
    pos = torch.from_numpy(d["pos"]).to(dev)
    g = CellGrid.build(pos, None, dom, d["h"], h_uniform=d["h"])
    assert (g.grid.ncell[0], g.grid.ncell[1]) == (g_o.ncx, g_o.ncy)
    np.testing.assert_array_equal(g.cell_start.cpu().numpy(), g_o.cell_start)
    np.testing.assert_array_equal(g.sort_index.cpu().numpy(), g_o.sort_index)
    np.testing.assert_array_equal(g.sorted_pos.cpu().numpy(), g_o.sorted_pos)
    nl = search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True)
    np.testing.assert_array_equal(nl.counts.cpu().numpy(), c_o)
    np.testing.assert_array_equal(nl.offsets.cpu().numpy(), off_o)
    np.testing.assert_array_equal(nl.csr_nbr().cpu().numpy(), nbr_o)
    # one-pass fixed-width layout, and its fallback when a row is longer
    for width, fits in [(67, True), (int(c_o.max()), True), (int(c_o.max()) - 1, False)]:
        print(f"  fixed width {width}", flush=True)
        nf = search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True, layout="fixed", width=width)
        assert nf.width == (width if fits else 0)
        assert nf.num_pairs == int(off_o[-1])
        np.testing.assert_array_equal(nf.counts.cpu().numpy(), c_o)
        np.testing.assert_array_equal(nf.offsets.cpu().numpy(), off_o)
        np.testing.assert_array_equal(nf.csr_nbr().cpu().numpy(), nbr_o)


def test_search_matches_brute_force_sets_small(orc):
    rng = np.random.default_rng(5)
    n = 3000
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    h = rng.uniform(0.04, 0.09, n).astype(np.float32)
    for per in [(True, True), (False, True), (False, False)]:
        for mode in ["scatter", "gather", "symmetric"]:
            print(f"case per={per} mode={mode}", flush=True)
            dom_o = orc.make_domain((-1, -1), (1, 1), per)
            m = {"scatter": 0, "gather": 1, "symmetric": 2}[mode]
            _, off_b, nbr_b = orc.neighbors_brute(dom_o, pos, pos, qh=h, rh=h, mode=m)
            brute = set()
            for i in range(n):
                for k in range(off_b[i], off_b[i + 1]):
                    brute.add((i, int(nbr_b[k])))
            dev = torch.device("cuda:0")
            dom = _lib.make_domain((-1, -1), (1, 1), per)
            tp = torch.from_numpy(pos).to(dev)
            th = torch.from_numpy(h).to(dev)
            g = CellGrid.build(tp, th, dom, float(h.max()))
            nl = search(g, tp, th, mode)
            off = nl.offsets.cpu().numpy()
            nbr = nl.csr_nbr().cpu().numpy()
            si = g.sort_index.cpu().numpy()
            got = set()
            for i in range(n):
                for k in range(off[i], off[i + 1]):
                    got.add((i, int(si[nbr[k]])))
            assert got == brute


def test_search_edge_cases():
    dev = torch.device("cuda:0")
    dom = _lib.make_domain((-1, -1), (1, 1), (True, True))
    # empty set
    pos = torch.zeros(0, 2, device=dev)
    g = CellGrid.build(pos, None, dom, 0.1, h_uniform=0.1)
    nl = search(g, pos, None, "scatter")
    assert nl.num_pairs == 0 and nl.offsets.cpu().tolist() == [0]
    nl = search(g, pos, None, "scatter", layout="fixed")
    assert nl.num_pairs == 0 and nl.offsets.cpu().tolist() == [0] and nl.width == 67
    # single particle, far outside the box (wrapped)
    pos = torch.tensor([[5.3, -7.9]], device=dev)
    g = CellGrid.build(pos, None, dom, 0.1, h_uniform=0.1)
    nl = search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True)
    assert nl.num_pairs == 1
    # all particles in one cell (degenerate clustering)
    pos = torch.full((500, 2), 0.25, device=dev) + torch.rand(500, 2, device=dev) * 1e-3
    g = CellGrid.build(pos, None, dom, 0.1, h_uniform=0.1)
    nl = search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True)
    assert nl.num_pairs == 500 * 500
    nl = search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True, layout="fixed")
    assert nl.num_pairs == 500 * 500 and nl.width == 0
    # too few periodic cells is an error, not a fault
    with pytest.raises(RuntimeError):
        CellGrid.build(pos, None, dom, 0.9, h_uniform=0.9)


@pytest.mark.parametrize("layout,width", [("fixed", 67), ("fixed", 8), ("sell", 0)])
def test_remapped_entries_equal_sort_index_of_slots(layout, width):
    """search(remap=True) writes grid.sort_index[j] for every entry j of the
    plain list: in the fixed-width fill itself, in a pass after the two-pass
    fill, and through the fallback when a row outgrows the width (8)."""
    dev = torch.device("cuda:0")
    d = sampling.dfsph_random_flow(96, seed=5, jitter_width=0.1)
    dom = _lib.make_domain((-1, -1), (1, 1), (True, True))
    g = CellGrid.build(torch.from_numpy(d["pos"]).to(dev), None, dom, d["h"], h_uniform=d["h"])
    # queries: the particles in their input order (not the grid's slots)
    q = torch.from_numpy(d["pos"]).to(dev)
    a = search(g, q, None, "scatter", layout=layout, width=width)
    b = search(g, q, None, "scatter", layout=layout, width=width, remap=True)
    print(f"layout {layout} width {width}: {a.width} / {b.width}, {a.pair_count()} pairs", flush=True)
    assert torch.equal(a.counts, b.counts) and a.width == b.width
    si = g.sort_index.cpu().numpy()
    ja, jb = a.csr_nbr().cpu().numpy(), b.csr_nbr().cpu().numpy()
    np.testing.assert_array_equal(si[ja], jb)


def _sorted_setup(orc, nside, seed, jit=0.1):
    d = sampling.dfsph_random_flow(nside, seed=seed, jitter_width=jit)
    dom_o = orc.make_domain((-1, -1), (1, 1), (True, True))
    g_o = orc.grid_build(dom_o, d["pos"], d["h"])
    si = g_o.sort_index
    _, off, nbr = orc.neighbors_cell(dom_o, g_o, g_o.sorted_pos, h0=d["h"])
    dev = torch.device("cuda:0")
    dom = _lib.make_domain((-1, -1), (1, 1), (True, True))
    g = CellGrid.build(torch.from_numpy(d["pos"]).to(dev), None, dom, d["h"], h_uniform=d["h"])
    nl = search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True)
    return d, dom_o, g_o, si, off, nbr, nl


@pytest.mark.parametrize("nside", [64, 160])
def test_density_and_operations_match_oracle(orc, nside):
    d, dom_o, g_o, si, off, nbr, nl = _sorted_setup(orc, nside, 11)
    dev = torch.device("cuda:0")
    wpos = g_o.sorted_pos
    m = d["mass"][si] * np.float32(1.0 + 0.1 * np.sin(np.arange(si.size)))  # non-uniform masses
    m = m.astype(np.float32)
    vel = d["vel"][si]
    h = d["h"]
    rho_o, vol_o = orc.density(dom_o, off, nbr, wpos, wpos, m, m, h0=h)
    L = _lib.lib()
    pr = nl.pairs()
    st = _lib.stream_handle(dev)
    tm = torch.from_numpy(m).to(dev)
    rho = torch.empty(m.size, device=dev)
    vol = torch.empty(m.size, device=dev)
    _lib.check(L.dsph_density(C.byref(pr), tm.data_ptr(), tm.data_ptr(), rho.data_ptr(), vol.data_ptr(), st))
    _close(rho.cpu().numpy(), rho_o)
    _close(vol.cpu().numpy(), vol_o)
    trho = torch.from_numpy(rho_o).to(dev)
    q1 = (np.cos(3 * wpos[:, 0]) * np.sin(2 * wpos[:, 1])).astype(np.float32)
    tq1 = torch.from_numpy(q1).to(dev)
    tv = torch.from_numpy(vel).to(dev)
    for op, gm, q, w in [(0, 0, tq1, 1), (0, 0, tv, 2), (1, 0, tq1, 1), (1, 1, tq1, 1), (1, 2, tq1, 1),
                         (1, 3, tq1, 1), (1, 1, tv, 2), (2, 0, tv, 2), (2, 1, tv, 2), (2, 2, tv, 2), (2, 3, tv, 2)]:
        print(f"case op={op} gm={gm} width={w}", flush=True)
        shape = {0: (m.size, w), 1: (m.size, w, 2), 2: (m.size,)}[op]
        out = torch.empty(shape, device=dev)
        _lib.check(L.dsph_sph_operation(C.byref(pr), op, gm, tm.data_ptr(), tm.data_ptr(), trho.data_ptr(),
                                        trho.data_ptr(), q.data_ptr(), q.data_ptr(), w, out.data_ptr(), st))
        qn = q.cpu().numpy()
        ref = orc.sph_operation(dom_o, off, nbr, wpos, wpos, op, gm, m, rho_o, qn, w, qm=m, qr=rho_o, qq=qn, h0=h)
        _close(out.cpu().numpy().reshape(ref.shape), ref)


def test_neighbor_dict_dropin_matches_brute_force(orc):
    """neighborSearch(...) returns the reference's neighborDict keys in
    original particle order; compare pair sets and per-pair values."""
    n_side = 40
    d = sampling.dfsph_random_flow(n_side, seed=4, jitter_width=0.3)
    dev = torch.device("cuda:0")
    state = {"positions": torch.from_numpy(d["pos"]).to(dev),
             "supports": torch.full((d["pos"].shape[0],), d["h"], device=dev)}
    config = {"domain": {"minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0], "periodicity": [True, True]}}
    _, nd = neighborSearch(state, state, config)
    i = nd["indices"][0].cpu().numpy()
    j = nd["indices"][1].cpu().numpy()
    dom_o = orc.make_domain((-1, -1), (1, 1), (True, True))
    _, off_b, nbr_b = orc.neighbors_brute(dom_o, d["pos"], d["pos"], h0=d["h"])
    brute = {(a, int(b)) for a in range(len(off_b) - 1) for b in nbr_b[off_b[a]:off_b[a + 1]]}
    assert set(zip(i.tolist(), j.tolist())) == brute
    assert np.all(np.diff(i) >= 0)  # rows in original order
    cnt = nd["numNeighbors"].cpu().numpy()
    np.testing.assert_array_equal(cnt, np.diff(off_b))
    # per-pair kernels vs the Wendland closed form at the exported q
    q = nd["distances"].cpu().numpy().astype(np.float64)
    w_ref, dw_ref = orc.wendland(q, d["h"])
    _close(nd["kernels"].cpu().numpy(), w_ref)
    g = nd["gradients"].cpu().numpy()
    x = nd["vectors"].cpu().numpy()
    _close(g, x * dw_ref[:, None])


def test_dropin_ops_match_torch_reference():
    """sphDensity / sphInterpolation / scatterAdd drop-ins vs a float32
    scatter_add formulation of their documented formulas."""
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0)
    n, P = 700, 9000
    i = np.sort(rng.integers(0, n, P))
    j = rng.integers(0, n, P)
    cnt = np.bincount(i, minlength=n).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
    m = rng.uniform(0.5, 1.5, n).astype(np.float32)
    rho = rng.uniform(0.9, 1.1, n).astype(np.float32)
    W = rng.uniform(0, 2, P).astype(np.float32)
    q = rng.standard_normal((n, 3)).astype(np.float32)
    # every device buffer is held in a variable for the whole call: a raw
    # pointer of a temporary tensor would dangle once the tensor is freed
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    tm, trho, tq, tj, tW, tcnt, toff = T(m), T(rho), T(q), T(j.astype(np.int64)), T(W), T(cnt), T(off)
    vals = rng.standard_normal((P, 2)).astype(np.float32)
    tvals = T(vals)
    L = _lib.lib()
    st = _lib.stream_handle(dev)
    out = torch.empty(n, device=dev)
    _lib.check(L.dsph_op_sph_density(tm.data_ptr(), n, tj.data_ptr(), tW.data_ptr(), P, n, tcnt.data_ptr(),
                                     toff.data_ptr(), out.data_ptr(), st))
    ref = np.zeros(n, np.float64)
    np.add.at(ref, i, m[j].astype(np.float64) * W)
    _close(out.cpu().numpy(), ref)
    out3 = torch.empty(n, 3, device=dev)
    _lib.check(L.dsph_op_sph_interpolation(tm.data_ptr(), trho.data_ptr(), tq.data_ptr(), 3, n, tj.data_ptr(),
                                           tW.data_ptr(), P, n, tcnt.data_ptr(), toff.data_ptr(), out3.data_ptr(),
                                           st))
    ref3 = np.zeros((n, 3), np.float64)
    np.add.at(ref3, i, (m[j] / rho[j] * W)[:, None].astype(np.float64) * q[j])
    _close(out3.cpu().numpy(), ref3)
    out2 = torch.empty(n, 2, device=dev)
    _lib.check(L.dsph_op_scatter_add(tvals.data_ptr(), 2, P, n, tcnt.data_ptr(), toff.data_ptr(), out2.data_ptr(),
                                     st))
    ref2 = np.zeros((n, 2), np.float64)
    np.add.at(ref2, i, vals.astype(np.float64))
    _close(out2.cpu().numpy(), ref2)
    # out-of-range neighbor indices and an inconsistent CSR are ignored, not faults
    bad_j = j.astype(np.int64).copy()
    bad_j[::7] = n + 12345
    bad_j[1::7] = -3
    tbad = T(bad_j)
    big_cnt = cnt.copy()
    big_cnt[-1] += 1000  # last segment runs past the pair arrays
    tbig = T(big_cnt)
    _lib.check(L.dsph_op_sph_density(tm.data_ptr(), n, tbad.data_ptr(), tW.data_ptr(), P, n, tbig.data_ptr(),
                                     toff.data_ptr(), out.data_ptr(), st))
    ok = (bad_j >= 0) & (bad_j < n)
    ref = np.zeros(n, np.float64)
    np.add.at(ref, i[ok], m[bad_j[ok]].astype(np.float64) * W[ok])
    _close(out.cpu().numpy(), ref)


def test_dropin_module_on_neighbor_dict(orc):
    """diffsph_amd.dropin with the reference's calling convention on a
    neighborDict: density and interpolation of a linear field vs the
    oracle's CSR density / interpolation, and the getAccessor-style errors."""
    from diffsph_amd import dropin
    d = sampling.dfsph_random_flow(48, seed=6, jitter_width=0.2)
    dev = torch.device("cuda:0")
    n = d["pos"].shape[0]
    state = {"positions": torch.from_numpy(d["pos"]).to(dev), "supports": torch.full((n,), d["h"], device=dev)}
    config = {"domain": {"minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0], "periodicity": [True, True]}}
    _, nd = neighborSearch(state, state, config)
    m = torch.from_numpy(d["mass"]).to(dev)
    idx = (nd["indices"][0].long().contiguous(), nd["indices"][1].long().contiguous())
    W = nd["kernels"].contiguous()
    cnt, off = nd["numNeighbors"].int().contiguous(), nd["neighborOffsets"].int().contiguous()
    rho = dropin.sphDensity((m, m), (m, m), (m, m), idx, W, n, cnt, off)
    dom_o = orc.make_domain((-1, -1), (1, 1), (True, True))
    g = orc.grid_build(dom_o, d["pos"], d["h"])
    wpos = np.empty_like(g.sorted_pos)
    wpos[g.sort_index] = g.sorted_pos
    _, off_b, nbr_b = orc.neighbors_brute(dom_o, wpos, wpos, h0=d["h"])
    rho_ref, _ = orc.density(dom_o, off_b, nbr_b, wpos, wpos, d["mass"], d["mass"], h0=d["h"])
    _close(rho.cpu().numpy(), rho_ref)
    q = torch.stack([torch.ones(n, device=dev), 2.0 * torch.ones(n, device=dev)], 1).contiguous()
    qi = dropin.sphInterpolation((m, m), (rho, rho), (q, q), idx, W, n, cnt, off)
    # interpolation of a constant field = its Shepard sum; here V_j = m_j / rho_j
    shep = dropin.scatterAdd((m[idx[1]] / rho[idx[1]] * W).contiguous(), n, cnt, off)
    _close(qi[:, 0].cpu().numpy(), shep.cpu().numpy())
    _close(qi[:, 1].cpu().numpy(), 2.0 * shep.cpu().numpy())
    with pytest.raises(RuntimeError, match="kernels"):
        dropin.sphDensity((m, m), (m, m), (m, m), idx, W.double(), n, cnt, off)
    with pytest.raises(RuntimeError, match="GPU"):
        dropin.scatterAdd(W.cpu(), n, cnt, off)
    with pytest.raises(RuntimeError, match="numNeighbors"):
        dropin.scatterAdd(W, n, cnt.long(), off)


def test_large_search_properties():
    """1M particles (the DFSPH workload size): size-independent properties."""
    d = sampling.dfsph_random_flow(1024, seed=0, jitter_width=0.1)
    dev = torch.device("cuda:0")
    dom = _lib.make_domain((-1, -1), (1, 1), (True, True))
    g = CellGrid.build(torch.from_numpy(d["pos"]).to(dev), None, dom, d["h"], h_uniform=d["h"])
    nl = search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True)
    cnt = nl.counts
    assert int(cnt.min()) >= 30 and int(cnt.max()) <= 80
    assert nl.num_pairs == int(cnt.long().sum())
    # symmetry (uniform h): count pairs (i,j) with i<j equals those with i>j
    off = nl.offsets.long()
    rows = torch.repeat_interleave(torch.arange(g.n, device=dev), cnt.long())
    cols = nl.csr_nbr().long()
    assert int((rows < cols).sum()) == int((rows > cols).sum())
    assert int((rows == cols).sum()) == g.n
    # symmetric set: sorted keys of (i,j) equal sorted keys of (j,i)
    k1 = torch.sort(rows * g.n + cols).values
    k2 = torch.sort(cols * g.n + rows).values
    assert torch.equal(k1, k2)
    del off


def test_search_1m_matches_oracle_bitexact(orc):
    """The 1M-particle workload (1024^2 jittered lattice on [-1,1]^2, h =
    3.99 dx): grid, counts, offsets and every list entry bit-exact in the
    two-pass and in the one-pass fixed-width layout, and the 16-bit modular
    deltas (deltas in the thousands, across the periodic seam) decode to the
    int32 list."""
    from diffsph_amd.neighborhood import Delta16Buffers
    print("case 1024^2 seed 0", flush=True)
    d, per = _case(1024, 0, 0.1)
    dom_o = orc.make_domain((-1, -1), (1, 1), per)
    g_o = orc.grid_build(dom_o, d["pos"], d["h"])
    c_o, off_o, nbr_o = orc.neighbors_cell(dom_o, g_o, g_o.sorted_pos, h0=d["h"])
    dev = torch.device("cuda:0")
    dom = _lib.make_domain((-1, -1), (1, 1), per)
    g = CellGrid.build(torch.from_numpy(d["pos"]).to(dev), None, dom, d["h"], h_uniform=d["h"])
    np.testing.assert_array_equal(g.cell_start.cpu().numpy(), g_o.cell_start)
    np.testing.assert_array_equal(g.sort_index.cpu().numpy(), g_o.sort_index)
    np.testing.assert_array_equal(g.sorted_pos.cpu().numpy(), g_o.sorted_pos)
    n = d["pos"].shape[0]
    for layout, width in [("sell", 0), ("fixed", 67)]:
        nl = search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True, layout=layout, width=width,
                    delta16=Delta16Buffers())
        print(f"  layout {layout}: pairs {nl.num_pairs} width {nl.width}", flush=True)
        assert nl.num_pairs == int(off_o[-1]) and nl.width == width
        np.testing.assert_array_equal(nl.counts.cpu().numpy(), c_o)
        np.testing.assert_array_equal(nl.offsets.cpu().numpy(), off_o)
        np.testing.assert_array_equal(nl.csr_nbr().cpu().numpy(), nbr_o)
        # 16-bit deltas: j = (i + d) mod n for every stored entry
        assert int(nl.nbr16_overflow.cpu()[0]) == 0
        full = nl.nbr.cpu().numpy().astype(np.int64)
        d16 = nl.nbr16[:full.shape[0]].cpu().numpy().astype(np.int64)
        so = nl.slice_off.cpu().numpy().astype(np.int64)
        rows = np.arange(n)
        col = np.arange(int(c_o.max()))
        idx = so[rows // 64][:, None] + 64 * col[None, :] + (rows % 64)[:, None]
        valid = col[None, :] < c_o[:, None]
        i_ = np.broadcast_to(rows[:, None], idx.shape)[valid]
        e = idx[valid]
        assert np.array_equal((i_ + d16[e]) % n, full[e])
        assert np.abs(d16[e]).max() > 1000


# cells of 16 (LDS rank sort), 32 +- 4 (both sort paths in one grid) and 64
# particles (insertion sort only)
@pytest.mark.parametrize("nside,hscale", [(96, 1.0), (96, 1.4), (96, 2.0)])
def test_grid_cell_order_matches_oracle(orc, nside, hscale):
    """The per-cell sort (entries ranked in LDS up to 32 per cell, longer
    cells insertion-sorted in place) gives the oracle's sort index: within
    each cell, increasing original index."""
    d, per = _case(nside, 9, 0.5)
    h = float(d["h"]) * hscale
    print(f"case nside={nside} h={h}", flush=True)
    dom_o = orc.make_domain((-1, -1), (1, 1), per)
    g_o = orc.grid_build(dom_o, d["pos"], h)
    g = CellGrid.build(torch.from_numpy(d["pos"]).to("cuda:0"), None, _lib.make_domain((-1, -1), (1, 1), per), h,
                       h_uniform=h)
    cs = g.cell_start.cpu().numpy()
    np.testing.assert_array_equal(cs, g_o.cell_start)
    m = np.diff(cs)
    print(f"cell sizes: max {m.max()}, cells over 32: {(m > 32).sum()}", flush=True)
    if hscale == 1.4:
        assert (m > 32).any() and (m <= 32).any()
    np.testing.assert_array_equal(g.sort_index.cpu().numpy(), g_o.sort_index)


def test_gather_set_permutes_every_array():
    """dsph_gather_set: arrays of 1, 2 and 4 words per element (float32,
    float32 pairs, int64, float32 quads) permuted by one index in one pass
    equal torch's indexing; misaligned or oversized requests are refused."""
    import ctypes as C
    from diffsph_amd._lib import DsphError, check, lib, ptr, stream_handle
    dev = torch.device("cuda:0")
    n = 100003
    g = torch.Generator().manual_seed(4)
    idx = torch.randperm(n, generator=g).to(torch.int32).to(dev)
    src = [torch.randn(n, generator=g).to(dev), torch.randn(n, 2, generator=g).to(dev),
           torch.randint(-2**40, 2**40, (n,), generator=g, dtype=torch.int64).to(dev),
           torch.randn(n, 4, generator=g).to(dev)]
    dst = [torch.empty_like(t) for t in src]
    words = [1, 2, 2, 4]
    k = len(src)
    mk = lambda ts: (C.c_void_p * k)(*[t.data_ptr() for t in ts])  # noqa: E731
    check(lib().dsph_gather_set(ptr(idx), n, k, mk(src), mk(dst), (C.c_int32 * k)(*words), stream_handle(dev)),
          "gather_set")
    for s, d in zip(src, dst):
        assert torch.equal(d, s[idx.long()])
    bad = [src[1][:, 1:]]  # 4-byte aligned, 2-word elements
    with pytest.raises(DsphError):
        check(lib().dsph_gather_set(ptr(idx), n - 1, 1, (C.c_void_p * 1)(bad[0].data_ptr()),
                                    (C.c_void_p * 1)(dst[1].data_ptr()), (C.c_int32 * 1)(2), stream_handle(dev)),
              "gather_set")


@pytest.mark.parametrize("kind", ["stacked", "nan"])
def test_overfull_cell_raises(kind):
    """A diverged or collapsed state (5000 particles at one point, or 5000
    NaN positions, which fall into cell 0) overfills a cell: the grid build
    and the search stay bounded and the search raises, in the one-pass and
    the two-pass layouts; a DFSPH step on such a state raises too."""
    from diffsph_amd._lib import DsphError
    from diffsph_amd.dfsph import DFSPHConfig, DFSPHSolver
    d, per = _case(96, 10, 0.1)
    extra = np.full((5000, 2), np.nan if kind == "nan" else 0.1, np.float32)
    pos = np.concatenate([d["pos"], extra]).astype(np.float32)
    print(f"case {kind}: {pos.shape[0]} particles", flush=True)
    dev = torch.device("cuda:0")
    dom = _lib.make_domain((-1, -1), (1, 1), per)
    g = CellGrid.build(torch.from_numpy(pos).to(dev), None, dom, d["h"], h_uniform=d["h"])
    assert int(torch.diff(g.cell_start).max()) >= 5000
    for layout in ("fixed", "sell"):
        with pytest.raises(DsphError, match="more than 4096"):
            search(g, g.sorted_pos, None, "scatter", qpos_wrapped=True, layout=layout)
    n = pos.shape[0]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)  # noqa: E731
    vel = np.concatenate([d["vel"], np.zeros((5000, 2), np.float32)])
    with pytest.raises(DsphError, match="more than 4096"):
        s = DFSPHSolver(T(pos), T(vel), T(np.full(n, d["mass"][0])), T(np.full(n, d["area"][0])),
                        DFSPHConfig(dt=3.0 * d["dx"] / 10.0, support=d["h"]))
        s.step()
