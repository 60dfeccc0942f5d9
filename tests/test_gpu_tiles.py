"""Neighbor tiles of the DFSPH pressure sweeps (dsph_neighbor_tiles and the
LDS-staged sweeps over v2/modules/dfsph.py:35-50).

The tiles are a layout of the neighbor list: a workgroup's runs must hold
exactly the particles of the 3 x 3 cells around its rows' cells (computed
here from the cell grid), every entry of its rows must name, through its
slot, the same neighbor as the list, and the sweeps that read the staged
records must give bit-identical steps to the sweeps that gather them from
global memory, whether every workgroup, some or none of them fit their
tile.
"""
import numpy as np
import pytest
import torch

from diffsph_amd import sampling
from diffsph_amd.dfsph import DFSPHConfig, DFSPHSolver
from diffsph_amd.neighborhood import TILE_RUN_STRIDE

pytestmark = pytest.mark.gpu


def _setup(nside, seed):
    d = sampling.dfsph_random_flow(nside, seed=seed, jitter_width=0.1)
    return d, 3.0 * d["dx"] / 10.0


def _solver(d, dt, **kw):
    dev = torch.device("cuda:0")
    cfg = DFSPHConfig(dt=dt, support=d["h"], **kw)
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    return DFSPHSolver(T(d["pos"]), T(d["vel"]), T(d["mass"]), T(d["area"]), cfg)


def _fields(s):
    return {k: getattr(s, k).cpu().numpy().copy() for k in ("pos", "vel", "rho", "p_div", "p_den", "a_div", "a_den",
                                                           "uid")}


def _expected_tile(cs, ncx, ncy, i0, i1):
    """The particles of the 3 x 3 cells around the cells of sorted
    particles i0..i1 (periodic grid), sorted."""
    c0 = int(np.searchsorted(cs, i0, "right") - 1)
    c1 = int(np.searchsorted(cs, i1, "right") - 1)
    cells = set()
    for c in range(c0, c1 + 1):
        x, y = c % ncx, c // ncx
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                cells.add(((y + dy) % ncy) * ncx + (x + dx) % ncx)
    return np.concatenate([np.arange(cs[c], cs[c + 1]) for c in sorted(cells)] + [np.zeros(0, np.int64)])


# 96^2 = 9216 particles (23 cells per row: a workgroup spans two rows of
# cells and wraps the periodic seam); 272^2 = 73984 particles
@pytest.mark.parametrize("nside,seed", [(96, 0), (272, 3)])
def test_tile_runs_and_slots(nside, seed):
    d, dt = _setup(nside, seed)
    print(f"case nside={nside} seed={seed}", flush=True)
    s = _solver(d, dt, tile_cap=4096)
    s.neighbor_step()
    nl = s.nl
    slot, truns, tn, cap = nl.tiles
    st = s.tile_stats()
    print(st, flush=True)
    assert st["tiled"] == 1.0
    n = s.n
    cs = nl.grid.cell_start.cpu().numpy()
    ncx, ncy = nl.grid.grid.ncell[0], nl.grid.grid.ncell[1]
    counts = nl.counts.cpu().numpy()
    so = nl.slice_off.cpu().numpy()
    nbr = nl.nbr.cpu().numpy()
    sl = slot.cpu().numpy().view(np.uint16)
    trh = truns.cpu().numpy().reshape(-1, TILE_RUN_STRIDE)
    tnh = tn.cpu().numpy()
    for b in range((n + 255) // 256):
        rows = np.arange(b * 256, min(n, b * 256 + 256))
        R = trh[b, 0]
        runs = trh[b, 1:1 + 2 * R].reshape(R, 2)
        lst = np.concatenate([np.arange(a, a + ln) for a, ln in runs])
        # disjoint sorted runs holding exactly the stencil's particles
        assert (np.diff(lst) > 0).all()
        assert (runs[1:, 0] > runs[:-1, 0] + runs[:-1, 1]).all()  # merged: not touching
        np.testing.assert_array_equal(lst, np.unique(_expected_tile(cs, ncx, ncy, rows[0], rows[-1])))
        assert tnh[b] == lst.size
        c = counts[rows]
        k = np.concatenate([so[i // 64] + 64 * np.arange(c[r]) + i % 64 for r, i in enumerate(rows)])
        assert (sl[k] < lst.size).all()
        np.testing.assert_array_equal(lst[sl[k]], nbr[k])


# cap 0: the median neighbor count of the first step's workgroups, so that
# some workgroups use their tile and the others gather
@pytest.mark.parametrize("nside,seed,cap", [(96, 1, 2048), (272, 2, 1024), (272, 2, 0), (96, 1, 64)])
def test_tiled_steps_bit_identical(nside, seed, cap):
    d, dt = _setup(nside, seed)
    if cap == 0:
        probe = _solver(d, dt, tile_cap=4096)
        probe.neighbor_step()
        cap = int(np.median(probe.nl.tiles[2].cpu().numpy()))
    print(f"case nside={nside} seed={seed} cap={cap}", flush=True)
    a = _solver(d, dt, tile_cap=0)
    b = _solver(d, dt, tile_cap=cap)
    for step in range(3):
        sa = a.step()
        sb = b.step()
        st = b.tile_stats()
        print(f"step {step}: iters {sa.div_iters}/{sa.den_iters} tiles {st}", flush=True)
        if cap == 64:
            assert st["tiled"] == 0.0
        if step == 0 and 64 < cap < 1024:
            assert 0.0 < st["tiled"] < 1.0
        assert (sa.div_iters, sa.den_iters) == (sb.div_iters, sb.den_iters)
        fa, fb = _fields(a), _fields(b)
        for k in fa:
            np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)


@pytest.mark.parametrize("nside,seed", [(96, 4), (272, 5)])
def test_tiled_feature_pass_bit_identical(nside, seed):
    """The warm-start feature pass (W_ij streamed, packed records) reads its
    neighbors from the tiles when the list has them."""
    from diffsph_amd.learned import WarmStartNet
    d, dt = _setup(nside, seed)
    print(f"case nside={nside} seed={seed}", flush=True)
    a = _solver(d, dt, tile_cap=0)
    b = _solver(d, dt, tile_cap=4096)
    for s in (a, b):
        s.keep_pair_w = True
        s.step()
        s.step()
    assert b.tile_stats()["tiled"] == 1.0
    net = WarmStartNet(torch.device("cuda:0"))
    xa = net.features(a, normalized=False).cpu().numpy()
    xb = net.features(b, normalized=False).cpu().numpy()
    np.testing.assert_array_equal(xa, xb)
    assert np.abs(xa[:, 1]).max() > 0  # the neighbor sums are not empty


def _fused_vs_two_pass(pos, domain, h, cap, d16, width=67):
    """The fixed-width list with the tiles built in its fill vs the same
    list filled first and tiled by dsph_neighbor_tiles afterwards."""
    from diffsph_amd.neighborhood import CellGrid, Delta16Buffers, search
    dev = pos.device
    grid = CellGrid.build(pos, None, domain, h, h_uniform=h)
    n = pos.shape[0]
    nt = (n + 255) // 256
    ent = ((n + 63) // 64) * 64 * width
    mk = lambda: (torch.full((ent,), -7, dtype=torch.int16, device=dev),  # noqa: E731
                  torch.full((nt * TILE_RUN_STRIDE,), -7, dtype=torch.int32, device=dev),
                  torch.full((nt,), -7, dtype=torch.int32, device=dev))
    ta, tb = mk(), mk()
    da, db = (Delta16Buffers(), Delta16Buffers()) if d16 else (None, None)
    a = search(grid, grid.sorted_pos, None, "scatter", qpos_wrapped=True, layout="fixed", width=width,
               delta16=da, tiles=(cap, *ta))
    b = search(grid, grid.sorted_pos, None, "scatter", qpos_wrapped=True, layout="fixed", width=width, delta16=db)
    assert a.width == width and b.width == width and a.tiles is not None
    b.attach_tiles(cap, *tb)
    return a, b, ta, tb


def _entries(nl):
    """Flat positions of every row's entries in the sliced layout."""
    counts = nl.counts.cpu().numpy().astype(np.int64)
    so = nl.slice_off.cpu().numpy().astype(np.int64)
    i = np.repeat(np.arange(counts.size), counts)
    c = np.arange(i.size) - np.repeat(np.cumsum(counts) - counts, counts)
    return i, so[i // 64] + 64 * c + i % 64


# 96^2: workgroups span two rows of cells and wrap the seam; 272^2; a
# 5 x 5 periodic grid (every candidate range takes the seam path); cap 0:
# the median tile size, so that some tiles do not fit and get no slots;
# cap 64: none fit
@pytest.mark.parametrize("nside,seed,cap,d16", [(96, 0, 4096, True), (272, 3, 1152, False), (272, 3, 0, True),
                                                 (24, 5, 4096, False), (96, 2, 64, False)])
def test_fill_built_tiles_equal_two_pass_tiles(nside, seed, cap, d16):
    from diffsph_amd import _lib
    d, _ = _setup(nside, seed)
    print(f"case nside={nside} seed={seed} cap={cap} d16={d16}", flush=True)
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(d["pos"]).to(dev)
    h = float(d["h"])
    if nside == 24:
        h = 0.34  # a 5 x 5 cell grid over [-1, 1]^2
    domain = _lib.make_domain((-1.0, -1.0), (1.0, 1.0), (True, True))
    median = cap == 0
    if median:
        cap = int(np.median(_fused_vs_two_pass(pos, domain, h, 4096, False)[2][2].cpu().numpy()))
        print(f"median tile {cap}", flush=True)
    a, b, ta, tb = _fused_vs_two_pass(pos, domain, h, cap, d16)
    np.testing.assert_array_equal(a.counts.cpu().numpy(), b.counts.cpu().numpy())
    i, k = _entries(a)
    np.testing.assert_array_equal(a.nbr.cpu().numpy()[k], b.nbr.cpu().numpy()[k])
    if d16:
        np.testing.assert_array_equal(a.nbr16.cpu().numpy()[k], b.nbr16.cpu().numpy()[k])
        assert int(a.nbr16_overflow.item()) == int(b.nbr16_overflow.item())
    tna, tnb = ta[2].cpu().numpy(), tb[2].cpu().numpy()
    ra = ta[1].cpu().numpy().reshape(-1, TILE_RUN_STRIDE)
    rb = tb[1].cpu().numpy().reshape(-1, TILE_RUN_STRIDE)
    np.testing.assert_array_equal(tna, tnb)
    for t in range(tna.size):
        if tna[t] >= 0:
            R = ra[t, 0]
            np.testing.assert_array_equal(ra[t, :1 + 2 * R], rb[t, :1 + 2 * R])
    usable = (tna >= 0) & (tna <= cap)
    if cap == 64:
        assert not usable.any()
    if median:
        assert 0 < usable.mean() < 1
    if cap == 4096:
        assert usable.all()
    sel = usable[i // 256]
    sa = ta[0].cpu().numpy().view(np.uint16)[k[sel]]
    sb = tb[0].cpu().numpy().view(np.uint16)[k[sel]]
    np.testing.assert_array_equal(sa, sb)
