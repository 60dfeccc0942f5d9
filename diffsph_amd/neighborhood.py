"""Cell-list neighbor search on the GPU (host side).

Mirrors the reference's v2 neighbor search interface
(src/diffSPH/v2/modules/neighborhood.py:352-593) on top of the HIP search in
``csrc/search.hip``:

* :class:`CellGrid` sorts a particle set into cell order (the reference's
  ``buildDataStructure`` / ``buildCompactHashMap``; a dense periodic grid
  replaces the hash map),
* :func:`search` builds the CSR neighbor list of a query set against a grid,
* :func:`neighborSearch` returns the reference's ``(datastructure,
  neighborDict)`` pair with the same dict keys (``indices``, ``distances``,
  ``vectors``, ``supports``, ``kernels``, ``gradients``, ``numNeighbors``,
  ``neighborOffsets``...), materialising per-pair arrays only there.

Pairs are kept when |mod(x_i - x_j)| / h_j <= 1 ('scatter' mode), evaluated
in float32 with separately rounded products and IEEE sqrt/div.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import DsphError, check, lib, ptr

MODES = {"scatter": 0, "gather": 1, "symmetric": 2}
TILE_MAX_RUNS = 16                      # DSPH_TILE_MAX_RUNS
TILE_RUN_STRIDE = 2 * TILE_MAX_RUNS + 1  # int32 per workgroup in tile_runs


class Workspace:
    """Grow-only scratch buffer per device (the C ABI never allocates)."""

    def __init__(self):
        self._buf = {}

    def get(self, nbytes: int, device) -> torch.Tensor:
        key = str(device)
        nbytes = max(int(nbytes), 256)
        b = self._buf.get(key)
        if b is None or b.numel() < nbytes:
            b = torch.empty(nbytes + nbytes // 4, dtype=torch.uint8, device=device)
            self._buf[key] = b
        return b


WORKSPACE = Workspace()


def _domain_from_config(config):
    d = config["domain"]
    return _lib.make_domain(_as_list(d["minExtent"]), _as_list(d["maxExtent"]),
                            _as_list(d.get("periodicity", d.get("periodic", False))))


def _as_list(v):
    if isinstance(v, torch.Tensor):
        return [float(x) for x in v.detach().cpu().reshape(-1).tolist()] if v.dtype != torch.bool else [
            bool(x) for x in v.detach().cpu().reshape(-1).tolist()]
    if isinstance(v, (list, tuple)):
        return list(v)
    return [v, v]


def _check_f32(t, name, shape_tail=None):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} is not defined")
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be float32")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} is not contiguous")
    if not t.is_cuda:
        raise RuntimeError(f"{name} is not on the GPU")
    if shape_tail is not None and tuple(t.shape[1:]) != tuple(shape_tail):
        raise RuntimeError(f"{name} is not of the correct dimension")


@dataclass
class CellGrid:
    """A particle set sorted into a dense cell grid."""
    domain: _lib.Domain
    grid: _lib.Grid
    n: int
    cell_start: torch.Tensor   # int32 [ncell+1]
    sort_index: torch.Tensor   # int32 [n]   sorted slot -> original index
    rank: torch.Tensor         # int32 [n]   original index -> sorted slot
    sorted_pos: torch.Tensor   # f32 [n,2]   wrapped
    sorted_h: torch.Tensor | None
    h_uniform: float

    @staticmethod
    def build(positions: torch.Tensor, supports: torch.Tensor | None, domain: _lib.Domain,
              hsearch: float, h_uniform: float = 0.0) -> "CellGrid":
        _check_f32(positions, "positions", (2,))
        n = positions.shape[0]
        dev = positions.device
        grid = _lib.Grid()
        check(lib().dsph_grid_dims(C.byref(domain), C.c_float(hsearch), C.byref(grid)), "grid_dims")
        ncell = grid.ncell[0] * grid.ncell[1]
        cell_start = torch.empty(ncell + 1, dtype=torch.int32, device=dev)
        sort_index = torch.empty(n, dtype=torch.int32, device=dev)
        rank = torch.empty(n, dtype=torch.int32, device=dev)
        sorted_pos = torch.empty(n, 2, dtype=torch.float32, device=dev)
        sorted_h = None
        if supports is not None:
            _check_f32(supports, "supports")
            sorted_h = torch.empty(n, dtype=torch.float32, device=dev)
        wsb = lib().dsph_grid_build_workspace(n, C.byref(grid))
        ws = WORKSPACE.get(wsb, dev)
        check(lib().dsph_grid_build(C.byref(domain), C.byref(grid), ptr(positions), ptr(supports), n,
                                    ptr(cell_start), ptr(sort_index), ptr(rank), ptr(sorted_pos),
                                    ptr(sorted_h), ptr(ws), ws.numel(), _lib.stream_handle(dev)),
              "grid_build")
        return CellGrid(domain, grid, n, cell_start, sort_index, rank, sorted_pos, sorted_h,
                        float(h_uniform))


@dataclass
class NeighborList:
    """Neighbor list; rows are query particles, entries sorted reference slots.

    ``nbr`` is in the sliced-ELL layout (64 rows per slice, column-major,
    padded with -1) when ``slice_off`` is set, else CSR.  ``offsets`` is the
    CSR scan of ``counts`` in both cases."""
    counts: torch.Tensor    # int32 [nq]
    offsets: torch.Tensor | None  # int32 [nq+1] (None until built: see ensure_offsets)
    nbr: torch.Tensor       # int32 [entries]
    num_pairs: int | None   # None until read (deferred overflow check)
    qpos: torch.Tensor      # wrapped query positions [nq,2] (row order)
    qh: torch.Tensor | None
    grid: CellGrid
    mode: int
    slice_off: torch.Tensor | None = None
    entries: int = 0        # length of the layout (pairs + padding)
    nbr16: torch.Tensor | None = None           # 16-bit modular deltas (pressure sweeps)
    nbr16_overflow: torch.Tensor | None = None  # int32 device flag
    width: int = 0          # row width of the one-pass fixed-width layout (0: two-pass)
    tiles: tuple | None = None  # (slot uint16 as int16, tile_runs, tile_n, cap): pressure-sweep tiles

    def attach_compact16(self, buf16: torch.Tensor, flag: torch.Tensor, limit: int = 32767):
        """Write the 16-bit delta copy of the list into `buf16` (int16, at
        least `entries` long) and use it in the pressure sweeps; `flag`
        (int32 [1]) is set on the device when a delta does not fit."""
        self.nbr16, self.nbr16_overflow = None, None
        pr = self.pairs()
        check(lib().dsph_neighbor_compact16(C.byref(pr), limit, ptr(buf16), buf16.numel(), ptr(flag),
                                            _lib.stream_handle(buf16.device)), "neighbor_compact16")
        self.nbr16, self.nbr16_overflow = buf16, flag

    def attach_tiles(self, cap: int, slot: torch.Tensor, tile_runs: torch.Tensor, tile_n: torch.Tensor):
        """Build the neighbor tiles of this (square) list for the pressure
        sweeps and use them there: `slot` (int16 storage of uint16 slots, at
        least `entries` long), `tile_runs` (int32, ceil(nq / 256) *
        TILE_RUN_STRIDE) and `tile_n` (int32, ceil(nq / 256)).  Workgroups
        whose neighbors do not fit `cap` read them from global memory."""
        self.tiles = None
        pr = self.pairs()
        check(lib().dsph_neighbor_tiles(C.byref(pr), C.byref(self.grid.grid), ptr(self.grid.cell_start), int(cap),
                                        ptr(slot), slot.numel(), ptr(tile_runs), ptr(tile_n),
                                        _lib.stream_handle(slot.device)), "neighbor_tiles")
        self.tiles = (slot, tile_runs, tile_n, int(cap))

    def ensure_offsets(self) -> torch.Tensor:
        """The CSR offsets (a list built with a deferred overflow check has
        none until asked for)."""
        if self.offsets is None:
            nq = self.counts.shape[0]
            dev = self.counts.device
            off = torch.empty(nq + 1, dtype=torch.int32, device=dev)
            ws = WORKSPACE.get(lib().dsph_scan_workspace(nq), dev)
            check(lib().dsph_scan_exclusive(ptr(self.counts), nq, ptr(off), None, ptr(ws), ws.numel(),
                                            _lib.stream_handle(dev)), "scan")
            self.offsets = off
        return self.offsets

    def pair_count(self) -> int:
        """num_pairs, read from the offsets when the list was built with a
        deferred overflow check (synchronises then)."""
        if self.num_pairs is None:
            self.num_pairs = int(self.ensure_offsets()[-1].item())
        return self.num_pairs

    def pairs(self) -> _lib.Pairs:
        p = _lib.Pairs()
        p.nq = self.counts.shape[0]
        p.nr = self.grid.n
        p.offsets = ptr(self.offsets)
        p.nbr = ptr(self.nbr)
        p.qpos = ptr(self.qpos)
        p.rpos = ptr(self.grid.sorted_pos)
        p.qh = ptr(self.qh)
        p.rh = ptr(self.grid.sorted_h)
        p.h_uniform = self.grid.h_uniform
        p.mode = self.mode
        p.domain = self.grid.domain
        p.counts = ptr(self.counts)
        p.slice_off = ptr(self.slice_off)
        p.nbr16 = ptr(self.nbr16)
        p.nbr16_overflow = ptr(self.nbr16_overflow)
        if self.tiles is not None:
            p.tile_slot, p.tile_runs, p.tile_n = ptr(self.tiles[0]), ptr(self.tiles[1]), ptr(self.tiles[2])
            p.tile_cap = self.tiles[3]
        return p

    def csr_nbr(self) -> torch.Tensor:
        """Neighbor indices in CSR order (rows back to back)."""
        if self.slice_off is None:
            return self.nbr
        self.ensure_offsets()
        out = torch.empty(max(self.pair_count(), 1), dtype=torch.int32, device=self.nbr.device)
        pr = self.pairs()
        check(lib().dsph_sell_to_csr(C.byref(pr), ptr(out), _lib.stream_handle(out.device)), "sell_to_csr")
        return out[: self.num_pairs]


class Delta16Buffers:
    """Grow-only int16 buffer and int32 overflow flag for the 16-bit
    neighbor deltas written by :func:`search` (``delta16=``)."""

    def __init__(self):
        self.buf = None
        self.flag = None

    def get(self, n: int, device):
        if self.buf is None or self.buf.numel() < n:
            self.buf = torch.empty(n + n // 8, dtype=torch.int16, device=device)
        if self.flag is None:
            self.flag = torch.zeros(1, dtype=torch.int32, device=device)
        return self.buf, self.flag


CELL_MAX = 4096  # DSPH_CELL_MAX
_OVERFULL = (f"search: a grid cell holds more than {CELL_MAX} particles (a diverged or collapsed state; NaN "
             "positions fall into cell 0): the list would be incomplete")


def search(grid: CellGrid, qpos: torch.Tensor, qh: torch.Tensor | None, mode: str = "scatter",
           qpos_wrapped: bool = False, layout: str = "sell", delta16: Delta16Buffers | None = None,
           width: int = 67, deferred: torch.Tensor | None = None, remap: bool = False,
           tiles: tuple | None = None) -> NeighborList:
    """Neighbor list of the query set `qpos` against the sorted set `grid`.

    Counts, scans (host reads of the pair and entry totals, as the
    reference's `.item()` does) and fills.  With ``qpos_wrapped`` the
    queries are the grid's own wrapped sorted positions.  ``layout`` is
    "sell" (sliced ELL, the layout the pair kernels stream), "csr", or
    "fixed": the sliced layout with every row `width` entries wide, filled
    in one pass without the count pass (one host read of the overflow flag
    and the pair total); a row longer than `width` makes it fall back to
    "sell" (the returned list then has ``width == 0``).  The entries, their
    order and the counts are the same in all three layouts.

    ``deferred`` (int32 [1] device tensor, "fixed" layout only): no host
    read at all -- the row-overflow flag is OR-ed into it on the device and
    the list is returned as is (rows longer than `width` truncated), with
    ``num_pairs`` None (read on demand by :meth:`NeighborList.pair_count`).
    The caller checks the flag later and redoes the work when it is set;
    this is what lets a whole step be captured into a graph.

    ``remap``: the entries are the reference particles' indices in the
    grid's input order (``grid.sort_index[j]``) instead of sorted slots, as
    :func:`dsph_remap_neighbors` would make them ("fixed" layout: in the
    same pass; other layouts: a pass after the fill).

    ``tiles`` = (cap, slot, tile_runs, tile_n), buffers as
    :meth:`NeighborList.attach_tiles` takes them ("fixed" layout, square
    list): the list's neighbor tiles are built with the fill (the run
    tables first, each entry's slot in the fill itself) and attached.  A
    list that falls back to "sell" comes back without tiles."""
    if delta16 is not None and not qpos_wrapped:
        raise DsphError("search: 16-bit neighbor deltas need the grid's own sorted positions as queries "
                        "(qpos_wrapped=True)")
    _check_f32(qpos, "queryPositions", (2,))
    m = MODES[mode] if isinstance(mode, str) else int(mode)
    nq = qpos.shape[0]
    dev = qpos.device
    st = _lib.stream_handle(dev)
    counts = torch.empty(nq, dtype=torch.int32, device=dev)
    L = lib()
    args = (C.byref(grid.domain), C.byref(grid.grid), ptr(grid.cell_start), ptr(grid.sorted_pos),
            ptr(grid.sorted_h), grid.n, ptr(qpos), ptr(qh), nq, 1 if qpos_wrapped else 0,
            C.c_float(grid.h_uniform), m)
    if deferred is not None and (layout != "fixed" or int(width) < 1):
        raise DsphError("search: deferred overflow checks need the fixed-width layout")
    if remap and delta16 is not None:
        raise DsphError("search: remapped entries have no 16-bit deltas")
    if tiles is not None and (remap or not qpos_wrapped or layout != "fixed"):
        raise DsphError("search: tiles are built for a square fixed-width list (qpos_wrapped, no remap)")
    if layout == "fixed":
        nl = _search_fixed(L, args, grid, qpos, qh, m, qpos_wrapped, counts, delta16, int(width), st, deferred,
                           grid.sort_index if remap else None, tiles)
        if nl is not None:
            return nl
        layout = "sell"
    cs = grid.cell_start
    if grid.n > 0 and int((cs[1:] - cs[:-1]).max()) > CELL_MAX:
        raise DsphError(_OVERFULL)
    check(L.dsph_neighbor_count(*args, ptr(counts), st), "neighbor_count")
    offsets = torch.empty(nq + 1, dtype=torch.int32, device=dev)
    ws = WORKSPACE.get(L.dsph_scan_workspace(nq), dev)
    total = C.c_int64(0)
    check(L.dsph_scan_exclusive(ptr(counts), nq, ptr(offsets), C.byref(total), ptr(ws), ws.numel(), st), "scan")
    npairs = int(total.value)
    slice_off = None
    entries = npairs
    if layout == "sell":
        slice_off = torch.empty((nq + 63) // 64 + 1, dtype=torch.int32, device=dev)
        ws2 = WORKSPACE.get(L.dsph_sell_workspace(nq), dev)
        tot2 = C.c_int64(0)
        check(L.dsph_sell_layout(ptr(counts), nq, ptr(slice_off), C.byref(tot2), ptr(ws2), ws2.numel(), st),
              "sell_layout")
        entries = int(tot2.value)
    elif layout != "csr":
        raise ValueError(f"unknown layout {layout}")
    cap = max(entries, 1)
    nbr = torch.empty(cap, dtype=torch.int32, device=dev)
    buf16 = flag16 = None
    if delta16 is not None:
        # rows are the grid's own slots (qpos_wrapped): the deltas are written in the same pass
        buf16, flag16 = delta16.get(cap, dev)
        check(L.dsph_neighbor_fill16(*args, ptr(offsets), ptr(slice_off), entries, ptr(nbr), ptr(buf16), cap,
                                     ptr(flag16), st), "neighbor_fill16")
    else:
        check(L.dsph_neighbor_fill(*args, ptr(offsets), ptr(slice_off), entries, ptr(nbr), cap, st),
              "neighbor_fill")
        if remap:
            check(L.dsph_remap_neighbors(ptr(nbr), nbr.numel(), ptr(grid.sort_index), grid.n, st),
                  "remap_neighbors")
    if qpos_wrapped:
        wq = qpos
    else:
        wq = wrap_positions(qpos, grid.domain)
    return NeighborList(counts, offsets, nbr[:max(entries, 1)] if layout == "sell" else nbr[:npairs], npairs, wq, qh,
                        grid, m, slice_off, entries, buf16, flag16)


def _search_fixed(L, args, grid, qpos, qh, m, qpos_wrapped, counts, delta16, width, st, deferred=None, remap=None,
                  tiles=None):
    nq = qpos.shape[0]
    dev = qpos.device
    nsl = (nq + 63) // 64
    cap = max(nsl * 64 * width, 1)
    slice_off = torch.empty(nsl + 1, dtype=torch.int32, device=dev)
    nbr = torch.empty(cap, dtype=torch.int32, device=dev)
    info = torch.empty(2, dtype=torch.int32, device=dev)   # [row overflow, pair total]
    buf16 = flag16 = None
    if delta16 is not None:
        buf16, flag16 = delta16.get(cap, dev)
    if tiles is not None:
        tcap, slot, truns, tn = tiles
        check(L.dsph_neighbor_fill_fixed_tiles(*args, width, ptr(counts), ptr(slice_off), ptr(nbr), cap, ptr(buf16),
                                               ptr(flag16), int(tcap), ptr(slot), slot.numel(), ptr(truns), ptr(tn),
                                               ptr(info), st), "neighbor_fill_fixed_tiles")
    elif remap is not None:
        check(L.dsph_neighbor_fill_fixed_map(*args, width, ptr(counts), ptr(slice_off), ptr(nbr), cap, ptr(remap),
                                             ptr(info), st), "neighbor_fill_fixed_map")
    else:
        check(L.dsph_neighbor_fill_fixed(*args, width, ptr(counts), ptr(slice_off), ptr(nbr), cap, ptr(buf16),
                                         ptr(flag16), ptr(info), st), "neighbor_fill_fixed")
    if deferred is not None:
        # the pair kernels read the sliced rows (counts + slice offsets): the
        # CSR offsets are built only if asked for (NeighborList.ensure_offsets)
        torch.bitwise_or(deferred, info[:1], out=deferred)
        offsets, npairs = None, None
    else:
        offsets = torch.empty(nq + 1, dtype=torch.int32, device=dev)
        ws = WORKSPACE.get(L.dsph_scan_workspace(nq), dev)
        check(L.dsph_scan_exclusive(ptr(counts), nq, ptr(offsets), None, ptr(ws), ws.numel(), st), "scan")
        info[1:2].copy_(offsets[nq:nq + 1])
        ovf, npairs = (int(v) for v in info.cpu())
        if ovf >= 2:
            raise DsphError(_OVERFULL)
        if ovf:
            return None
    wq = qpos if qpos_wrapped else wrap_positions(qpos, grid.domain)
    nl = NeighborList(counts, offsets, nbr, npairs, wq, qh, grid, m, slice_off, cap, buf16, flag16, width)
    if tiles is not None:
        nl.tiles = (tiles[1], tiles[2], tiles[3], int(tiles[0]))
    return nl


def wrap_positions(pos: torch.Tensor, domain: _lib.Domain) -> torch.Tensor:
    """getPeriodicPosition (v2/modules/neighborhood.py:407-430) in float32."""
    _check_f32(pos, "positions", (2,))
    out = torch.empty_like(pos)
    check(lib().dsph_wrap_positions(C.byref(domain), ptr(pos), pos.shape[0], ptr(out),
                                    _lib.stream_handle(pos.device)), "wrap_positions")
    return out


def export_pairs(nl: NeighborList, q_sort_index=None, q_rank=None, r_sort_index=None,
                 want=("i", "j", "distances", "vectors", "supports", "kernels", "gradients")):
    """Per-pair arrays of the reference's neighborDict in original indices."""
    nq = nl.counts.shape[0]
    dev = nl.counts.device
    P = nl.pair_count()
    out_off = torch.empty(nq + 1, dtype=torch.int32, device=dev)
    res = {}

    def mk(name, shape, dtype):
        if name in want:
            res[name] = torch.empty(shape, dtype=dtype, device=dev)
        return res.get(name)

    oi = mk("i", (P,), torch.int64)
    oj = mk("j", (P,), torch.int64)
    od = mk("distances", (P,), torch.float32)
    ov = mk("vectors", (P, 2), torch.float32)
    oh = mk("supports", (P,), torch.float32)
    ow = mk("kernels", (P,), torch.float32)
    og = mk("gradients", (P, 2), torch.float32)
    L = lib()
    ws = WORKSPACE.get(L.dsph_pairs_export_workspace(nq), dev)
    pr = nl.pairs()
    check(L.dsph_pairs_export(C.byref(pr), ptr(q_sort_index), ptr(q_rank), ptr(r_sort_index), ptr(out_off),
                              ptr(oi), ptr(oj), ptr(od), ptr(ov), ptr(oh), ptr(ow), ptr(og), ptr(ws), ws.numel(),
                              _lib.stream_handle(dev)), "pairs_export")
    res["offsets"] = out_off
    return res


def filter_pairs(domain: _lib.Domain, mode, qpos_wrapped: torch.Tensor, qh: torch.Tensor | None,
                 rpos_wrapped: torch.Tensor, rh: torch.Tensor | None, indices_j: torch.Tensor,
                 num_neighbors: torch.Tensor, neighbor_offsets: torch.Tensor, kernels: bool = True) -> dict:
    """computeNeighborhood (v2/modules/neighborhood.py:433-460) on a candidate
    CSR (a Verlet list): keeps the entries with |mod(x_i - x_j)| / h_ij <= 1
    and returns their per-pair arrays (i, j int64; distances = r / h_ij,
    vectors = d / (r + 1e-7), supports = h_ij; kernels / gradients of the
    Wendland C2 kernel as neighborSearch computes them, :585-589) with the
    CSR counts and offsets of the kept entries.  One count pass, a scan (one
    host read of the total, as the reference's tensor sizes need) and one
    fill pass, each a wave per row."""
    m = MODES[mode] if isinstance(mode, str) else int(mode)
    _check_f32(qpos_wrapped, "queryPositions", (2,))
    _check_f32(rpos_wrapped, "referencePositions", (2,))
    n = qpos_wrapped.shape[0]
    dev = qpos_wrapped.device
    st = _lib.stream_handle(dev)
    L = lib()
    jj = indices_j.to(torch.int64).contiguous()
    nn = num_neighbors.to(torch.int32).contiguous()
    off = neighbor_offsets.to(torch.int32).contiguous()
    args = (C.byref(domain), m, ptr(qpos_wrapped), ptr(qh), n, ptr(rpos_wrapped), ptr(rh), rpos_wrapped.shape[0],
            ptr(jj), jj.numel(), ptr(nn), ptr(off))
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    check(L.dsph_pairs_filter_count(*args, ptr(counts), st), "pairs_filter_count")
    offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
    ws = WORKSPACE.get(L.dsph_scan_workspace(n), dev)
    total = C.c_int64(0)
    check(L.dsph_scan_exclusive(ptr(counts), n, ptr(offsets), C.byref(total), ptr(ws), ws.numel(), st), "scan")
    P = int(total.value)
    f = dict(dtype=torch.float32, device=dev)
    res = {"i": torch.empty(P, dtype=torch.int64, device=dev), "j": torch.empty(P, dtype=torch.int64, device=dev),
           "distances": torch.empty(P, **f), "vectors": torch.empty(P, 2, **f), "supports": torch.empty(P, **f)}
    if kernels:
        res["kernels"] = torch.empty(P, **f)
        res["gradients"] = torch.empty(P, 2, **f)
    check(L.dsph_pairs_filter_fill(*args, ptr(offsets), P, ptr(res["i"]), ptr(res["j"]), ptr(res["distances"]),
                                   ptr(res["vectors"]), ptr(res["supports"]), ptr(res.get("kernels")),
                                   ptr(res.get("gradients")), st), "pairs_filter_fill")
    res["numNeighbors"] = counts
    res["neighborOffsets"] = offsets[:-1]
    return res


def max_displacement(prior: torch.Tensor, now: torch.Tensor, supports: torch.Tensor) -> tuple[float, float]:
    """(max_k |now_k - prior_k|, min supports) of the Verlet reuse tests
    (v2/modules/neighborhood.py:470-471, 511-512); one host read."""
    _check_f32(prior, "prior positions", (2,))
    _check_f32(now, "positions", (2,))
    sup = supports.to(torch.float32).contiguous()
    L = lib()
    dev = now.device
    ws = WORKSPACE.get(L.dsph_verlet_check_workspace(), dev)
    out = (C.c_float * 2)()
    check(L.dsph_verlet_check(ptr(prior), ptr(now), now.shape[0], ptr(sup), sup.numel(), out, ptr(ws), ws.numel(),
                              _lib.stream_handle(dev)), "verlet_check")
    return float(out[0]), float(out[1])


def _moved(max_dist: float, min_support: float, verlet_scale: float) -> bool:
    # maxDistance * 2 > (verletScale - 1) * minSupport in float32 tensors
    f = np.float32
    return bool(f(max_dist) * f(2) > f(verlet_scale - 1) * f(min_support))


def buildDataStructure(positions, supports, hMax, minExtent, maxExtent, periodic, numParticles, verbose=False):
    """torchCompactRadius.buildDataStructure as the reference calls it
    (v2/modules/neighborhood.py:403: supports already scaled by verletScale):
    the reference set sorted into the dense cell grid, with the reference's
    dict keys (referencePositions, referenceSupports, sortIndex,
    sortedPositions, sortedSupports)."""
    pos = positions.detach().to(torch.float32).contiguous()
    sup = supports.detach().to(torch.float32).contiguous()
    domain = _lib.make_domain(_as_list(minExtent), _as_list(maxExtent), _as_list(periodic))
    hsearch = max(float(hMax), float(sup.max().item()) if sup.numel() else float(hMax))
    grid = CellGrid.build(pos, sup, domain, hsearch)
    si = grid.sort_index.long()
    return {"referencePositions": pos, "referenceSupports": sup, "sortIndex": si, "sortedPositions": pos[si],
            "sortedSupports": grid.sorted_h, "numParticles": int(numParticles), "_grid": grid, "_domain": domain,
            "_hsearch": hsearch}


def _candidates(queryPositions, querySupports, ds, mode="scatter"):
    """(i, j, counts, offsets[:-1]) of the query set against the data
    structure's reference set; see neighborSearchExisting."""
    grid = ds["_grid"]
    if grid is None:
        grid = CellGrid.build(ds["_positions"], ds["_supports"], ds["_domain"],
                              max(ds["_hsearch"], float(ds["_supports"].max().item())))
        ds["_grid"] = grid
    qpos = queryPositions.detach().to(torch.float32).contiguous()
    qh = querySupports.detach().to(torch.float32).contiguous() if querySupports is not None else None
    nl = search(grid, qpos, qh, mode, layout="csr")
    ex = export_pairs(nl, None, None, grid.sort_index, want=("i", "j"))
    return ex["i"], ex["j"], nl.counts, ex["offsets"][:-1]


def neighborSearchExisting(queryPositions, querySupports, datastructure, mode="scatter", searchRange=1,
                           algorithm="cpp"):
    """torchCompactRadius.neighborSearchExisting (v2/modules/neighborhood.py:526):
    (i, j) int64 of the query set against the data structure's reference set
    (rows in query order, j original reference indices), kept where
    |mod(x_i - x_j)| / h <= 1 with h the reference's (scatter), the query's
    (gather) or their mean (symmetric) support.  A data structure whose
    positions were updated since it was built is re-sorted first."""
    i, j, _, _ = _candidates(queryPositions, querySupports, datastructure, mode)
    return i, j


def updateDataStructure(referenceState, config, priorDatastructure, verbose=False):
    """v2/modules/neighborhood.py:462-478: rebuild when the reference set
    changed size or moved by more than (verletScale - 1) min(h) / 2, else keep
    the sort and take the current positions."""
    vs = _verlet_scale(config)
    ds = priorDatastructure
    pos = referenceState["positions"]
    if ds["referencePositions"].shape != pos.shape:
        return _build_ds(referenceState, config), True
    dmax, hmin = max_displacement(ds["referencePositions"], pos.detach().to(torch.float32).contiguous(),
                                  ds["referenceSupports"])
    if _moved(dmax, hmin, vs):
        return _build_ds(referenceState, config), True
    si = ds["sortIndex"]
    ds["sortedPositions"] = pos[si]
    ds["sortedSupports"] = referenceState["supports"][si]
    ds["_positions"] = pos.detach().to(torch.float32).contiguous()
    ds["_supports"] = (referenceState["supports"] * vs).detach().to(torch.float32).contiguous()
    ds["_grid"] = None
    return ds, False


def _verlet_scale(config) -> float:
    return float(config.get("neighborhood", {}).get("verletScale", 1.5))


def _build_ds(state, config):
    """getNeighborSearchDataStructure (v2/modules/neighborhood.py:402-403)."""
    d = config["domain"]
    sup = state["supports"]
    return buildDataStructure(state["positions"], sup * _verlet_scale(config), sup.max(), d["minExtent"],
                              d["maxExtent"], d.get("periodicity", d.get("periodic", False)),
                              state["positions"].shape[0])


def neighborSearch(queryState, referenceState, config, computeKernels=True, priorState=None,
                   neighborDatastructure=None, verbose=False):
    """Drop-in for v2/modules/neighborhood.py:491-593 (scatter mode, kernel
    values from the Wendland C2 kernel), with the reference's Verlet reuse:
    the candidate list (support x verletScale, 'fullIndices') is rebuilt
    when there is no prior state, the query set changed size, or it moved by
    more than (verletScale - 1) min(h) / 2 (or the data structure was
    rebuilt); otherwise the prior candidates are re-filtered at the current
    positions.  Returns ``(datastructure, neighborDict)`` with the reference's
    keys."""
    vs = _verlet_scale(config)
    if neighborDatastructure is None:
        neighborDatastructure = _build_ds(referenceState, config)
        dirty = True
    else:
        neighborDatastructure, dirty = updateDataStructure(referenceState, config, neighborDatastructure, verbose)
    qpos = queryState["positions"]
    if priorState is not None:
        if priorState["initialPositions"][0].shape != qpos.shape:
            dirty = True
        else:
            dmax, hmin = max_displacement(priorState["initialPositions"][0].detach().to(torch.float32).contiguous(),
                                          qpos.detach().to(torch.float32).contiguous(), queryState["supports"])
            if _moved(dmax, hmin, vs):
                dirty = True
    else:
        dirty = True
    if dirty:
        fi, fj, full_num, full_off = _candidates(qpos, queryState["supports"], neighborDatastructure, "scatter")
        full = (fi, fj)
        initial = (queryState["positions"], referenceState["positions"])
    else:
        full = priorState["fullIndices"]
        full_num = priorState["fullNumNeighbors"]
        full_off = priorState["fullNeighborOffsets"]
        initial = priorState["initialPositions"]
    domain = _domain_from_config(config)
    pos_x = wrap_positions(qpos.detach().to(torch.float32).contiguous(), domain)
    pos_y = wrap_positions(referenceState["positions"].detach().to(torch.float32).contiguous(), domain)
    h_i = queryState["supports"].detach().to(torch.float32).contiguous()
    h_j = referenceState["supports"].detach().to(torch.float32).contiguous()
    fr = filter_pairs(domain, "scatter", pos_x, h_i, pos_y, h_j, full[1], full_num, full_off, kernels=computeKernels)
    nd = {
        "indices": (fr["i"], fr["j"]),
        "fullIndices": full,
        "distances": fr["distances"],
        "vectors": fr["vectors"],
        "supports": fr["supports"],
        "initialPositions": initial,
        "numNeighbors": fr["numNeighbors"],
        "neighborOffsets": fr["neighborOffsets"],
        "fullNumNeighbors": full_num,
        "fullNeighborOffsets": full_off,
    }
    if computeKernels:
        nd["kernels"] = fr["kernels"]
        nd["gradients"] = fr["gradients"]
    return neighborDatastructure, nd


def searchNeighbors(state, config):
    """v2/modules/neighborhood.py:352-399: the fluid-fluid search (with the
    Verlet reuse of the prior neighborhood and data structure) into
    state['fluid']['datastructure' / 'neighborhood' / 'numNeighbors'], and
    with ``config['boundary']['active']`` the boundary-boundary,
    boundary->fluid and fluid->boundary searches into the same keys as the
    reference."""
    f = state["fluid"]
    f["datastructure"], f["neighborhood"] = neighborSearch(
        f, f, config, computeKernels=True, priorState=f.get("neighborhood"),
        neighborDatastructure=f.get("datastructure"), verbose=False)
    f["numNeighbors"] = f["neighborhood"]["numNeighbors"]
    if config["boundary"]["active"]:
        b = state["boundary"]
        b["datastructure"], b["neighborhood"] = neighborSearch(
            b, b, config, computeKernels=True, priorState=b.get("neighborhood"),
            neighborDatastructure=b.get("datastructure"), verbose=False)
        b["numNeighbors"] = b["neighborhood"]["numNeighbors"]
        _, state["fluidToBoundaryNeighborhood"] = neighborSearch(
            b, f, config, True, state.get("fluidToBoundaryNeighborhood"), f["datastructure"], False)
        _, state["boundaryToFluidNeighborhood"] = neighborSearch(
            f, b, config, True, state.get("boundaryToFluidNeighborhood"), b["datastructure"], False)
        state["boundaryToFluidNumNeighbors"] = state["boundaryToFluidNeighborhood"]["numNeighbors"]
        state["fluidToBoundaryNumNeighbors"] = state["fluidToBoundaryNeighborhood"]["numNeighbors"]
