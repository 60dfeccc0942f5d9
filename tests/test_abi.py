"""The C-ABI library loads and exports every entry point include/diffsph.h
declares (no compute: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diffsph.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(dsph_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "dsph_neighbor_count" in names and "dsph_dfsph_solve" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from diffsph_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libdiffsph_amd.so is not built (run python -m diffsph_amd.build)")
    h = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(h, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    assert not [n for n in declared_functions() if n not in _lib.SIGNATURES]


def test_host_only_entry_points_run_without_gpu():
    from diffsph_amd import _lib
    L = _lib.lib()
    assert L.dsph_abi_version() == 4
    assert L.dsph_status_string(5) == b"output buffer or workspace too small"
    dom = _lib.make_domain((-1, -1), (1, 1), (True, True))
    g = _lib.Grid()
    assert L.dsph_grid_dims(C.byref(dom), C.c_float(2.0 / 1024 * 3.989), C.byref(g)) == 0
    assert (g.ncell[0], g.ncell[1]) == (256, 256) and g.hcell == pytest.approx(2.0 / 256)
    # invalid arguments are rejected with a status, not a fault
    assert L.dsph_grid_dims(C.byref(dom), C.c_float(0.9), C.byref(g)) == 2
    assert L.dsph_grid_dims(C.byref(dom), C.c_float(-1.0), C.byref(g)) == 1
    assert L.dsph_neighbor_count(None, None, None, None, None, 0, None, None, 0, 0, 0.1, 0, None, None) == 1
    assert L.dsph_scan_workspace(10_000_000) >= 4 * (10_000_000 // 2048)
    # the MLP forward reads input rows as float4: a misaligned X is rejected
    # on the host before any launch (the pointers are never dereferenced)
    assert L.dsph_mlp_forward(C.c_void_p(0x1004), 16, C.c_void_p(0x2000), C.c_void_p(0x3000),
                              C.c_void_p(0x4000), None) == 1


def test_grid_dims_match_oracle(orc):
    from diffsph_amd import _lib
    L = _lib.lib()
    for (mn, mx, per, h) in [((-1, -1), (1, 1), (True, True), 0.0078),
                             ((0, 0), (4, 2), (False, False), 0.0156),
                             ((-1, -1), (1, 1), (True, True), 2.0 / 12 * 3.989),
                             ((0, 0), (3.7, 1.3), (True, True), 0.05)]:
        dom = _lib.make_domain(mn, mx, per)
        g = _lib.Grid()
        assert L.dsph_grid_dims(C.byref(dom), C.c_float(h), C.byref(g)) == 0
        ncx, ncy, hc = orc.grid_dims(orc.make_domain(mn, mx, per), h)
        assert (g.ncell[0], g.ncell[1], g.hcell) == (ncx, ncy, hc)
        assert g.hcell >= h


def test_batch_workspace_host_checks():
    """The batch layout is computed on the host: invalid boxes, counts and
    widths give 0 / an error status before any device call."""
    from diffsph_amd import _lib
    L = _lib.lib()
    doms = (_lib.Domain * 2)(_lib.make_domain((-0.25, -0.25), (0.25, 0.25), (True, True)),
                             _lib.make_domain((-0.5, -0.5), (0.5, 0.5), (True, True)))
    counts = (C.c_int64 * 2)(1024, 4096)
    h = 2.0 / 256 * 3.99
    ws = L.dsph_batch_workspace(doms, counts, 2, C.c_float(h), 64)
    assert ws > 4 * 64 * 5120
    assert L.dsph_batch_workspace(doms, counts, 2, C.c_float(h), 0) == 0
    # fewer than 3 cells per periodic axis
    assert L.dsph_batch_workspace(doms, counts, 2, C.c_float(0.2), 64) == 0
    walled = (_lib.Domain * 1)(_lib.make_domain((0, 0), (1, 1), (False, True)))
    assert L.dsph_batch_workspace(walled, counts, 1, C.c_float(h), 64) == 0
    b = _lib.Batch()
    assert L.dsph_batch_init(doms, counts, 2, C.c_float(h), C.c_float(h), 64, 256, None, ws, C.byref(b)) == 1
    assert L.dsph_batch_init(doms, counts, 2, C.c_float(h), C.c_float(h), 64, 100, C.c_void_p(256), ws,
                             C.byref(b)) == 1
    assert L.dsph_batch_init(doms, counts, 2, C.c_float(h), C.c_float(0.5 * h), 64, 256, C.c_void_p(256), ws,
                             C.byref(b)) == 1


def test_delta16_requires_rows_equal_reference_slots():
    """16-bit neighbor deltas decode j = i + d (mod n): a query set other
    than the grid's own sorted particles is rejected on the host before any
    launch (the fake pointers are never dereferenced)."""
    from diffsph_amd import _lib
    L = _lib.lib()
    dom = _lib.make_domain((-1, -1), (1, 1), (True, True))
    g = _lib.Grid()
    assert L.dsph_grid_dims(C.byref(dom), C.c_float(0.1), C.byref(g)) == 0
    P = C.c_void_p
    common = (C.byref(dom), C.byref(g), P(0x1000), P(0x2000), None)
    # nq != nref
    assert L.dsph_neighbor_fill16(*common, 100, P(0x3000), None, 99, 1, C.c_float(0.1), 0, P(0x4000), None, 1000,
                                  P(0x5000), P(0x6000), 1000, P(0x7000), None) == 1
    # queries not the grid's wrapped positions
    assert L.dsph_neighbor_fill16(*common, 100, P(0x3000), None, 100, 0, C.c_float(0.1), 0, P(0x4000), None, 1000,
                                  P(0x5000), P(0x6000), 1000, P(0x7000), None) == 1
    for nq, wrapped in [(99, 1), (100, 0)]:
        assert L.dsph_neighbor_fill_fixed(*common, 100, P(0x3000), None, nq, wrapped, C.c_float(0.1), 0, 67,
                                          P(0x4000), P(0x5000), P(0x6000), 2 * 64 * 67, P(0x8000), P(0x9000),
                                          P(0xa000), None) == 1


def test_search_rejects_delta16_for_foreign_queries():
    import torch
    from diffsph_amd import neighborhood
    from diffsph_amd._lib import DsphError
    with pytest.raises(DsphError, match="16-bit"):
        neighborhood.search(None, torch.zeros((4, 2)), None, delta16=object(), qpos_wrapped=False)


def test_persistent_solve_flag_is_rejected():
    """The cooperative-launch solve is not built; asking for it raises
    before any GPU work."""
    from diffsph_amd._lib import DsphError
    from diffsph_amd.dfsph import DFSPHConfig, DFSPHSolver
    import torch
    x = torch.zeros(4, 2)
    cfg = DFSPHConfig(dt=1e-3, support=0.1, persistent=True)
    with pytest.raises(DsphError, match="persistent"):
        DFSPHSolver(x, x, torch.ones(4), torch.ones(4), cfg)
