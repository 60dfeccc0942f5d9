"""The exclusive scan behind CSR offsets and sliced-ELL slice offsets:
exact against numpy, and a total beyond INT32_MAX is reported as an error
(DSPH_ERR_DIM -> DsphError) instead of wrapping into negative offsets that
a pair kernel would then index with."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _scan(counts_np, with_total=True):
    from diffsph_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda:0")
    n = counts_np.shape[0]
    counts = torch.from_numpy(counts_np.astype(np.int32)).to(dev)
    out = torch.empty(n + 1, dtype=torch.int32, device=dev)
    ws = torch.empty(L.dsph_scan_workspace(n), dtype=torch.uint8, device=dev)
    tot = C.c_int64(-1)
    _lib.check(L.dsph_scan_exclusive(_lib.ptr(counts), n, _lib.ptr(out), C.byref(tot) if with_total else None,
                                     _lib.ptr(ws), ws.numel(), _lib.stream_handle(dev)), "scan")
    torch.cuda.synchronize()
    return out.cpu().numpy(), tot.value


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 1_000_003])
def test_scan_matches_numpy(n):
    rng = np.random.default_rng(n)
    c = rng.integers(0, 200, n).astype(np.int64)
    out, tot = _scan(c)
    ref = np.concatenate([[0], np.cumsum(c)])
    assert np.array_equal(out.astype(np.int64), ref) and tot == ref[-1]


def test_scan_total_at_int32_max_is_accepted():
    c = np.array([2 ** 30, 2 ** 30 - 1], np.int64)
    out, tot = _scan(c)
    assert tot == 2 ** 31 - 1 and list(out) == [0, 2 ** 30, 2 ** 31 - 1]


@pytest.mark.parametrize("counts", [[2 ** 30, 2 ** 30, 5], [2 ** 31 - 1] * 3 + [0] * 5000])
def test_scan_overflow_is_an_error(counts):
    from diffsph_amd._lib import DsphError
    with pytest.raises(DsphError, match="status 2"):
        _scan(np.array(counts, np.int64))


def test_sell_layout_overflow_is_an_error():
    """128 rows of 2^25 entries: each 64-row slice needs 2^31 padded entries,
    beyond the int32 slice offsets."""
    from diffsph_amd import _lib
    from diffsph_amd._lib import DsphError
    L = _lib.lib()
    dev = torch.device("cuda:0")
    nq = 128
    counts = torch.full((nq,), 2 ** 25, dtype=torch.int32, device=dev)
    so = torch.empty(nq // 64 + 1, dtype=torch.int32, device=dev)
    ws = torch.empty(L.dsph_sell_workspace(nq), dtype=torch.uint8, device=dev)
    tot = C.c_int64(0)
    with pytest.raises(DsphError, match="status 2"):
        _lib.check(L.dsph_sell_layout(_lib.ptr(counts), nq, _lib.ptr(so), C.byref(tot), _lib.ptr(ws), ws.numel(),
                                      _lib.stream_handle(dev)), "sell_layout")
    # a layout that fits still works
    counts.fill_(40)
    _lib.check(L.dsph_sell_layout(_lib.ptr(counts), nq, _lib.ptr(so), C.byref(tot), _lib.ptr(ws), ws.numel(),
                                  _lib.stream_handle(dev)), "sell_layout")
    assert tot.value == 128 * 40 and so.cpu().tolist() == [0, 64 * 40, 128 * 40]
