"""Stand-in for the third-party ``torchCompactRadius`` package the reference
imports (v2/modules/neighborhood.py:8,350), backed by the HIP cell-list
search, so an unmodified v2 driver can run on it::

    import sys, diffsph_amd.compact_radius as cr
    sys.modules["torchCompactRadius"] = cr
    sys.modules["torchCompactRadius.neighborhood"] = cr

``torchCompactRadius`` itself is not part of the reference tree; the
interface here follows its call sites (v2/modules/neighborhood.py:403,526,
v2/modules/neighborhood.py:45-52) and the reference's own fallback
(v2/neighborhoodFallback/neighborhood.py:10-113).  Pairs are kept where
|mod(x_i - x_j)| / h <= 1 in float32 (the pair predicate of the whole
library).  Only 2-D is supported.
"""
from __future__ import annotations

import torch

from . import _lib
from .neighborhood import (CellGrid, MODES, buildDataStructure, export_pairs, neighborSearchExisting,  # noqa: F401
                           search)

__all__ = ["radiusSearch", "buildDataStructure", "neighborSearchExisting"]


def radiusSearch(queryPositions, referencePositions=None, support=None, fixedSupport=None, mode="gather",
                 domainMin=None, domainMax=None, periodicity=None, hashMapLength=4096, algorithm="compact",
                 verbose=False, returnStructure=False):
    """(i, j) int64 neighbor pairs of the query set against the reference
    set (rows in query order, j original reference indices).

    ``fixedSupport`` gives one radius for all pairs; ``support`` is either
    the query supports (mode 'gather') or a (query, reference) tuple.
    Periodic axes wrap into [domainMin, domainMax) (default [0, 1))."""
    if mode not in MODES:
        raise ValueError(f"mode = {mode} not supported")
    x = queryPositions.detach().to(torch.float32).contiguous()
    y = x if referencePositions is None else referencePositions.detach().to(torch.float32).contiguous()
    if x.dim() != 2 or x.shape[1] != 2 or y.shape[1] != 2:
        raise RuntimeError(f"Unsupported dimension: {x.shape[1] if x.dim() == 2 else x.dim()}")
    dev = x.device
    qh = rh = None
    h0 = 0.0
    if fixedSupport is not None:
        h0 = float(fixedSupport)
    elif isinstance(support, torch.Tensor):
        if mode != "gather":
            raise ValueError(f"mode = {mode} != gather")
        qh = support.detach().to(torch.float32).contiguous()
    elif isinstance(support, tuple):
        qh = support[0].detach().to(torch.float32).contiguous()
        rh = support[1].detach().to(torch.float32).contiguous()
    else:
        raise ValueError("either support or fixedSupport is required")
    if qh is not None and qh.shape[0] != x.shape[0]:
        raise ValueError("querySupport does not match the query positions")
    if rh is not None and rh.shape[0] != y.shape[0]:
        raise ValueError("referenceSupport does not match the reference positions")
    mn = [0.0, 0.0] if domainMin is None else [float(v) for v in torch.as_tensor(domainMin).reshape(-1).tolist()]
    mx = [1.0, 1.0] if domainMax is None else [float(v) for v in torch.as_tensor(domainMax).reshape(-1).tolist()]
    if periodicity is None:
        per = [False, False]
    elif isinstance(periodicity, bool):
        per = [periodicity, periodicity]
    else:
        per = [bool(v) for v in torch.as_tensor(periodicity).reshape(-1).tolist()]
    domain = _lib.make_domain(mn, mx, per)
    radii = [h0] + [float(t.max().item()) for t in (qh, rh) if t is not None and t.numel()]
    grid = CellGrid.build(y, rh, domain, max(radii), h_uniform=h0)
    nl = search(grid, x, qh, mode, layout="csr")
    ex = export_pairs(nl, None, None, grid.sort_index, want=("i", "j"))
    if returnStructure:
        return ex["i"], ex["j"], grid
    return ex["i"], ex["j"]
