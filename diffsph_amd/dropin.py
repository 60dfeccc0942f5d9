"""Drop-in replacements for the reference's compiled pair ops.

Same names, argument order, dtypes and return shapes as the reference's
torch extensions; each checks its tensors the way the reference's
``getAccessor`` does (defined, contiguous, on the GPU, dtype, rank) and
raises ``RuntimeError`` with the tensor's name, then calls the C ABI on the
current stream.  Outputs are allocated here (torch's caching allocator);
the C ABI never allocates.

* :func:`sphDensity`          -- partiBench/cppOp.cpp:5-106
* :func:`sphInterpolation`    -- partiBench/cppOp.cpp:110-222,
  sphOperations/cppOp.cpp:112-132
* :func:`scatterAdd`          -- sphOperations/cppOp.cpp:364-423
* :func:`viscosityKernel`     -- partiBench/viscosityKernel.cpp:5-21
* :func:`mDBC`                -- v2/cppSrc/mDBC.cpp:3-147
* :func:`computeNeighborhood` -- v2/cppSrc/computeNeighborhood.cpp:5-94
* :func:`countNeighbors`, :func:`buildNeighborList` -- neighborhood.cpp:65-435

Pair arguments are ``(A, B)`` tuples as in the reference (only the B side
of masses/densities/quantities and the j side of indices are read, as there).
Unlike the reference, CSR segments are clipped to the pair arrays and
neighbor indices outside the reference set are skipped, so inconsistent
inputs cannot read out of bounds.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import check, lib, ptr


def _acc(t, name, dtype, ndim=None):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} is not defined")
    if not t.is_cuda:
        raise RuntimeError(f"{name} is not on the GPU")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} is not of the correct type")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} is not contiguous")
    if ndim is not None and t.dim() != ndim:
        raise RuntimeError(f"{name} is not of the correct dimension")
    return t


def _csr(numParticles, numNeighbors, neighborOffset):
    n = int(numParticles)
    if n < 0:
        raise RuntimeError("numParticles must be non-negative")
    _acc(numNeighbors, "numNeighbors", torch.int32, 1)
    _acc(neighborOffset, "neighborOffset", torch.int32, 1)
    if numNeighbors.numel() < n or neighborOffset.numel() < n:
        raise RuntimeError("numNeighbors / neighborOffset shorter than numParticles")
    return n


def sphDensity(masses, densities, quantities, indices, kernels, numParticles, numNeighbors, neighborOffset):
    """rho_i = sum_k m_j[j_k] W_k over CSR row i (partiBench/cppOp.cpp:5-106)."""
    mj = _acc(masses[1], "masses_j", torch.float32, 1)
    jj = _acc(indices[1], "indices_j", torch.int64, 1)
    W = _acc(kernels, "kernels", torch.float32, 1)
    n = _csr(numParticles, numNeighbors, neighborOffset)
    if W.numel() != jj.numel():
        raise RuntimeError("kernels and indices_j differ in length")
    out = torch.zeros(n, dtype=torch.float32, device=mj.device)
    check(lib().dsph_op_sph_density(ptr(mj), mj.numel(), ptr(jj), ptr(W), jj.numel(), n, ptr(numNeighbors),
                                    ptr(neighborOffset), ptr(out), _lib.stream_handle(mj.device)), "sphDensity")
    return out


def sphInterpolation(masses, densities, quantities, indices, kernels, numParticles, numNeighbors, neighborOffset):
    """q_i = sum_k m_j/rho_j q_j W_k over CSR row i (partiBench/cppOp.h:90-125).

    Quantities of shape [N] or [N, c] with c <= 4 (the reference handles
    [N] only, sphOperations/cppOp.cpp:104-105)."""
    mj = _acc(masses[1], "masses_j", torch.float32, 1)
    rj = _acc(densities[1], "densities_j", torch.float32, 1)
    qj = _acc(quantities[1], "quantities_j", torch.float32)
    jj = _acc(indices[1], "indices_j", torch.int64, 1)
    W = _acc(kernels, "kernels", torch.float32, 1)
    n = _csr(numParticles, numNeighbors, neighborOffset)
    if qj.dim() not in (1, 2) or (qj.dim() == 2 and not 1 <= qj.shape[1] <= 4):
        raise RuntimeError("quantities_j is not of the correct dimension")
    if rj.numel() != mj.numel() or qj.shape[0] != mj.numel():
        raise RuntimeError("masses_j, densities_j and quantities_j differ in length")
    if W.numel() != jj.numel():
        raise RuntimeError("kernels and indices_j differ in length")
    width = 1 if qj.dim() == 1 else qj.shape[1]
    out = torch.zeros((n,) + tuple(qj.shape[1:]), dtype=torch.float32, device=mj.device)
    check(lib().dsph_op_sph_interpolation(ptr(mj), ptr(rj), ptr(qj), width, mj.numel(), ptr(jj), ptr(W), jj.numel(),
                                          n, ptr(numNeighbors), ptr(neighborOffset), ptr(out),
                                          _lib.stream_handle(mj.device)), "sphInterpolation")
    return out


def scatterAdd(input, numParticles, numNeighbors, neighborOffset):
    """out_i = sum over CSR row i of input[k] (sphOperations/cppOp.cpp:364-423).
    input is [P] or [P, c] with c <= 4; the output is [numParticles, ...]."""
    x = _acc(input, "input", torch.float32)
    n = _csr(numParticles, numNeighbors, neighborOffset)
    if x.dim() not in (1, 2) or (x.dim() == 2 and not 1 <= x.shape[1] <= 4):
        raise RuntimeError("input is not of the correct dimension")
    width = 1 if x.dim() == 1 else x.shape[1]
    out = torch.zeros((n,) + tuple(x.shape[1:]), dtype=torch.float32, device=x.device)
    check(lib().dsph_op_scatter_add(ptr(x), width, x.shape[0], n, ptr(numNeighbors), ptr(neighborOffset), ptr(out),
                                    _lib.stream_handle(x.device)), "scatterAdd")
    return out


def _pair_arrays(P, **arrays):
    for name, (t, shape_tail) in arrays.items():
        _acc(t, name, torch.float32, 1 + len(shape_tail))
        if t.shape[0] != P or tuple(t.shape[1:]) != shape_tail:
            raise RuntimeError(f"{name} does not match the number of pairs")


def _rows(t, name, n, dims):
    _acc(t, name, torch.float32, dims)
    if t.shape[0] < n or (dims == 2 and t.shape[1] != 2):
        raise RuntimeError(f"{name} is not of the correct dimension")


def viscosityKernel(indices, supports, kernels, gradients, distances, vectors, numParticles, numNeighbors,
                    neighborOffset, velocities, densities, masses, particle_supports, cs, rho0, alpha, eps,
                    pi_switch):
    """deltaSPH artificial viscosity of the C++ op
    (partiBench/viscosityKernel.cpp:5-21, viscosityFn viscosityKernel.h:98-145):
    out_i = alpha h_i cs rho0 / rho_i sum_k m_j/(rho_i+rho_j) pi_ij gradW_k with
    pi_ij = (v_ij . xhat) / ((r + eps) + h^2).  Returns a [N_A, 2] tensor
    (rows >= numParticles stay zero, as the reference's zeros(...) output)."""
    jj = _acc(indices[1], "indices (B)", torch.int64, 1)
    P = jj.numel()
    _pair_arrays(P, supports=(supports, ()), gradients=(gradients, (2,)), distances=(distances, ()),
                 vectors=(vectors, (2,)))
    n = _csr(numParticles, numNeighbors, neighborOffset)
    vi, vj = velocities
    _rows(vi, "particle velocities (A)", n, 2)
    _rows(densities[0], "particle densities (A)", n, 1)
    _rows(particle_supports[0], "particle supports (A)", n, 1)
    nref = vj.shape[0] if isinstance(vj, torch.Tensor) else 0
    _rows(vj, "particle velocities (B)", nref, 2)
    _rows(masses[1], "particle masses (B)", nref, 1)
    _rows(densities[1], "particle densities (B)", nref, 1)
    out = torch.zeros(vi.shape, dtype=torch.float32, device=vi.device)
    check(lib().dsph_op_viscosity(ptr(vi), ptr(vj), ptr(masses[1]), ptr(densities[0]), ptr(densities[1]),
                                  ptr(particle_supports[0]), nref, ptr(jj), ptr(supports), ptr(gradients),
                                  ptr(distances), ptr(vectors), P, n, ptr(numNeighbors), ptr(neighborOffset),
                                  float(cs), float(rho0), float(alpha), float(eps), 1 if pi_switch else 0, ptr(out),
                                  _lib.stream_handle(vi.device)), "viscosityKernel")
    return out


def mDBC(masses, densities, positions, numBoundaryParticles, numFluidParticles, indices, numNeighbors,
         neighborOffset, kernels, kernelGradients, distances, vectors, supports, rho0):
    """mDBC wall densities of the C++ op (v2/cppSrc/mDBC.cpp:3-147,
    mDBCDensityKernel v2/cppSrc/mDBC.h:129-244).  positions = (boundary
    positions, ghost positions), both indexed by boundary particle; masses /
    densities = (., fluid).  Returns [mDBCDensity, shepardDensity]."""
    jj = _acc(indices[1], "indices (B)", torch.int64, 1)
    P = jj.numel()
    _pair_arrays(P, kernels=(kernels, ()), kernelGradients=(kernelGradients, (2,)), distances=(distances, ()),
                 vectors=(vectors, (2,)), supports=(supports, ()))
    nb = _csr(numBoundaryParticles, numNeighbors, neighborOffset)
    _rows(positions[0], "particle positions (A)", nb, 2)
    _rows(positions[1], "particle positions (B)", nb, 2)
    mj = _acc(masses[1], "particle masses (B)", torch.float32, 1)
    nref = mj.shape[0]
    _rows(densities[1], "particle densities (B)", nref, 1)
    if int(numFluidParticles) > nref:
        raise RuntimeError("numFluidParticles exceeds the fluid arrays")
    dev = numNeighbors.device
    out = torch.zeros(nb, dtype=torch.float32, device=dev)
    shep = torch.zeros(nb, dtype=torch.float32, device=dev)
    check(lib().dsph_op_mdbc(ptr(mj), ptr(densities[1]), nref, ptr(positions[0]), ptr(positions[1]), nb, ptr(jj),
                             ptr(kernels), ptr(kernelGradients), ptr(distances), ptr(vectors), ptr(supports), P,
                             ptr(numNeighbors), ptr(neighborOffset), float(rho0), ptr(out), ptr(shep),
                             _lib.stream_handle(dev)), "mDBC")
    return [out, shep]


def _domain_tensors(minDomain, maxDomain, periodicity, device):
    mn = torch.as_tensor(minDomain, dtype=torch.float32, device=device).contiguous()
    mx = torch.as_tensor(maxDomain, dtype=torch.float32, device=device).contiguous()
    per = torch.as_tensor(periodicity, device=device).to(torch.bool).contiguous()
    if mn.numel() != 2 or mx.numel() != 2 or per.numel() != 2:
        raise RuntimeError(f"Unsupported dimension: {mn.numel()}")
    return mn, mx, per


def _exclusive_scan(counts):
    """offsets[0..n] of counts and the total (one host synchronisation)."""
    n = counts.numel()
    L = lib()
    dev = counts.device
    off = torch.empty(n + 1, dtype=torch.int32, device=dev)
    ws = torch.empty(max(int(L.dsph_scan_workspace(n)), 1), dtype=torch.uint8, device=dev)
    total = C.c_int64()
    check(L.dsph_scan_exclusive(ptr(counts), n, ptr(off), C.byref(total), ptr(ws), ws.numel(),
                                _lib.stream_handle(dev)), "scan")
    return off, int(total.value)


def computeNeighborhood(indices, numParticles, numNeighbors, neighborOffset, positions, supports, minDomain,
                        maxDomain, periodicity):
    """C++ pair filter (v2/cppSrc/computeNeighborhood.cpp:5-94): keeps the
    entries of each CSR row whose periodic offset d (double moduloOp) has
    |d|^2 <= h_j^2 and returns [counts, offsets, i, j, rij, xij, hij] with
    rij = |d|/h_j and xij = d/(|d| + 1e-6 h_i)."""
    jj = _acc(indices[1], "indices (B)", torch.int64, 1)
    n = _csr(numParticles, numNeighbors, neighborOffset)
    pi, pj = positions
    hi, hj = supports
    _acc(pi, "particle positions (A)", torch.float32, 2)
    _acc(pj, "particle positions (B)", torch.float32, 2)
    _acc(hi, "particle supports (A)", torch.float32, 1)
    _acc(hj, "particle supports (B)", torch.float32, 1)
    if pi.shape[0] < n or hi.shape[0] < n or pi.shape[1] != 2 or pj.shape[1] != 2 or hj.shape[0] < pj.shape[0]:
        raise RuntimeError("particle positions / supports do not match numParticles")
    dev = pi.device
    mn, mx, per = _domain_tensors(minDomain, maxDomain, periodicity, dev)
    L = lib()
    st = _lib.stream_handle(dev)
    P = jj.numel()
    nref = pj.shape[0]
    common = (ptr(pi), ptr(hi), n, ptr(pj), ptr(hj), nref, ptr(jj), P, ptr(numNeighbors), ptr(neighborOffset),
              ptr(mn), ptr(mx), ptr(per))
    counts = torch.zeros(hi.shape[0], dtype=torch.int32, device=dev)
    check(L.dsph_op_neighborhood_count(*common, ptr(counts), st), "computeNeighborhood (count)")
    off, total = _exclusive_scan(counts)
    f = dict(dtype=torch.float32, device=dev)
    oi = torch.zeros(total, dtype=torch.int64, device=dev)
    oj = torch.zeros(total, dtype=torch.int64, device=dev)
    rij, xij, hij = torch.zeros(total, **f), torch.zeros(total, 2, **f), torch.zeros(total, **f)
    check(L.dsph_op_neighborhood_fill(*common, ptr(off), total, ptr(oi), ptr(oj), ptr(rij), ptr(xij), ptr(hij), st),
          "computeNeighborhood (fill)")
    return [counts, off[:-1], oi, oj, rij, xij, hij]


_MODES = {"scatter": 0, "gather": 1, "symmetric": 2}


def _hash_args(queryPositions, querySupport, searchRange, sortedPositions, sortedSupport, hashTable, hashMapLength,
               numCells, cellTable, hCell, maxDomain, minDomain, periodicity, mode):
    if mode not in _MODES:
        raise RuntimeError("Invalid support mode: " + str(mode))
    m = _MODES[mode]
    qp = _acc(queryPositions, "queryPositions", torch.float32, 2)
    sp = _acc(sortedPositions, "sortedPositions", torch.float32, 2)
    if qp.shape[1] != 2 or sp.shape[1] != 2:
        raise RuntimeError(f"Unsupported dimension: {qp.shape[1]}")
    qh = None if m == 0 and querySupport is None else _acc(querySupport, "querySupport", torch.float32, 1)
    sh = None if m == 1 and sortedSupport is None else _acc(sortedSupport, "sortedSupport", torch.float32, 1)
    if (qh is not None and qh.shape[0] < qp.shape[0]) or (sh is not None and sh.shape[0] < sp.shape[0]):
        raise RuntimeError("supports shorter than the positions")
    ht = _acc(hashTable, "hashTable", torch.int32, 2)
    ct = _acc(cellTable, "cellTable", torch.int64, 2)
    nc = _acc(numCells, "numCells", torch.int32, 1)
    if ht.shape[1] != 2 or ht.shape[0] < int(hashMapLength) or ct.shape[1] != 3 or nc.numel() != 2:
        raise RuntimeError("hashTable / cellTable / numCells are not of the correct dimension")
    if int(hashMapLength) < 1 or int(searchRange) < 0:
        raise RuntimeError("hashMapLength and searchRange must be positive")
    mn, mx, per = _domain_tensors(minDomain, maxDomain, periodicity, qp.device)
    return (ptr(qp), ptr(qh), qp.shape[0], int(searchRange), ptr(sp), ptr(sh), sp.shape[0], ptr(ht),
            int(hashMapLength), ptr(nc), ptr(ct), ct.shape[0], float(hCell), ptr(mn), ptr(mx), ptr(per), m), \
        (qp, qh, sp, sh, ht, ct, nc, mn, mx, per)


def countNeighbors(queryPositions, querySupport, searchRange, sortedPositions, sortedSupport, hashTable,
                   hashMapLength, numCells, cellTable, qMin, hCell, maxDomain, minDomain, periodicity, mode,
                   verbose=False):
    """Neighbor counts over a compact hash map (neighborhood.cpp:65-227):
    int32 [nQuery].  qMin is accepted for the signature; the query cell is
    floor((x - minDomain) / hCell) as in the reference's kernels."""
    args, keep = _hash_args(queryPositions, querySupport, searchRange, sortedPositions, sortedSupport, hashTable,
                            hashMapLength, numCells, cellTable, hCell, maxDomain, minDomain, periodicity, mode)
    qp = keep[0]
    counts = torch.zeros(qp.shape[0], dtype=torch.int32, device=qp.device)
    check(lib().dsph_hash_count_neighbors(*args, ptr(counts), _lib.stream_handle(qp.device)), "countNeighbors")
    return counts


def buildNeighborList(neighborCounter, neighborOffsets, neighborListLength, queryPositions, querySupport,
                      searchRange, sortedPositions, sortedSupport, hashTable, hashMapLength, numCells, cellTable,
                      qMin, hCell, maxDomain, minDomain, periodicity, mode, verbose=False):
    """Neighbor list over a compact hash map (neighborhood.cpp:253-435):
    (i, j) int32 of length neighborListLength, j indexing the sorted arrays;
    row i is written from neighborOffsets[i]."""
    args, keep = _hash_args(queryPositions, querySupport, searchRange, sortedPositions, sortedSupport, hashTable,
                            hashMapLength, numCells, cellTable, hCell, maxDomain, minDomain, periodicity, mode)
    qp = keep[0]
    _acc(neighborCounter, "neighborCounter", torch.int32, 1)
    _acc(neighborOffsets, "neighborOffsets", torch.int32, 1)
    if neighborOffsets.numel() < qp.shape[0]:
        raise RuntimeError("neighborOffsets shorter than the query set")
    P = int(neighborListLength)
    oi = torch.zeros(P, dtype=torch.int32, device=qp.device)
    oj = torch.zeros(P, dtype=torch.int32, device=qp.device)
    check(lib().dsph_hash_build_neighbors(ptr(neighborOffsets), P, *args, ptr(oi), ptr(oj),
                                          _lib.stream_handle(qp.device)), "buildNeighborList")
    return oi, oj
