"""Drop-in replacements for the reference's compiled pair ops.

Same names, argument order, dtypes and return shapes as the reference's
torch extensions; each checks its tensors the way the reference's
``getAccessor`` does (defined, contiguous, on the GPU, dtype, rank) and
raises ``RuntimeError`` with the tensor's name, then calls the C ABI on the
current stream.  Outputs are allocated here (torch's caching allocator);
the C ABI never allocates.

* :func:`sphDensity`       -- partiBench/cppOp.cpp:5-106
* :func:`sphInterpolation` -- partiBench/cppOp.cpp:110-222,
  sphOperations/cppOp.cpp:112-132
* :func:`scatterAdd`       -- sphOperations/cppOp.cpp:364-423

Pair arguments are ``(A, B)`` tuples as in the reference (only the B side
of masses/densities/quantities and the j side of indices are read, as there).
Unlike the reference, CSR segments are clipped to the pair arrays and
neighbor indices outside the reference set are skipped, so inconsistent
inputs cannot read out of bounds.
"""
from __future__ import annotations

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
