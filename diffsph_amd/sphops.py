"""The reference's generic SPH operation on a neighborDict's per-pair arrays.

Drop-ins for ``sphOperation`` and ``sphOperationStates``
(src/diffSPH/v2/sphOps.py:253-303): same names, argument order, argument
meaning and return shapes.  Every operation runs as one HIP kernel
(``dsph_pp_operation``: one thread per row, the row's pairs summed in list
order); there is no CPU path.

Arguments follow the reference: ``masses`` and ``densities`` are (set A,
set B) tuples, ``quantities`` is either a (q_i, q_j) tuple of per-particle
fields or one per-pair tensor, ``neighborhood`` is the (i, j) index pair
(int64, rows sorted ascending as every neighbor search returns them),
``kernels`` W_ij [P], ``kernelGradients`` [P, 2], ``radialDistances``
r_ij / h_ij [P], ``directions`` x_ij [P, 2], ``supports`` h_ij [P].
Quantities may be scalar [N], vector [N, 2] or a 2x2 tensor [N, 2, 2].
"""
from __future__ import annotations

import ctypes as C
import weakref

import torch

from . import _lib
from ._lib import DsphError, check, lib, ptr

OPERATIONS = {"density": 0, "interpolate": 1, "gradient": 2, "divergence": 3, "divergenceConsistent": 4,
              "curl": 5, "laplacian": 6}
_GRAD_MODES = {"naive": 0, "difference": 1, "summation": 2, "symmetric": 3}
_LAP_MODES = {"naive": 0, "conserving": 4, "divergenceFree": 5, "dot": 6}

# row offsets of a row index tensor, cached per tensor object (validated by a
# weak reference to it and its version counter)
_OFFSETS: dict = {}
_OFFSETS_MAX = 16


def _f32(t, name):
    if not isinstance(t, torch.Tensor):
        raise DsphError(f"{name} is not a tensor")
    if not t.is_cuda:
        raise DsphError(f"{name} is not on the GPU")
    if t.dtype != torch.float32:
        raise DsphError(f"{name} must be float32 (got {t.dtype})")
    return t.contiguous()


def row_offsets(rows: torch.Tensor, num_rows: int) -> torch.Tensor:
    """CSR offsets [num_rows + 1] (int32) of a row-sorted pair list."""
    if rows.dtype != torch.int64 or not rows.is_cuda:
        raise DsphError("neighborhood indices must be int64 GPU tensors")
    hit = _OFFSETS.get(id(rows))
    if hit is not None and hit[0]() is rows and hit[1] == rows._version and hit[2] == num_rows:
        return hit[3]
    rows_c = rows.contiguous()
    off = torch.empty(num_rows + 1, dtype=torch.int32, device=rows.device)
    flag = torch.empty(1, dtype=torch.int32, device=rows.device)
    check(lib().dsph_pp_offsets(ptr(rows_c), rows_c.numel(), num_rows, ptr(off), ptr(flag),
                                _lib.stream_handle(rows.device)), "pp_offsets")
    if int(flag.item()) != 0:
        raise DsphError("neighborhood indices[0] must be sorted ascending and lie in [0, numParticles) "
                        "(the row order every neighbor search returns)")
    if len(_OFFSETS) >= _OFFSETS_MAX:
        for k in [k for k, v in _OFFSETS.items() if v[0]() is None] or list(_OFFSETS)[:1]:
            del _OFFSETS[k]
    _OFFSETS[id(rows)] = (weakref.ref(rows), rows._version, num_rows, off)
    return off


def _pair_arrays(neighborhood, num_rows, num_ref, kernels=None, gradients=None, distances=None, vectors=None,
                 supports=None, laplacians=None):
    i, j = neighborhood
    if j.dtype != torch.int64 or not j.is_cuda:
        raise DsphError("neighborhood indices must be int64 GPU tensors")
    if i.numel() != j.numel():
        raise DsphError("neighborhood index tensors differ in length")
    P = j.numel()
    keep = []
    pa = _lib.PairArrays()
    pa.num_rows, pa.num_ref, pa.num_pairs = num_rows, num_ref, P
    off = row_offsets(i, num_rows)
    j = j.contiguous()
    keep += [off, j]
    pa.offsets, pa.j = ptr(off), ptr(j)
    for name, t, tail in (("kernels", kernels, ()), ("gradients", gradients, (2,)), ("distances", distances, ()),
                          ("vectors", vectors, (2,)), ("supports", supports, ()), ("laplacians", laplacians, ())):
        if t is None:
            continue
        t = _f32(t, name)
        if tuple(t.shape) != (P, *tail):
            raise DsphError(f"{name} has shape {tuple(t.shape)}, expected {(P, *tail)}")
        keep.append(t)
        setattr(pa, name, ptr(t))
    return pa, keep


def _width(q, n, name):
    """Components per row of a quantity with n rows: 1, 2 or 4 (2x2)."""
    if q.shape[0] != n:
        raise DsphError(f"{name} has {q.shape[0]} rows, expected {n}")
    tail = tuple(q.shape[1:])
    if tail == ():
        return 1
    if tail == (2,):
        return 2
    if tail == (2, 2):
        return 4
    raise DsphError(f"{name}: unsupported quantity shape {tuple(q.shape)} (scalar, [N,2] or [N,2,2])")


def _run(pa, op, mode, dot, mA, mB, rA, rB, qA, qB, qP, qw, out_shape, device):
    out = torch.empty(out_shape, dtype=torch.float32, device=device)
    check(lib().dsph_pp_operation(C.byref(pa), op, mode, 1 if dot else 0, ptr(mA), ptr(mB), ptr(rA), ptr(rB),
                                  ptr(qA), ptr(qB), ptr(qP), qw, ptr(out), _lib.stream_handle(device)),
          "sphOperation")
    return out


def sphOperation(masses, densities, quantities, neighborhood, kernels, kernelGradients, radialDistances, directions,
                 supports, numParticles: int, operation: str = "interpolate", gradientMode: str = "symmetric",
                 divergenceMode: str = "div", kernelLaplacians=None) -> torch.Tensor:
    """v2/sphOps.py:253-282: 'density', 'interpolate', 'gradient',
    'divergence', 'divergenceConsistent', 'curl', 'laplacian' and
    'directLaplacian' (gradient then divergence)."""
    if operation == "directLaplacian":
        grad = sphOperation(masses, densities, quantities, neighborhood, kernels, kernelGradients, radialDistances,
                            directions, supports, numParticles, "gradient", gradientMode, divergenceMode)
        return sphOperation(masses, densities, (grad, grad), neighborhood, kernels, kernelGradients,
                            radialDistances, directions, supports, numParticles, "divergence", gradientMode,
                            divergenceMode)
    if operation not in OPERATIONS:
        raise ValueError(f"Operation {operation} not supported!")
    if divergenceMode not in ("div", "dot"):
        raise AssertionError("Only supports div F and nabla dot F")
    op = OPERATIONS[operation]
    n = int(numParticles)
    mA, mB = (_f32(masses[0], "masses[0]"), _f32(masses[1], "masses[1]"))
    nb = mB.shape[0]
    dev = mB.device
    rA = _f32(densities[0], "densities[0]") if densities is not None and densities[0] is not None else None
    rB = _f32(densities[1], "densities[1]") if densities is not None and densities[1] is not None else None
    qA = qB = qP = None
    per_pair = False
    if op != 0:
        if isinstance(quantities, (tuple, list)):
            qA = _f32(quantities[0], "quantities[0]") if quantities[0] is not None else None
            qB = _f32(quantities[1], "quantities[1]")
            qw = _width(qB, nb, "quantities[1]")
            if qA is not None and _width(qA, n, "quantities[0]") != qw:
                raise DsphError("quantities[0] and quantities[1] differ in shape")
            tail = tuple(qB.shape[1:])
        else:
            qP = _f32(quantities, "quantities")
            per_pair = True
            qw = _width(qP, neighborhood[1].numel(), "quantities")
            tail = tuple(qP.shape[1:])
    else:
        qw, tail = 1, ()

    if op == 6:  # laplacian (sphOps.py:191-250)
        if qw == 4:
            grad = sphOperation(masses, densities, quantities, neighborhood, kernels, kernelGradients,
                                radialDistances, directions, supports, n, "gradient", "difference")
            return sphOperation(masses, densities, (grad, grad), neighborhood, kernels, kernelGradients,
                                radialDistances, directions, supports, n, "divergence", "difference", "div")
        mode = _LAP_MODES.get(gradientMode, 1)  # any other mode: the default branch
        if mode == 0 and kernelLaplacians is None:
            raise AssertionError("Laplace Kernel Values required for naive sph Laplacian operation")
        if not per_pair and qA is None:
            raise DsphError("the Laplacian needs (q_i, q_j)")
        pa, keep = _pair_arrays(neighborhood, n, nb, gradients=kernelGradients, distances=radialDistances,
                                vectors=directions, supports=supports,
                                laplacians=kernelLaplacians if mode == 0 else None)
        return _run(pa, op, mode, False, mA, mB, rA, rB, qA, qB, qP, qw, (n, *tail), dev)

    mode = _GRAD_MODES.get(gradientMode, 0)  # the reference's else-branch is the naive form
    if op in (3, 4):
        if qw == 1:
            raise AssertionError("Cannot compute divergence on non vector fields!")
        out_shape = (n,) if qw == 2 else (n, 2)
    elif op == 5:
        if per_pair or qw != 2:
            raise AssertionError("Cannot compute curl on non vector fields!")
        out_shape = (n,)
    elif op == 2:
        out_shape = (n, *tail, 2)
    elif op == 1:
        out_shape = (n, *tail)
    else:
        out_shape = (n,)
    if mode == 3 and op in (2, 3, 4, 5) and (per_pair or qA is None):
        raise AssertionError("Symmetric gradient only supports two inputs for quantities!")
    if op in (2, 3, 4, 5) and not per_pair and mode in (1, 2) and qA is None:
        raise DsphError("difference / summation modes need (q_i, q_j)")
    pa, keep = _pair_arrays(neighborhood, n, nb, kernels=kernels if op in (0, 1) else None,
                            gradients=kernelGradients if op in (2, 3, 4, 5) else None)
    return _run(pa, op, mode, divergenceMode == "dot", mA, mB, rA, rB, qA, qB, qP, qw, out_shape, dev)


def sphOperationStates(stateA, stateB, quantities, neighborhood: dict, operation: str = "interpolate",
                       gradientMode: str = "symmetric", divergenceMode: str = "div") -> torch.Tensor:
    """v2/sphOps.py:285-303: the operation on two state dicts and a
    neighborDict ('density' sums m_j W_ij, the masses standing in for the
    densities as the reference passes them)."""
    if operation == "density":
        return sphOperation((stateA["masses"], stateB["masses"]), (stateA["masses"], stateB["masses"]),
                            (stateA["masses"], stateB["masses"]), neighborhood["indices"], neighborhood["kernels"],
                            None, None, None, None, stateA["numParticles"], operation="density")
    return sphOperation((stateA["masses"], stateB["masses"]), (stateA["densities"], stateB["densities"]), quantities,
                        neighborhood["indices"], neighborhood["kernels"], neighborhood["gradients"],
                        neighborhood["distances"], neighborhood["vectors"], neighborhood["supports"],
                        stateA["numParticles"], operation=operation, gradientMode=gradientMode,
                        divergenceMode=divergenceMode,
                        kernelLaplacians=neighborhood["laplacians"] if "laplacians" in neighborhood else None)
