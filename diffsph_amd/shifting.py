"""Particle shifting with the reference's interface.

* ``computeShifting(particleState, config, computeRho=False, scheme=...)``
  (v2/modules/shifting.py:361-460): implicit particle shifting (IPS) of a
  periodic fluid set: the right-hand side B = sum_j A_j gradW_ij and the
  kernel-Hessian blocks A_j H_ij are assembled by one HIP pass over the
  neighbor list (``dsph_ips_assemble``), and the reference's BiCGStab
  (bicgstab_shifting, :318-...) runs on the GPU: block matvecs, float32
  vector updates in the reference's order and float64 dot products, with
  its stopping rules (atol = rtol |B|, rho / omega breakdowns, the
  per-particle solver threshold).  As the reference reads every dot product
  back (``.item()`` in its conditions), the loop reads the scalars it
  branches on.
* ``solveShifting(simulationState, config)`` (:463-600): ``maxIterations``
  shift iterations ('IPS' or the delta+ scheme), each clamped to
  threshold * dx and applied to the positions, which are then restored; the
  summed displacement is returned, as the runner adds it
  (runner.py:40-42).

Fluid-only sets with a uniform support (``config['boundary']`` inactive,
no free surface).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import DsphError, check, lib, ptr
from .neighborhood import WORKSPACE, CellGrid, search, wrap_positions


@dataclass
class IPSResult:
    update: torch.Tensor      # [n, 2] = -x (the reference's update)
    rhs: torch.Tensor         # B [2n]
    code: int                 # iterations, 0 (converged at s), -10/-11 breakdowns, -12 threshold
    residual: float           # |H x - B|
    convergence: list         # |H x_k - B| per iteration
    num_pairs: int


class _Vec:
    """Device scalars of the BiCGStab loop (dot products read back)."""

    def __init__(self, n2, dev):
        self.n2 = n2
        self.dev = dev
        self.ws = torch.empty(lib().dsph_vec_reduce_workspace(n2), dtype=torch.uint8, device=dev)
        self.out = torch.zeros(1, dtype=torch.float64, device=dev)
        self.st = _lib.stream_handle(dev)

    def dot(self, a, b) -> float:
        check(lib().dsph_vec_dot(self.n2, ptr(a), ptr(b), ptr(self.out), ptr(self.ws), self.ws.numel(), self.st),
              "vec_dot")
        return float(self.out.item())

    def norm(self, a) -> float:
        return float(np.sqrt(max(self.dot(a, a), 0.0)))

    def max_norm2(self, x) -> float:
        check(lib().dsph_vec_max_norm2(self.n2 // 2, ptr(x), ptr(self.out), ptr(self.ws), self.ws.numel(),
                                       self.st), "vec_max_norm2")
        return float(self.out.item())

    def lincomb(self, out, x, a=0.0, y=None, b=0.0, z=None):
        check(lib().dsph_vec_lincomb(self.n2, ptr(x), C.c_float(float(a)), ptr(y), C.c_float(float(b)), ptr(z),
                                     ptr(out), self.st), "vec_lincomb")
        return out


def ips_shift(positions: torch.Tensor, areas: torch.Tensor, support: float, domain: _lib.Domain, rtol=1e-4,
              atol=0.0, maxiter=64, threshold=-0.5) -> IPSResult:
    """One IPS solve for the set `positions` (caller order) with rest areas
    `areas` and uniform support."""
    if not positions.is_cuda:
        raise DsphError("ips_shift needs GPU tensors")
    pos = positions.detach().to(torch.float32).contiguous()
    A = areas.detach().to(torch.float32).contiguous()
    n = pos.shape[0]
    dev = pos.device
    st = _lib.stream_handle(dev)
    L = lib()
    grid = CellGrid.build(pos, None, domain, support, h_uniform=support)
    wpos = wrap_positions(pos, domain)
    nl = search(grid, wpos, None, "scatter", qpos_wrapped=True, layout="sell")
    check(L.dsph_remap_neighbors(ptr(nl.nbr), nl.nbr.numel(), ptr(grid.sort_index), n, st), "remap_neighbors")
    pr = nl.pairs()
    pr.rpos = ptr(wpos)
    pr.rh = None
    entries = max(nl.entries, 1)
    B = torch.empty(2 * n, device=dev)
    Hs = torch.zeros(entries, 4, device=dev)
    check(L.dsph_ips_assemble(C.byref(pr), ptr(A), ptr(B), ptr(Hs), entries, st), "ips_assemble")

    V = _Vec(2 * n, dev)

    def matvec(x, y):
        check(L.dsph_ips_spmv(C.byref(pr), ptr(Hs), entries, ptr(x), ptr(y), st), "ips_spmv")
        return y

    x = torch.zeros(2 * n, device=dev)
    conv = []
    bnrm2 = V.norm(B)
    atol_ = max(float(atol), float(rtol) * bnrm2)

    def result(xv, code):
        r = matvec(xv, torch.empty_like(B))
        V.lincomb(r, r, -1.0, B)
        upd = -xv.view(n, 2)
        return IPSResult(upd.contiguous(), B, code, V.norm(r), conv, nl.num_pairs)

    if bnrm2 == 0:
        return result(x, 0)
    eps = float(np.finfo(np.float32).eps)
    rhotol = omegatol = eps * eps
    r = B.clone()
    r0 = r.clone()
    p = r.clone()
    ap = torch.empty_like(B)
    s = torch.empty_like(B)
    as_ = torch.empty_like(B)
    tmp = torch.empty_like(B)
    res = torch.empty_like(B)
    for it in range(maxiter):
        if V.norm(r) < atol_:
            return result(x, it)
        rho = V.dot(r, r0)
        if abs(rho) < rhotol:
            return result(x, -10)
        matvec(p, ap)
        rv = V.dot(ap, r0)
        if rv == 0:
            return result(x, -11)
        alpha = np.float32(rho / rv)
        V.lincomb(s, r, -float(alpha), ap)
        if V.norm(s) < atol_:
            V.lincomb(x, x, float(alpha), p)
            return result(x, 0)
        matvec(s, as_)
        omega = np.float32(V.dot(as_, s) / V.dot(as_, as_))
        V.lincomb(x, x, float(alpha), p, float(omega), s)
        rho_prev = V.dot(r, r0)
        V.lincomb(r, s, -float(omega), as_)
        beta = np.float32((V.dot(r, r0) / rho_prev) * (float(alpha) / float(omega)))
        V.lincomb(tmp, p, -float(omega), ap)
        V.lincomb(p, r, float(beta), tmp)
        if it > 0 and abs(float(omega)) < omegatol:
            return result(x, -11)
        matvec(x, res)
        V.lincomb(res, res, -1.0, B)
        conv.append(V.norm(res))
        if V.max_norm2(x) > threshold:
            return result(x, -12)
    return result(x, maxiter)


def _domain(config):
    from .schemes import _domain as dom
    mn, mx, per = dom(config)
    return _lib.make_domain(mn, mx, per)


def _support(state):
    h = state["supports"]
    hmin, hmax = float(h.min()), float(h.max())
    if hmin != hmax:
        raise DsphError("IPS shifting needs a uniform support")
    return hmin


def computeShifting(particleState, config, computeRho=False, scheme="BiCG"):
    """computeShifting (v2/modules/shifting.py:361-460) for a fluid set:
    returns (update, K, J, H, B, convergence, iters, residual) with the
    per-pair K and H left out (None) and J the assembled per-particle sum."""
    sc = config["shifting"]
    if sc.get("freeSurface", False):
        raise DsphError("computeShifting: free-surface shifting is not built")
    if torch.any(particleState.get("boundaryMarker", torch.zeros(1)) > 0):
        raise DsphError("computeShifting: fluid-only sets (no boundary particles)")
    if sc.get("initialization", "zero") != "zero":
        raise DsphError("computeShifting: only the 'zero' initialization is built")
    if computeRho or sc.get("summationDensity", False):
        raise DsphError("computeShifting: rest areas only (summationDensity False)")
    r = ips_shift(particleState["positions"], particleState["areas"], _support(particleState), _domain(config),
                  rtol=float(sc.get("rtol", 1e-4)), maxiter=int(sc.get("maxSolveIter", 64)),
                  threshold=float(sc.get("solverThreshold", -0.5)))
    J = r.rhs.view(-1, 2)
    return r.update, None, J, None, r.rhs, r.convergence, r.code, r.residual


def solveShifting(simulationState, config):
    """solveShifting (v2/modules/shifting.py:463-600) for a fluid-only
    state: returns (dx, overallStates); the state's positions are left as
    they were (the runner adds dx)."""
    if "boundary" in simulationState and simulationState["boundary"] is not None:
        raise DsphError("solveShifting: fluid-only states")
    sc = config["shifting"]
    f = simulationState["fluid"]
    init = f["positions"].detach().to(torch.float32).contiguous().clone()
    pos = init.clone()
    spacing = float(config["particle"]["dx"])
    thr = float(sc.get("threshold", 0.5)) * spacing
    states = []
    for _ in range(int(sc.get("maxIterations", 1))):
        ps = {"positions": pos, "areas": f["areas"], "supports": f["supports"], "numParticles": pos.shape[0],
              "masses": f["masses"], "densities": f["densities"], "velocities": f["velocities"]}
        if sc.get("scheme", "deltaSPH") == "IPS":
            upd, _, _, _, _, conv, iters, res = computeShifting(ps, config)
            states.append((conv, iters, res))
        else:
            upd = -_deltaplus_displacement(ps, config)
        upd = torch.clamp(upd, -thr, thr)
        pos = pos - upd
    dx = pos - init
    return dx, states


def _deltaplus_displacement(ps, config):
    """deltaPlusShifting (shifting.py:119-147) for one iteration: the
    library's delta+ kernel on a fresh neighbor list, clamped far outside
    any threshold (solveShifting clamps afterwards)."""
    from .deltasph import DeltaSPHConfig, DeltaSPHSolver
    from .schemes import _domain as dom
    sc = config["shifting"]
    mn, mx, per = dom(config)
    h = _support(ps)
    ks = config["kernel"]["kernelScale"]
    ks = float(ks.item()) if isinstance(ks, torch.Tensor) else float(ks)
    cfg = DeltaSPHConfig(support=h, dx=float(config["particle"]["dx"]), cs=float(config["fluid"]["cs"]),
                         domain_min=mn, domain_max=mx, periodic=per, shift_R=float(sc.get("R", 0.25)),
                         shift_n=float(sc.get("n", 4)), shift_cfl=float(sc.get("CFL", 1.5)),
                         shift_compute_mach=bool(sc.get("computeMach", True)), shift_threshold=1e30, dt=1e-3)
    if abs(cfg.kernel_scale - ks) > 1e-6 * ks:
        raise DsphError("deltaPlusShifting: kernelScale must be support / (2 dx)")
    s = DeltaSPHSolver(ps["positions"], ps["velocities"], ps["masses"], ps["densities"], cfg)
    out = torch.empty_like(s.pos)
    s.shift(shift_out=out)
    return out
