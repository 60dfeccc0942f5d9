"""Hand-written adjoint of the delta-SPH RK4 rollout, with checkpointing.

The reference differentiates its solver with torch autograd through every
module of a step (v2/simulationSchemes/deltaPlus.py:176-422,
v2/modules/integration.py:231-307).  Here the forward rollout runs the HIP
solver (:class:`diffsph_amd.deltasph.DeltaSPHSolver`), stores the state
every `checkpoint_every` steps, and the backward pass

* re-runs each segment forward from its checkpoint, recording the sorted
  pre-step states and the per-step cell-order permutations,
* walks the steps in reverse with the adjoint of RK4
  (k1 = F(s), s2 = s + dt/2 k1, ..., s' = s + dt/6 (k1 + 2 k2 + 2 k3 + k4)):
  lambda_k1 = lambda_k4 = dt/6 lambda', lambda_k2 = lambda_k3 = dt/3 lambda',
  then stage by stage from s4 back to s the vector-Jacobian products of F
  (the HIP kernel `dsph_deltasph_rhs_vjp`), feeding each stage's cotangent
  into the stage before it,
* undoes each step's permutation.

The time steps of the forward run and the neighbor sets are constants of
the rollout.  :func:`rollout` exposes it as a torch.autograd.Function.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr
from .deltasph import DeltaSPHConfig, DeltaSPHSolver
from .neighborhood import WORKSPACE


def _zero_like(*ts):
    return [torch.zeros_like(t) for t in ts]


class Rollout:
    """Forward rollout with checkpoints and its adjoint (fluid only)."""

    def __init__(self, mass: torch.Tensor, cfg: DeltaSPHConfig, steps: int, checkpoint_every: int = 8, dts=None):
        """`dts`: optional fixed time steps (one per step); by default the
        solver's adaptive step of the forward run is used (and then held
        fixed for the backward pass)."""
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        if dts is not None and len(dts) != steps:
            raise ValueError("need one time step per step")
        self.mass = mass.detach().float().contiguous()
        self.cfg = cfg
        self.given_dts = None if dts is None else [float(d) for d in dts]
        self.steps = int(steps)
        self.every = int(checkpoint_every)
        self.dts = []
        self.checkpoints = {}
        self._fixed_cfg = dataclasses.replace(cfg, adaptive=False)

    # -- forward -------------------------------------------------------------
    def forward(self, pos0, vel0, rho0):
        s = DeltaSPHSolver(pos0, vel0, self.mass, rho0, self._fixed_cfg if self.given_dts else self.cfg)
        self.n = s.n
        self.dts = []
        self.checkpoints = {}
        for k in range(self.steps):
            if self.given_dts:
                s.dt = self.given_dts[k]
            if k % self.every == 0:
                self.checkpoints[k] = (s.pos.clone(), s.vel.clone(), s.rho.clone(), s.mass.clone(), s.uid.clone(),
                                       s.dt)
            self.dts.append(s.step())
        self.final_uid = s.uid.clone()
        order = torch.argsort(s.uid.long())
        return s.pos[order].clone(), s.vel[order].clone(), s.rho[order].clone()

    # -- one RHS VJP -----------------------------------------------------------
    def _vjp(self, sv: DeltaSPHSolver, x, u, r, lx_dot, lu_dot, lr_dot):
        """(dl/dx, dl/du, dl/drho) at the stage state (x, u, r) for the
        cotangents of (dx/dt = u, dudt, drhodt)."""
        nl, pr, wpos, grid = sv.neighbors(x)
        st = _lib.stream_handle(x.device)
        grad = torch.empty(sv.n, 2, device=x.device)
        L = lib()
        if self.cfg.diffusion:
            check(L.dsph_deltasph_normalization(C.byref(pr), None, ptr(sv.mass), ptr(r), None, None, ptr(grad), st),
                  "deltasph_normalization")
        ws = WORKSPACE.get(L.dsph_deltasph_vjp_workspace(sv.n), x.device)
        gx, gu, gr = torch.empty_like(x), torch.empty_like(u), torch.empty_like(r)
        check(L.dsph_deltasph_rhs_vjp(C.byref(pr), C.byref(sv.prm), ptr(u), ptr(sv.mass), ptr(r), ptr(grad),
                                      ptr(lx_dot), ptr(lu_dot.contiguous()), ptr(lr_dot.contiguous()), ptr(gx),
                                      ptr(gu), ptr(gr), ptr(ws), ws.numel(), st), "deltasph_rhs_vjp")
        self._keep = (nl, wpos, grid)
        return gx, gu, gr

    def _step_backward(self, sv: DeltaSPHSolver, x0, u0, r0, dt, lx, lu, lr):
        """Adjoint of one RK4 step from the sorted pre-step state."""
        n = sv.n
        dev = x0.device
        dtf = float(np.float32(dt))
        half = float(np.float32(dtf) * np.float32(0.5))
        # recompute the stage states
        states = [(x0, u0, r0)]
        ku = torch.empty(n, 2, device=dev)
        kr = torch.empty(n, device=dev)
        x, u, r = x0, u0, r0
        for c in (half, half, dtf):
            sv.rhs(x, u, r, ku, kr)
            kx = u.clone()
            xn, un, rn = torch.empty_like(x0), torch.empty_like(u0), torch.empty_like(r0)
            check(lib().dsph_rk_update(n, ptr(x0), ptr(kx), ptr(xn), ptr(u0), ptr(ku), ptr(un), ptr(r0), ptr(kr),
                                       ptr(rn), C.c_float(c), _lib.stream_handle(dev)), "rk_update")
            x, u, r = xn, un, rn
            states.append((x, u, r))
        w = [dtf / 6.0, dtf / 3.0, dtf / 3.0, dtf / 6.0]
        lk = [[w[q] * lx, w[q] * lu, w[q] * lr] for q in range(4)]  # cotangents of k_q = (dx, du, drho)
        ls = [lx.clone(), lu.clone(), lr.clone()]
        coef = [half, half, dtf]  # s_{q+1} = s + coef[q-1] k_{q-1}
        for q in (3, 2, 1, 0):
            xs, us, rs = states[q]
            gx, gu, gr = self._vjp(sv, xs, us, rs, lk[q][0], lk[q][1], lk[q][2])
            ls[0] += gx
            ls[1] += gu
            ls[2] += gr
            if q > 0:
                c = coef[q - 1]
                lk[q - 1][0] += c * gx
                lk[q - 1][1] += c * gu
                lk[q - 1][2] += c * gr
        return ls

    # -- backward ------------------------------------------------------------
    def backward(self, gx_T, gu_T, gr_T):
        """Cotangents of the final (positions, velocities, densities), in
        input order (None = 0) -> cotangents of the initial ones."""
        dev = self.mass.device
        n = self.n
        zero2 = torch.zeros(n, 2, device=dev)
        zero1 = torch.zeros(n, device=dev)
        uid = self.final_uid.long()
        lx = (gx_T if gx_T is not None else zero2).detach().float()[uid].contiguous()
        lu = (gu_T if gu_T is not None else zero2).detach().float()[uid].contiguous()
        lr = (gr_T if gr_T is not None else zero1).detach().float()[uid].contiguous()
        starts = sorted(self.checkpoints)
        for c0 in reversed(starts):
            c1 = min(c0 + self.every, self.steps)
            pos, vel, rho, mass, uid0, _ = self.checkpoints[c0]
            sv = DeltaSPHSolver(pos, vel, mass, rho, self._fixed_cfg, uid=uid0)
            rec = []

            def keep(s):
                rec.append((s.pos.clone(), s.vel.clone(), s.rho.clone(), s.mass.clone(), s.last_sort.long()))

            sv.on_sorted = keep
            for k in range(c0, c1):
                sv.dt = self.dts[k]
                sv.step()
            sv.on_sorted = None
            for k in range(c1 - 1, c0 - 1, -1):
                x0, u0, r0, m0, si = rec[k - c0]
                sv.mass = m0
                lx, lu, lr = self._step_backward(sv, x0, u0, r0, self.dts[k], lx, lu, lr)
                # undo the step's sort: sorted[s] = pre[si[s]]
                lx = torch.zeros_like(lx).index_copy_(0, si, lx)
                lu = torch.zeros_like(lu).index_copy_(0, si, lu)
                lr = torch.zeros_like(lr).index_copy_(0, si, lr)
        return lx, lu, lr


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos0, vel0, rho0, ro: Rollout):
        xT, uT, rT = ro.forward(pos0, vel0, rho0)
        ctx.ro = ro
        return xT, uT, rT

    @staticmethod
    def backward(ctx, gx, gu, gr):
        lx, lu, lr = ctx.ro.backward(gx, gu, gr)
        return lx, lu, lr, None


def rollout(pos0, vel0, rho0, mass, cfg: DeltaSPHConfig, steps: int, checkpoint_every: int = 8, dts=None):
    """Differentiable rollout: final (positions, velocities, densities) in
    input order; gradients through the hand-written adjoint."""
    ro = Rollout(mass, cfg, steps, checkpoint_every, dts)
    return _RolloutFn.apply(pos0, vel0, rho0, ro)


def kinetic_energy(vel, mass):
    """E_k = sum 0.5 m |u|^2 (runner.py:44-45, areas * densities = m)."""
    return 0.5 * (mass * (vel * vel).sum(1)).sum()
