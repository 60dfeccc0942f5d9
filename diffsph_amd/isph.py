"""ISPH on the GPU: the reference's incompressible scheme over the HIP
kernels.

Follows v2/modules/incompressible.py:366-460 (simulationStep):

1. neighbor search (cell sort + sliced-ELL list, every step),
2. summation density, V = m/rho; alpha (density form), delta-SPH viscosity,
   advection = gravity + viscosity,
3. projectVelocities (:173-240): Laplacian A (per list entry, self entry
   -a_ii), source B = rho0 div_consistent(v*) - mean, Jacobi
   preconditioner -1/a_ii, x0 = 0.5 p*_prev, BiCGStab (rtol 1e-3, 512
   iterations) -> p*,
4. solveIncompressible_relaxedJacobi (:308-364): alpha/dt, omega 0.5,
   warm start 0.5 p**_prev, error mean |r| <= 5e-4, 2..256 iterations ->
   p**,
5. a = -grad_summation(p*)/rho0, v' = v* + a dt,
   shift = -dt^2 grad_symmetric(p**)/rho0,
   v = v' + grad_difference(v') shift, x = x + v' dt + shift.

The state is kept in cell order and carries a UID, as in the DFSPH solver.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import check, lib, ptr
from .neighborhood import WORKSPACE, CellGrid, search


@dataclass
class ISPHConfig:
    dt: float
    support: float
    rho0: float = 1.0
    domain_min: tuple = (-1.0, -1.0)
    domain_max: tuple = (1.0, 1.0)
    periodic: tuple = (True, True)
    viscosity: str | None = "deltaSPH_inviscid"
    visc_alpha: float = 0.01
    visc_nu: float = 0.01
    cs: float = 10.0
    visc_eps: float = 1e-6
    pi_switch: bool = False
    gravity: tuple | None = None
    # projectVelocities -> bicgstab(A, B, x0, rtol=1e-3, maxiter=512, M)
    bicg_rtol: float = 1e-3
    bicg_atol: float = 0.0
    bicg_maxiter: int = 512
    # projectVelocities(solverPrecision=...): float32 as the reference's
    # default, or float64 (constants are a near-null mode of the Laplacian;
    # a strongly non-solenoidal source can make the float32 solve diverge)
    solver_precision: str = "float32"
    # solveIncompressible_relaxedJacobi (incompressible.py:340-343)
    jacobi_omega: float = 0.5
    jacobi_min_iters: int = 2
    jacobi_max_iters: int = 256
    jacobi_threshold: float = 5e-4
    check_every: int = 8


@dataclass
class ISPHStats:
    bicg_iters: int = 0
    bicg_status: int = 0
    bicg_residual: float = 0.0
    bicg_rhs: float = 0.0
    jacobi_iters: int = 0
    jacobi_error: float = 0.0
    num_pairs: int = 0


class ISPHSolver:
    def __init__(self, positions, velocities, masses, areas, cfg: ISPHConfig, uid=None):
        if not positions.is_cuda:
            raise RuntimeError("ISPHSolver needs GPU tensors")
        dev = positions.device
        self.cfg = cfg
        self.n = positions.shape[0]
        self.domain = _lib.make_domain(cfg.domain_min, cfg.domain_max, cfg.periodic)
        f = dict(dtype=torch.float32, device=dev)
        self.pos = positions.detach().to(**f).contiguous().clone()
        self.vel = velocities.detach().to(**f).contiguous().clone()
        self.mass = masses.detach().to(**f).contiguous().clone()
        self.areas = areas.detach().to(**f).contiguous().clone()
        self.uid = (uid if uid is not None else torch.arange(self.n, device=dev, dtype=torch.int32)).to(
            torch.int32).contiguous().clone()
        n = self.n
        self.p_star = None   # 'pressure' of the previous step
        self.p_ss = None     # 'pressureIncompressible' of the previous step
        for name, shape in (("rho", (n,)), ("vol", (n,)), ("alpha_div", (n,)), ("alpha_den", (n,)),
                            ("alpha_isph", (n,)), ("div_src", (n,)), ("visc", (n, 2)), ("adv", (n, 2)),
                            ("B", (n,)), ("aii", (n,)), ("precond", (n,)), ("acc", (n, 2)), ("residual", (n,)),
                            ("vprime", (n, 2)), ("shift", (n, 2))):
            setattr(self, name, torch.zeros(shape, **f))
        self.kgrad = None
        self.lap = None
        self.nl = None
        self.step_count = 0

    def _permute(self, t, idx):
        if t is None:
            return None
        out = torch.empty_like(t)
        if t.dtype == torch.float32:
            w = 1 if t.dim() == 1 else t.shape[1]
            check(lib().dsph_gather_f32(ptr(t), ptr(idx), self.n, w, ptr(out), _lib.stream_handle(t.device)),
                  "gather")
            return out
        return t[idx.long()]

    def neighbor_step(self):
        cfg = self.cfg
        grid = CellGrid.build(self.pos, None, self.domain, cfg.support, h_uniform=cfg.support)
        si = grid.sort_index
        for name in ("pos", "vel", "mass", "areas", "uid", "p_star", "p_ss"):
            setattr(self, name, self._permute(getattr(self, name), si))
        self.nl = search(grid, grid.sorted_pos, None, "scatter", qpos_wrapped=True)
        need = max(self.nl.entries, 1)
        if self.kgrad is None or self.kgrad.shape[0] < need:
            self.kgrad = torch.empty(need + need // 8, dtype=torch.float32, device=self.pos.device)
            self.lap = torch.empty(need + need // 8, dtype=torch.float32, device=self.pos.device)
        return self.nl

    def _visc_params(self):
        cfg = self.cfg
        if cfg.viscosity is None:
            return None
        vp = _lib.ViscParams()
        vp.alpha = cfg.visc_alpha if cfg.viscosity == "deltaSPH_inviscid" else cfg.visc_nu
        vp.cs, vp.rho0, vp.eps = cfg.cs, cfg.rho0, cfg.visc_eps
        vp.pi_switch = 1 if cfg.pi_switch else 0
        vp.scheme = 0 if cfg.viscosity == "deltaSPH_inviscid" else 1
        return vp

    def step(self) -> ISPHStats:
        cfg = self.cfg
        L = lib()
        dev = self.pos.device
        st = _lib.stream_handle(dev)
        nl = self.neighbor_step()
        pairs = nl.pairs()
        out = ISPHStats(num_pairs=nl.num_pairs)
        check(L.dsph_density(C.byref(pairs), ptr(self.mass), ptr(self.mass), ptr(self.rho), ptr(self.vol), st),
              "density")
        vp = self._visc_params()
        check(L.dsph_dfsph_prepare(C.byref(pairs), ptr(self.vel), ptr(self.mass), ptr(self.areas), ptr(self.vol),
                                   ptr(self.rho), C.c_float(cfg.dt), C.c_float(cfg.rho0),
                                   C.byref(vp) if vp is not None else None, ptr(self.alpha_div),
                                   ptr(self.alpha_den), ptr(self.div_src), ptr(self.visc) if vp else None,
                                   ptr(self.kgrad), st), "dfsph_prepare")
        if vp is None:
            self.visc.zero_()
        # advection = gravity + velocity diffusion (incompressible.py:373)
        if cfg.gravity is not None:
            g = torch.tensor(cfg.gravity, dtype=torch.float32, device=dev)
            torch.add(g, self.visc, out=self.adv)
        else:
            self.adv.copy_(self.visc)
        # divergence-free system (projectVelocities)
        check(L.dsph_isph_laplacian(C.byref(pairs), ptr(self.mass), ptr(self.rho), C.c_float(cfg.dt),
                                    ptr(self.lap), ptr(self.aii), ptr(self.precond), st), "isph_laplacian")
        ws = WORKSPACE.get(L.dsph_isph_source_workspace(self.n), dev)
        check(L.dsph_isph_source(C.byref(pairs), ptr(self.vel), ptr(self.adv), ptr(self.mass), ptr(self.rho),
                                 C.c_float(cfg.dt), C.c_float(cfg.rho0), ptr(self.B), ptr(ws), ws.numel(), st),
              "isph_source")
        p_star = self.p_star.clone() if self.p_star is not None else torch.zeros_like(self.rho)
        stats = _lib.BicgstabStats()
        ws = WORKSPACE.get(L.dsph_bicgstab_workspace(self.n), dev)
        check(L.dsph_bicgstab(C.byref(pairs), ptr(self.lap), ptr(self.precond), ptr(self.B), ptr(p_star),
                              C.c_float(0.5 if self.p_star is not None else 1.0), cfg.bicg_rtol, cfg.bicg_atol,
                              cfg.bicg_maxiter, cfg.check_every, 1 if cfg.solver_precision == "float64" else 0,
                              C.byref(stats), ptr(ws), ws.numel(), st),
              "bicgstab")
        out.bicg_iters, out.bicg_status = stats.iterations, stats.status
        out.bicg_residual, out.bicg_rhs = stats.residual_norm, stats.rhs_norm
        # density-invariant relaxed Jacobi
        p_ss = torch.empty_like(self.rho)
        check(L.dsph_isph_jacobi_init(self.n, ptr(self.alpha_den), C.c_float(cfg.dt), ptr(self.p_ss),
                                      ptr(self.alpha_isph), ptr(p_ss), st), "isph_jacobi_init")
        prm = _lib.DfsphParams()
        prm.dt, prm.rho0, prm.omega = cfg.dt, cfg.rho0, cfg.jacobi_omega
        prm.error_threshold = cfg.jacobi_threshold
        prm.min_iters, prm.max_iters = cfg.jacobi_min_iters, cfg.jacobi_max_iters
        prm.clamp_pressure = 0
        prm.source_density = 1
        prm.check_every = cfg.check_every
        prm.scheme = 1
        js = _lib.SolveStats()
        ws = WORKSPACE.get(L.dsph_dfsph_workspace(self.n), dev)
        check(L.dsph_dfsph_solve(C.byref(pairs), C.byref(prm), ptr(self.kgrad), ptr(self.vel), ptr(self.adv),
                                 ptr(self.rho), ptr(self.vol), ptr(self.areas), ptr(self.alpha_isph), None,
                                 ptr(p_ss), ptr(self.acc), ptr(self.residual), C.byref(js), ptr(ws), ws.numel(), st),
              "isph_jacobi")
        out.jacobi_iters, out.jacobi_error = js.iterations, js.final_error
        # closing passes: velocities and positions
        check(L.dsph_isph_close(C.byref(pairs), ptr(self.mass), ptr(self.rho), ptr(self.vel), ptr(self.adv),
                                ptr(p_star), ptr(p_ss), C.c_float(cfg.dt), C.c_float(cfg.rho0), ptr(self.vprime),
                                ptr(self.shift), ptr(self.pos), ptr(self.vel), st), "isph_close")
        self.p_star, self.p_ss = p_star, p_ss
        self.step_count += 1
        return out

    def state_by_uid(self):
        order = torch.argsort(self.uid.long())
        return self.pos[order], self.vel[order], self.rho[order]
