"""DFSPH on the GPU: the reference's scheme over the HIP pair kernels.

Follows src/diffSPH/v2/simulationSchemes/dfsph.py:148-227 and
src/diffSPH/v2/modules/dfsph.py:52-109 with semi-implicit Euler
(src/diffSPH/v2/modules/integration.py:111-138):

1. neighbor search (cell sort + CSR list, every step),
2. summation density, V = m/rho,
3. one sweep: alpha for both solves, div(v) and delta-SPH viscosity,
4. divergence solve (omega 0.3, threshold 1e-2/rho0, warm start 0.75 p_prev),
5. density solve (omega 0.3, threshold 5e-4/rho0, warm start 0.75 p_prev,
   advection = gravity + viscosity + divergence acceleration),
6. dudt = a_p + g + visc + a_div; v += dt dudt; x += dt v.

The particle state is kept in cell order: every step re-sorts it (the
permutation is near the identity after the first step) and carries a UID
so callers can map back to the input order.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import torch

from . import _lib
from ._lib import DsphError, check, lib, ptr
from .neighborhood import TILE_RUN_STRIDE, WORKSPACE, CellGrid, Delta16Buffers, NeighborList, search


@dataclass
class DFSPHConfig:
    dt: float
    rho0: float = 1.0
    support: float = 0.0            # h (uniform)
    domain_min: tuple = (-1.0, -1.0)
    domain_max: tuple = (1.0, 1.0)
    periodic: tuple = (True, True)
    omega_div: float = 0.3
    omega_den: float = 0.3
    threshold_div: float = 1e-2     # divided by rho0 as the reference does
    threshold_den: float = 5e-4
    min_iters: int = 2
    max_iters: int = 256
    clamp_pressure: bool = True
    warm_start: float = 0.75        # p0 = 0.75 p_prev (modules/dfsph.py:70-75)
    viscosity: str | None = "deltaSPH_inviscid"
    visc_alpha: float = 0.01
    visc_nu: float = 0.01
    cs: float = 10.0
    visc_eps: float = 1e-6
    pi_switch: bool = False
    gravity: tuple | None = None
    check_every: int = 0
    # the pressure sweeps read a 16-bit modular-delta copy of the neighbor
    # list (falls back to the int32 list on the device when a delta does not
    # fit); results are identical either way
    compact_index: bool = True
    compact_limit: int = 32767      # largest |delta| kept in 16 bits (smaller values force the fallback)
    # row width of the one-pass neighbor list (no count pass); a longer row
    # falls back to the two-pass list and widens it.  0: always two-pass.
    # Odd: a power-of-two slice stride (64 x 64 entries) costs the pressure
    # sweeps 4-6 % at 1M particles, an odd one about 2 %
    neighbor_width: int = 67
    # neighbor tiles for the pressure sweeps: each workgroup of 256 rows
    # stages its neighbors' records in LDS (at most tile_cap of them, 16 B
    # each; a workgroup with more reads them from global memory) and the
    # sweeps read them there; 0: the sweeps gather from global memory.
    # Results are identical either way.  A workgroup of the 1M random flow
    # stages ~910 particles (the 3 x 3 cells around its rows' cells, ~1020
    # at most); 18 KB of LDS per workgroup keeps 8 waves per SIMD
    tile_cap: int = 1152
    # one cooperative (persistent) launch per pressure solve: disabled, see
    # DFSPHSolver.__init__
    persistent: bool = False


@dataclass
class StepStats:
    div_iters: int = 0
    den_iters: int = 0
    div_error: float = 0.0
    den_error: float = 0.0
    den_error0: float = 0.0
    num_pairs: int = 0


@dataclass
class JudgedGuess:
    """An initial divergence pressure under review ('div_init' may return
    one): the solve runs `check_after` iterations from `p`; when it has not
    converged by then, ``verdict(solver, error)`` (error: the solver's
    error metric of the last iteration, NaN when the solve stopped on one)
    returns None to continue the same solve, or a pressure to solve from
    afresh instead."""
    p: torch.Tensor
    check_after: int
    verdict: object


class DFSPHSolver:
    """Owns the particle state (cell-ordered, on the GPU) and steps it."""

    def __init__(self, positions, velocities, masses, areas, cfg: DFSPHConfig, uid=None):
        if cfg.persistent:
            # A cooperative launch holding the whole Jacobi loop (grid-wide
            # barriers, loop state on the device) matched the chunked solve
            # bit for bit at 9k particles but did not return at 1M particles
            # (1792 blocks, the occupancy the runtime reports); the cause was
            # not found, so the path is not built.
            raise DsphError("the persistent (cooperative-launch) solve is disabled")
        dev = positions.device
        if not positions.is_cuda:
            raise RuntimeError("DFSPHSolver needs GPU tensors")
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
        self.p_div = None
        self.p_den = None
        n = self.n
        self.rho = torch.empty(n, **f)
        self.vol = torch.empty(n, **f)
        self.alpha_div = torch.empty(n, **f)
        self.alpha_den = torch.empty(n, **f)
        self.div_src = torch.empty(n, **f)
        self.visc = torch.zeros(n, 2, **f)
        self.a_div = torch.empty(n, 2, **f)
        self.a_den = torch.empty(n, 2, **f)
        self.adv = torch.empty(n, 2, **f)
        self.dudt = torch.empty(n, 2, **f)
        self.residual = torch.empty(n, **f)
        self.kgrad = None   # V_j gradW_ij per neighbor-list entry, [entries, 2]
        self.wpair = None   # W_ij per neighbor-list entry (density pass), [entries]
        self.keep_pair_w = False  # set by the learned warm start, which reads wpair
        self.grid = None
        self.nl = None
        self.step_count = 0
        self._last_iters = {}
        # optional hooks: 'div_init' / 'den_init' -> callable(solver) returning
        # an initial pressure; 'after_div' -> callable(solver); 'div_split' ->
        # (k1, callable(solver, p) returning a replacement pressure after k1
        # divergence iterations)
        self.hooks = {}
        self._d16 = Delta16Buffers()
        self._tile_bufs = None
        self._nbr_width = max(int(cfg.neighbor_width), 0)

    # -- helpers ---------------------------------------------------------
    def _permute_state(self, idx):
        """The per-particle state into the new cell order, one pass over all
        arrays (dsph_gather_set)."""
        names = [k for k in ("pos", "vel", "mass", "areas", "uid", "p_div", "p_den") if getattr(self, k) is not None]
        src = [getattr(self, k) for k in names]
        dst = [torch.empty_like(t) for t in src]
        for t in src:
            if not t.is_contiguous() or t.shape[0] != self.n:
                raise DsphError("state arrays must be contiguous with one row per particle")
        words = [t.element_size() * (t.numel() // self.n) // 4 for t in src]
        k = len(src)
        check(lib().dsph_gather_set(ptr(idx), self.n, k, (C.c_void_p * k)(*[t.data_ptr() for t in src]),
                                    (C.c_void_p * k)(*[t.data_ptr() for t in dst]), (C.c_int32 * k)(*words),
                                    _lib.stream_handle(idx.device)), "gather_set")
        for name, t in zip(names, dst):
            setattr(self, name, t)

    def neighbor_step(self):
        cfg = self.cfg
        grid = CellGrid.build(self.pos, None, self.domain, cfg.support, h_uniform=cfg.support)
        self._permute_state(grid.sort_index)
        self.grid = grid
        fused = cfg.compact_index and cfg.compact_limit >= 32767
        w = self._nbr_width
        tiles = None
        if cfg.tile_cap > 0 and w > 0:
            # the fixed-width list's tiles, built by its fill; the 16-bit
            # deltas are left out there (only a workgroup without a usable
            # tile would read them, and it reads the int32 entries instead)
            tiles = (cfg.tile_cap, *self._tile_buffers(((self.n + 63) // 64) * 64 * w))
        self.nl = search(grid, grid.sorted_pos, None, "scatter", qpos_wrapped=True,
                         delta16=self._d16 if fused and tiles is None else None, layout="fixed" if w > 0 else "sell",
                         width=w, tiles=tiles)
        if w > 0 and self.nl.width == 0:
            # a row outgrew the width: this step used the two-pass list
            self._nbr_width = (int(self.nl.counts.max()) + 16) | 1
        need = max(self.nl.entries, 1)
        if self.kgrad is None or self.kgrad.shape[0] < need:
            self.kgrad = torch.empty(need + need // 8, dtype=torch.float32, device=self.pos.device)
            self.wpair = torch.empty(need + need // 8, dtype=torch.float32, device=self.pos.device)
        if cfg.compact_index and not fused:
            buf, flag = self._d16.get(need, self.pos.device)
            self.nl.attach_compact16(buf, flag, cfg.compact_limit)
        if cfg.tile_cap > 0 and self.nl.tiles is None:
            self.nl.attach_tiles(cfg.tile_cap, *self._tile_buffers(need))
        return self.nl

    def _tile_buffers(self, need: int):
        """(slot, tile_runs, tile_n) with room for `need` entries."""
        nt = (self.n + 255) // 256
        t = self._tile_bufs
        if t is None or t[0].numel() < need:
            dev = self.pos.device
            t = self._tile_bufs = (torch.empty(need + need // 8, dtype=torch.int16, device=dev),
                                   torch.empty(nt * TILE_RUN_STRIDE, dtype=torch.int32, device=dev),
                                   torch.empty(nt, dtype=torch.int32, device=dev))
        return t

    def tile_stats(self) -> dict:
        """Neighbor-tile occupancy of the current list (synchronises): the
        share of workgroups whose neighbors fit their tile and the mean /
        largest neighbor count per workgroup."""
        if self.nl is None or self.nl.tiles is None:
            return {}
        cap = self.nl.tiles[3]
        tn = self.nl.tiles[2][: (self.n + 255) // 256].cpu()
        ok = (tn >= 0) & (tn <= cap)
        v = tn[tn >= 0].float()
        return {"cap": cap, "tiled": float(ok.float().mean()), "mean": float(v.mean()) if v.numel() else 0.0,
                "max": int(tn.max())}

    def _visc_params(self):
        cfg = self.cfg
        if cfg.viscosity is None:
            return None
        vp = _lib.ViscParams()
        vp.alpha = cfg.visc_alpha if cfg.viscosity == "deltaSPH_inviscid" else cfg.visc_nu
        vp.cs = cfg.cs
        vp.rho0 = cfg.rho0
        vp.eps = cfg.visc_eps
        vp.pi_switch = 1 if cfg.pi_switch else 0
        vp.scheme = 0 if cfg.viscosity == "deltaSPH_inviscid" else 1
        return vp

    def _solve(self, pairs, density: bool, p: torch.Tensor, accel: torch.Tensor, adv=None, min_iters=None,
               max_iters=None, check_every=None):
        cfg = self.cfg
        prm = _lib.DfsphParams()
        prm.dt = cfg.dt
        prm.rho0 = cfg.rho0
        prm.omega = cfg.omega_den if density else cfg.omega_div
        prm.error_threshold = (cfg.threshold_den if density else cfg.threshold_div) / cfg.rho0
        prm.min_iters = cfg.min_iters if min_iters is None else min_iters
        prm.max_iters = cfg.max_iters if max_iters is None else max_iters
        prm.clamp_pressure = 1 if cfg.clamp_pressure else 0
        prm.source_density = 1 if density else 0
        # Queue as many iterations as the previous solve of this kind needed
        # before the first host check (iterations past convergence are no-ops).
        prev = self._last_iters.get(density, 0)
        prm.check_every = cfg.check_every if cfg.check_every > 0 else max(4, min(prev + 2, prm.max_iters))
        if check_every is not None:
            prm.check_every = check_every
        stats = _lib.SolveStats()
        L = lib()
        ws = WORKSPACE.get(L.dsph_dfsph_workspace(self.n), p.device)
        alpha = self.alpha_den if density else self.alpha_div
        check(L.dsph_dfsph_solve(C.byref(pairs), C.byref(prm), ptr(self.kgrad), ptr(self.vel), ptr(adv),
                                 ptr(self.rho),
                                 ptr(self.vol), ptr(self.areas), ptr(alpha), ptr(self.div_src), ptr(p),
                                 ptr(accel), ptr(self.residual), C.byref(stats), ptr(ws), ws.numel(),
                                 _lib.stream_handle(p.device)), "dfsph_solve")
        return stats

    def probe_guess(self, p0: torch.Tensor, iterations: int = 1):
        """Run `iterations` Jacobi iterations of the divergence solve from p0
        on copies (the solver state is unchanged).  Returns (error, residual,
        natural) of the last iteration: the solver's one-sided error metric
        mean(max(r/rho0, -thr)), the two-sided mean |r| / rho0 of the
        residual r at that iteration's start, and the natural residual of the
        clamped Jacobi fixed point, mean |(p_new - p) alpha| / (omega rho0),
        which is zero exactly at a solution with p >= 0."""
        cfg = self.cfg
        prm = _lib.DfsphParams()
        prm.dt, prm.rho0, prm.omega = cfg.dt, cfg.rho0, cfg.omega_div
        prm.error_threshold = cfg.threshold_div / cfg.rho0
        prm.min_iters, prm.max_iters = iterations, iterations
        prm.clamp_pressure = 1 if cfg.clamp_pressure else 0
        prm.source_density = 0
        L = lib()
        ws = WORKSPACE.get(L.dsph_dfsph_workspace(self.n), p0.device)
        pairs = self.nl.pairs()
        st = _lib.stream_handle(p0.device)
        p = p0.clone()
        acc = torch.empty(self.n, 2, device=p.device)
        res = torch.empty(self.n, device=p.device)
        err = C.c_double(0.0)
        if iterations > 1:
            check(L.dsph_dfsph_iterate(C.byref(pairs), C.byref(prm), ptr(self.kgrad), ptr(self.rho), ptr(self.vol),
                                       ptr(self.alpha_div), ptr(self.div_src), ptr(p), ptr(acc), None,
                                       iterations - 1, None, ptr(ws), ws.numel(), st), "dfsph_iterate")
        p_last = p.clone()
        check(L.dsph_dfsph_iterate(C.byref(pairs), C.byref(prm), ptr(self.kgrad), ptr(self.rho), ptr(self.vol),
                                   ptr(self.alpha_div), ptr(self.div_src), ptr(p), ptr(acc), ptr(res), 1,
                                   C.byref(err), ptr(ws), ws.numel(), st), "dfsph_iterate")
        nat = ((p - p_last) * self.alpha_div).abs().mean() / (cfg.omega_div * cfg.rho0)
        return float(err.value), float(res.abs().mean()) / cfg.rho0, float(nat)

    def probe(self, p0: torch.Tensor, iterations: int):
        """(error, p): `iterations` Jacobi iterations of the divergence solve
        from p0 on a copy (the solver state is unchanged), p the pressure
        after them and error the solver's error metric of the last one (one
        host synchronisation)."""
        cfg = self.cfg
        prm = _lib.DfsphParams()
        prm.dt, prm.rho0, prm.omega = cfg.dt, cfg.rho0, cfg.omega_div
        prm.error_threshold = cfg.threshold_div / cfg.rho0
        prm.min_iters, prm.max_iters = iterations, iterations
        prm.clamp_pressure = 1 if cfg.clamp_pressure else 0
        prm.source_density = 0
        L = lib()
        ws = WORKSPACE.get(L.dsph_dfsph_workspace(self.n), p0.device)
        p = p0.clone()
        acc = torch.empty(self.n, 2, device=p.device)
        err = C.c_double(0.0)
        check(L.dsph_dfsph_iterate(C.byref(self.nl.pairs()), C.byref(prm), ptr(self.kgrad), ptr(self.rho),
                                   ptr(self.vol), ptr(self.alpha_div), ptr(self.div_src), ptr(p), ptr(acc), None,
                                   iterations, C.byref(err), ptr(ws), ws.numel(), _lib.stream_handle(p.device)),
              "dfsph_iterate")
        return float(err.value), p

    def first_iteration_error(self, p0: torch.Tensor) -> float:
        """Solver error metric after one divergence iteration from p0."""
        return self.probe_guess(p0)[0]

    # -- one step --------------------------------------------------------
    def step(self) -> StepStats:
        """One DFSPH step: search, density, solves, semi-implicit Euler."""
        out = self._forces()
        cfg = self.cfg
        g = None
        if cfg.gravity is not None:
            g = (C.c_float * 2)(*[float(v) for v in cfg.gravity])  # host pointer
        check(lib().dsph_integrate_semi_implicit(ptr(self.pos), ptr(self.vel), ptr(self.a_den), ptr(self.visc),
                                                 ptr(self.a_div), None, g, self.n, C.c_float(cfg.dt),
                                                 ptr(self.dudt), _lib.stream_handle(self.pos.device)), "integrate")
        self.step_count += 1
        return out

    def derivatives(self) -> StepStats:
        """The step without its integration: fills every per-step field and
        dudt = a_p + g + visc + a_div (v2/simulationSchemes/dfsph.py:209), in
        the state's (cell) order; positions and velocities are unchanged."""
        out = self._forces()
        d = self.a_den
        if self.cfg.gravity is not None:
            d = d + torch.tensor(self.cfg.gravity, dtype=torch.float32, device=d.device)
        torch.add(torch.add(d, self.visc), self.a_div, out=self.dudt)
        return out

    def _forces(self) -> StepStats:
        cfg = self.cfg
        L = lib()
        st = _lib.stream_handle(self.pos.device)
        nl = self.neighbor_step()
        pairs = nl.pairs()
        out = StepStats(num_pairs=nl.num_pairs)
        if self.keep_pair_w:
            # W_ij per list entry is kept for the learned warm start's features
            check(L.dsph_density_w(C.byref(pairs), ptr(self.mass), ptr(self.mass), ptr(self.rho), ptr(self.vol),
                                   ptr(self.wpair), st), "density")
        else:
            check(L.dsph_density(C.byref(pairs), ptr(self.mass), ptr(self.mass), ptr(self.rho), ptr(self.vol),
                                 st), "density")
        vp = self._visc_params()
        check(L.dsph_dfsph_prepare(C.byref(pairs), ptr(self.vel), ptr(self.mass), ptr(self.areas), ptr(self.vol),
                                   ptr(self.rho), C.c_float(cfg.dt), C.c_float(cfg.rho0),
                                   C.byref(vp) if vp is not None else None, ptr(self.alpha_div),
                                   ptr(self.alpha_den), ptr(self.div_src), ptr(self.visc) if vp else None,
                                   ptr(self.kgrad), st),
              "dfsph_prepare")
        if vp is None:
            self.visc.zero_()
        # divergence solve (advection = 0: v* = v)
        judged = None
        if "div_init" in self.hooks:
            p_div = self.hooks["div_init"](self)
            if isinstance(p_div, JudgedGuess):
                judged, p_div = p_div, p_div.p
            p_div = p_div.contiguous()
        else:
            p_div = (cfg.warm_start * self.p_div) if self.p_div is not None else torch.zeros_like(self.rho)
        split = self.hooks.get("div_split")
        if judged is not None and split is None and 0 < judged.check_after < cfg.max_iters:
            # the guess's first iterations are those of the solve itself (two
            # phases, as below): kept, the solve goes on; rejected, it starts
            # again from the verdict's pressure
            k1 = judged.check_after
            s = self._solve(pairs, False, p_div, self.a_div, None, max_iters=k1, check_every=k1)
            iters = s.iterations
            if not s.converged:
                # k1 iterations without convergence, or fewer when the error
                # is not a number (the stopping rule ends the solve then)
                restart = judged.verdict(self, s.final_error)
                if restart is None:
                    prev = self._last_iters.get(False, 0)
                    self._last_iters[False] = max(prev - k1, 0)
                    s = self._solve(pairs, False, p_div, self.a_div, None, min_iters=max(cfg.min_iters - k1, 1),
                                    max_iters=cfg.max_iters - k1)
                    iters += s.iterations
                else:
                    p_div = restart.contiguous()
                    s = self._solve(pairs, False, p_div, self.a_div, None)
                    iters = s.iterations
        elif split is None:
            s = self._solve(pairs, False, p_div, self.a_div, None)
            iters = s.iterations
        else:
            # two-phase solve: k1 iterations, then the hook may replace the
            # pressure (a learned correction), then Jacobi to the same
            # stopping rule (at least one more iteration)
            k1, fn = split
            s = self._solve(pairs, False, p_div, self.a_div, None, max_iters=k1, check_every=k1)
            iters = s.iterations
            if s.iterations >= k1 and not s.converged and cfg.max_iters > k1:
                p_div = fn(self, p_div).contiguous()
                prev = self._last_iters.get(False, 0)
                self._last_iters[False] = max(prev - k1, 0)
                s = self._solve(pairs, False, p_div, self.a_div, None, min_iters=1, max_iters=cfg.max_iters - k1)
                iters += s.iterations
        out.div_iters, out.div_error = iters, s.final_error
        self._last_iters[False] = iters
        self.p_div = p_div
        if "after_div" in self.hooks:
            self.hooks["after_div"](self)
        # advection for the density solve: g + visc + a_div (dfsph.py:199)
        adv = self.adv
        torch.add(self.visc, self.a_div, out=adv)
        if cfg.gravity is not None:
            g = torch.tensor(cfg.gravity, dtype=torch.float32, device=adv.device)
            adv.copy_((g + self.visc) + self.a_div)
        if "den_init" in self.hooks:
            p_den = self.hooks["den_init"](self)
        else:
            p_den = (cfg.warm_start * self.p_den) if self.p_den is not None else torch.zeros_like(self.rho)
        s = self._solve(pairs, True, p_den, self.a_den, adv)
        out.den_iters, out.den_error, out.den_error0 = s.iterations, s.final_error, s.initial_error
        self._last_iters[True] = s.iterations
        self.p_den = p_den
        return out

    def density_error(self):
        """(mean, max) |rho/rho0 - 1| of the current summation density."""
        out = (C.c_double * 2)()
        ws = WORKSPACE.get(1 << 16, self.rho.device)
        check(lib().dsph_density_error(ptr(self.rho), self.n, C.c_float(self.cfg.rho0), out, ptr(ws), ws.numel(),
                                       _lib.stream_handle(self.rho.device)), "density_error")
        return float(out[0]), float(out[1])

    def state_by_uid(self):
        """Positions, velocities and densities in input (UID) order."""
        order = torch.argsort(self.uid.long())
        return self.pos[order], self.vel[order], self.rho[order]


# ---------------------------------------------------------------------------
# dfsphSolve on a reference neighborDict (v2/modules/dfsph.py:52-109)
# ---------------------------------------------------------------------------

def dfsphSolve(stateA, stateB, neighborhood, config):
    """Drop-in for the reference's ``dfsphSolve(stateA, stateB, neighborhood,
    config)`` (v2/modules/dfsph.py:52-109): the same state keys, config keys,
    warm start (0.75 of ``pressureIncompressible`` / ``pressureDivergence``),
    stopping rule, clamping and side effects (``predictedVelocities``,
    ``actualArea``, ``sourceTerm``, ``alpha``, ``pressureA``, ``pressureB``,
    ``pressureAccel``, ``residual``, ``convergence_<sourceTerm>``); returns
    the pressure acceleration.  It runs on the neighborDict's own per-pair
    gradients: alpha, the source, the pressure acceleration and each Jacobi
    update are HIP kernels over the pair arrays, and every iteration reads
    the float64 convergence metric back, as the reference's ``.item()`` does
    (the fused solver of :class:`DFSPHSolver` keeps the loop on the
    device)."""
    from .sphops import _pair_arrays, sphOperationStates
    dfc = config["dfsph"]
    solve_density = dfc["sourceTerm"] == "density"
    dt = float(config["timestep"]["dt"])
    rho0 = float(config["fluid"]["rho0"])
    dt_pow = dt * dt if solve_density else dt
    n = int(stateA["numParticles"])
    nb = int(stateB["numParticles"])
    vel = stateA["velocities"]
    stateA["predictedAcceleration"] = torch.zeros_like(vel)
    stateA["predictedVelocities"] = vel + dt * stateA["advection"]
    stateA["actualArea"] = stateA["masses"] / stateA["densities"]
    div = sphOperationStates(stateA, stateB, (stateA["predictedVelocities"], stateB["predictedVelocities"]),
                             neighborhood, operation="divergence", gradientMode="difference")
    if solve_density:
        stateA["sourceTerm"] = 1 - stateA["areas"] / stateA["actualArea"] - dt * div
    else:
        stateA["sourceTerm"] = div
    dev = vel.device
    stateA["pressureB"] = torch.zeros(n, device=dev)
    stateA["pressureA"] = torch.zeros(n, device=dev)
    if "pressureIncompressible" in stateA and solve_density:
        stateA["pressureB"] = 0.75 * stateA["pressureIncompressible"]
        stateA["pressureA"] = 0.75 * stateA["pressureIncompressible"]
    elif "pressureDivergence" in stateA and not solve_density:
        stateA["pressureB"] = 0.75 * stateA["pressureDivergence"]
        stateA["pressureA"] = 0.75 * stateA["pressureDivergence"]
    pa, keep = _pair_arrays(neighborhood["indices"], n, nb, gradients=neighborhood["gradients"])
    L = lib()
    st = _lib.stream_handle(dev)
    alpha = torch.empty(n, device=dev)
    f32 = lambda t: t.contiguous().to(torch.float32)  # noqa: E731
    actA, actB = f32(stateA["actualArea"]), f32(stateB["actualArea"])
    check(L.dsph_pp_dfsph_alpha(C.byref(pa), ptr(actA), ptr(actB), ptr(f32(stateA["areas"])),
                                ptr(f32(stateB["areas"])), C.c_float(dt_pow), C.c_float(rho0), ptr(alpha), st),
          "dfsph alpha")
    stateA["alpha"] = alpha
    min_iters, max_iters = int(dfc["minIters"]), int(dfc["maxIters"])
    thr = float(dfc["errorThreshold"])
    omega = float(dfc["omega"])
    clampp = 1 if dfc["clampPressure"] else 0
    ws = WORKSPACE.get(L.dsph_pp_dfsph_update_workspace(n), dev)
    err_dev = torch.zeros(1, dtype=torch.float64, device=dev)

    def pressure_accel():
        g = sphOperationStates(stateA, stateB, (stateA["pressureB"], stateB["pressureB"]), neighborhood,
                               operation="gradient", gradientMode="summation")
        return -g / stateA["densities"].view(-1, 1)

    errors = []
    it = 0
    error = 0.0
    while it < max_iters and (it < min_iters or error > thr):
        stateA["pressureAccel"] = pressure_accel()
        stateA["pressureA"] = stateA["pressureB"].clone()
        p_new = torch.empty(n, device=dev)
        res = torch.empty(n, device=dev)
        check(L.dsph_pp_dfsph_update(C.byref(pa), ptr(f32(stateB["masses"])), ptr(f32(stateB["densities"])),
                                     ptr(f32(stateA["pressureAccel"])), ptr(f32(stateB["pressureAccel"])),
                                     ptr(f32(stateA["sourceTerm"])), ptr(alpha), ptr(f32(stateA["pressureA"])),
                                     C.c_float(dt_pow), C.c_float(omega), clampp, C.c_float(rho0),
                                     C.c_float(thr), ptr(p_new), ptr(res), ptr(err_dev), ptr(ws), ws.numel(), st),
              "dfsph update")
        stateA["pressureB"], stateA["residual"] = p_new, res
        error = float(err_dev.item())
        errors.append(error)
        it += 1
    a_p = pressure_accel()
    stateA[f"convergence_{dfc['sourceTerm']}"] = errors
    return a_p
