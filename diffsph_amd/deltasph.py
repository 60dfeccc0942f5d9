"""delta-SPH (weakly compressible) on the GPU: the reference's deltaPlus
scheme over the HIP pair kernels.

Follows v2/simulationSchemes/deltaPlus.py:176-422 (fluid only, density by
continuity) integrated with RK4 (v2/modules/integration.py:231-307) and the
adaptive time step of v2/modules/timestep.py:8-39:

1. neighbor search on the stage positions (rows in the state's order,
   entries mapped back to state indices),
2. normalization matrices, their pseudo-inverses L and the renormalized
   density gradient (one kernel, two passes over each row),
3. one fused pass: density diffusion, viscosity, continuity, EOS, pressure
   acceleration, gravity -> dudt, drho/dt (dx/dt is the stage velocity),
4. RK4 stage updates and the final combination,
5. dt = computeTimestep(max |dudt|) for the next step.

The state is re-sorted into cell order at the start of every step and
carries a UID so callers can map back to the input order.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import DsphError, check, lib, ptr
from .neighborhood import WORKSPACE, CellGrid, search, wrap_positions

VISCOSITY = {None: -1, "none": -1, "deltaSPH_inviscid": 0, "deltaSPH_viscid": 1}
EOS = {"isoThermal": 0, "Tait": 1}
# wall density: "mDBCDensity" = v2/modules/mDBC.py:50-136 (pinv, unclamped);
# "mDBC_op" = the C++ op v2/cppSrc/mDBC.h:129-244 (Cramer, clamped at rho0)
MDBC = {"mDBCDensity": 0, "mDBC_op": 1}


@dataclass
class DeltaSPHConfig:
    support: float                  # h (uniform)
    dx: float                       # particle spacing (kernelScale = h / (2 dx), parameters.py:205)
    rho0: float = 1.0
    cs: float = 10.0
    domain_min: tuple = (-1.0, -1.0)
    domain_max: tuple = (1.0, 1.0)
    periodic: tuple = (True, True)
    delta: float = 0.1              # density diffusion (densityDiffusion.py:82)
    diffusion: bool = True
    viscosity: str | None = "deltaSPH_viscid"
    nu: float = 0.01
    alpha: float = 0.01
    eps: float = 1e-6
    pi_switch: bool = False
    eos: str = "isoThermal"
    kappa: float = 1.3
    gravity: tuple | None = None
    # computeTimestep (timestep.py:46-54 defaults)
    dt: float | None = None         # None: the first step uses computeTimestep without the acceleration term
    adaptive: bool = True
    cfl: float = 1.5
    max_dt: float = 0.1
    min_dt: float = 0.001
    viscosity_constraint: bool = True
    acceleration_constraint: bool = True
    acoustic_constraint: bool = True
    # wall viscosity (boundaryDiffusion, boundaryViscosity.py:219-241)
    boundary_viscosity: str | None = "deltaSPH_inviscid"
    boundary_alpha: float = 0.01
    boundary_nu: float = 0.01
    boundary_eps: float = 1e-6
    boundary_pi_switch: bool = False
    mdbc: str = "mDBCDensity"
    # "global": |dudt| over all components, as computeTimestep evaluates
    # torch.linalg.norm(dudt[~isnan(dudt)]) on a flattened tensor
    # (timestep.py:22-25); "max": the largest per-particle |dudt|
    accel_norm: str = "global"
    # cell side of the search grids = 1.001 * support * cell_scale (>= 1);
    # changes the order of the neighbor entries, not the neighbor sets
    cell_scale: float = 1.0
    # row width of the one-pass neighbor lists (no count pass); a longer row
    # falls back to the two-pass list for that search and widens the width
    # of that list kind.  0: always two-pass
    neighbor_width: int = 67
    # delta+ particle shifting after every step (solveShifting with scheme
    # 'deltaSPH', runner.py:40-42; shifting.py:735-765 defaults); off in the
    # harness configurations, as in the reference's benchmark notebooks
    shifting: bool = False
    shift_R: float = 0.25
    shift_n: float = 4.0
    shift_cfl: float = 1.5
    shift_compute_mach: bool = True
    shift_threshold: float = 0.5

    @property
    def kernel_scale(self) -> float:
        return float(np.float32(self.support) / (np.float32(2.0) * np.float32(self.dx)))

    def params(self) -> _lib.DeltaSPHParams:
        p = _lib.DeltaSPHParams()
        if self.viscosity not in VISCOSITY:
            raise ValueError(f"unsupported viscosity scheme {self.viscosity}")
        if self.eos not in EOS:
            raise ValueError(f"unsupported EOS {self.eos}")
        p.rho0, p.cs, p.delta, p.kernel_scale = self.rho0, self.cs, self.delta, self.kernel_scale
        p.visc_coeff = self.nu if self.viscosity == "deltaSPH_viscid" else self.alpha
        p.visc_eps = self.eps
        p.eos_kappa = self.kappa
        g = self.gravity or (0.0, 0.0)
        p.gravity[0], p.gravity[1] = float(g[0]), float(g[1])
        p.viscosity = VISCOSITY[self.viscosity]
        p.pi_switch = 1 if self.pi_switch else 0
        p.eos = EOS[self.eos]
        p.diffusion = 1 if self.diffusion else 0
        return p

    def timestep_nu(self) -> float:
        """The viscosity constraint's nu (timestep.py:13-14)."""
        if self.viscosity == "deltaSPH_viscid":
            return self.nu
        return self.alpha * self.cs * self.support / (2 * (2 + 2))


def timestep_constants(cfg: DeltaSPHConfig):
    """The state-independent bounds of computeTimestep (timestep.py:13-20)
    in float32: (viscous dt, acoustic dt)."""
    F = np.float32
    h, ks = F(cfg.support), F(cfg.kernel_scale)
    dt_v = F(0.125) * h * h / F(cfg.timestep_nu()) / (ks * ks)
    dt_c = F(cfg.cfl) * h / F(cfg.cs) / ks
    return dt_v, dt_c


def timestep_params(cfg: DeltaSPHConfig) -> _lib.TimestepParams:
    """The computeTimestep constants for the device form of the step
    (dsph_timestep_dev, and per domain the batched step)."""
    dt_v, dt_c = timestep_constants(cfg)
    tp = _lib.TimestepParams()
    tp.support = cfg.support
    tp.kernel_scale = cfg.kernel_scale
    tp.dt_visc = float(dt_v)
    tp.dt_acoustic = float(dt_c)
    tp.max_dt = float(np.float32(cfg.max_dt))
    tp.min_dt = float(np.float32(cfg.min_dt))
    tp.adaptive = 1 if cfg.adaptive else 0
    tp.viscosity_constraint = 1 if cfg.viscosity_constraint else 0
    tp.acceleration_constraint = 1 if cfg.acceleration_constraint else 0
    tp.acoustic_constraint = 1 if cfg.acoustic_constraint else 0
    return tp


def compute_timestep(cfg: DeltaSPHConfig, dt_prev: float, max_accel: float | None) -> float:
    """computeTimestep (v2/modules/timestep.py:8-39) in float32 arithmetic."""
    F = np.float32
    h, ks = F(cfg.support), F(cfg.kernel_scale)
    dt_v, dt_c = timestep_constants(cfg)
    dt_a = F(cfg.max_dt) if max_accel is None else F(0.25) * np.sqrt(h / (F(max_accel) + F(1e-7))) / ks
    new = F(dt_prev)
    if cfg.viscosity_constraint:
        new = dt_v
    if cfg.acceleration_constraint:
        new = min(new, dt_a)
    if cfg.acoustic_constraint:
        new = min(new, dt_c)
    new = min(new, F(cfg.max_dt))
    new = max(new, F(cfg.min_dt))
    return float(F(new))


class Walls:
    """Static wall particles with mDBC densities (v2/modules/mDBC.py).

    The set is sorted once into its own cell order (so sorted slots are wall
    indices), its ghosts are mirrored into the fluid once, and its density
    is re-extrapolated from the fluid at every right-hand side."""

    def __init__(self, positions, distances, normals, areas, cfg: "DeltaSPHConfig", domain):
        dev = positions.device
        f = dict(dtype=torch.float32, device=dev)
        pos = positions.detach().to(**f).contiguous()
        self.grid = CellGrid.build(pos, None, domain, cfg.support, h_uniform=cfg.support)
        si = self.grid.sort_index.long()
        self.n = pos.shape[0]
        self.pos = self.grid.sorted_pos
        self.dist = distances.detach().to(**f)[si].contiguous()
        self.normal = normals.detach().to(**f)[si].contiguous()
        self.mass = (areas.detach().to(**f).expand(self.n)[si] * cfg.rho0).contiguous()
        self.vel = torch.zeros(self.n, 2, **f)
        self.rho = torch.full((self.n,), float(cfg.rho0), **f)
        self.shepard = torch.empty(self.n, **f)
        self.ghost = torch.empty(self.n, 2, **f)
        check(lib().dsph_boundary_ghosts(ptr(self.pos), ptr(self.dist), ptr(self.normal), self.n,
                                         C.c_float(cfg.support), ptr(self.ghost), _lib.stream_handle(dev)),
              "boundary_ghosts")
        self.viscosity = VISCOSITY[cfg.boundary_viscosity]
        self.visc_coeff = cfg.boundary_nu if cfg.boundary_viscosity == "deltaSPH_viscid" else cfg.boundary_alpha
        self.eps = cfg.boundary_eps
        self.pi_switch = 1 if cfg.boundary_pi_switch else 0


@dataclass
class RHS:
    dudt: torch.Tensor
    drhodt: torch.Tensor
    grad_rho: torch.Tensor
    num_pairs: int


class CapturedSteps:
    """`steps` RK4 steps of a :class:`DeltaSPHSolver` recorded in one graph
    (:meth:`DeltaSPHSolver.capture`)."""

    def __init__(self, solver, graph, steps, flag, widths, bufs):
        self.solver = solver
        self.graph = graph
        self.steps = steps
        self.flag = flag
        self.widths = widths
        self.bufs = bufs            # the state buffers the graph reads and writes
        self.fallbacks = 0          # launches redone eagerly after a row overflow

    @property
    def valid(self) -> bool:
        """False once the solver's list widths changed (recapture)."""
        return self.solver._widths == self.widths

    def launch(self):
        """Advance the solver by `steps` steps; one host read at the end."""
        s = self.solver
        if not self.valid:
            raise DsphError("captured steps are stale: the neighbor list widths changed; capture again")
        # eager steps since the capture may have moved the state to other
        # tensors and changed the clock: bring both back to the graph's buffers
        for name, buf in zip(("pos", "vel", "mass", "rho", "uid"), self.bufs):
            cur = getattr(s, name)
            if cur.data_ptr() != buf.data_ptr():
                buf.copy_(cur)
                setattr(s, name, buf)
        s._dt_dev.fill_(float(np.float32(s.dt)))
        s._time_dev.fill_(float(s.time))
        snap = s._snapshot()
        self.graph.replay()
        if int(self.flag.item()):
            # a row did not fit its fixed width: redo the steps with the
            # host-checked path from the state before the replay
            s._restore(snap)
            s.dt = float(snap[5].item())
            s.time = float(snap[6].item())
            for _ in range(self.steps):
                s.step()
            self.fallbacks += 1
            return
        s.dt = float(s._dt_dev.item())
        s.time = float(s._time_dev.item())
        s.step_count += self.steps


class DeltaSPHSolver:
    """Owns the particle state (cell-ordered, on the GPU) and steps it with RK4."""

    def __init__(self, positions, velocities, masses, densities, cfg: DeltaSPHConfig, uid=None, walls=None):
        """`walls`: optional dict(pos, dist, normal, area) of static wall
        particles handled with mDBC."""
        if not positions.is_cuda:
            raise RuntimeError("DeltaSPHSolver needs GPU tensors")
        self.cfg = cfg
        self.n = positions.shape[0]
        dev = positions.device
        f = dict(dtype=torch.float32, device=dev)
        self.domain = _lib.make_domain(cfg.domain_min, cfg.domain_max, cfg.periodic)
        self.pos = positions.detach().to(**f).contiguous().clone()
        self.vel = velocities.detach().to(**f).contiguous().clone()
        self.mass = masses.detach().to(**f).contiguous().clone()
        self.rho = densities.detach().to(**f).contiguous().clone()
        self.uid = (uid if uid is not None else torch.arange(self.n, device=dev, dtype=torch.int32)).to(
            torch.int32).contiguous().clone()
        self.prm = cfg.params()
        self.walls = None
        if walls is not None:
            T = lambda a: a if isinstance(a, torch.Tensor) else torch.as_tensor(a)  # noqa: E731
            self.walls = Walls(T(walls["pos"]).to(dev), T(walls["dist"]).to(dev), T(walls["normal"]).to(dev),
                               T(walls["area"]).to(dev), cfg, self.domain)
        self.dt = cfg.dt if cfg.dt is not None else compute_timestep(cfg, cfg.max_dt, None)
        self.time = 0.0
        self.step_count = 0
        self.dudt = None            # RK4-averaged derivatives of the last step
        self.drhodt = None
        self._last_nl = None
        self.last_sort = None
        self.on_sorted = None       # optional callback(solver) after the per-step sort
        n = self.n
        self._grad = torch.empty(n, 2, **f)
        self._vol = torch.empty(n, **f)
        self._k = [(torch.empty(n, 2, **f), torch.empty(n, 2, **f), torch.empty(n, **f)) for _ in range(4)]
        self._stage = (torch.empty(n, 2, **f), torch.empty(n, 2, **f), torch.empty(n, **f))
        self._widths = {}           # list kind -> row width of the one-pass list
        self._deferred = None       # int32 [1] device flag: searches check overflow later (captured steps)

    @property
    def last_pairs(self) -> int:
        """Fluid pairs of the last right-hand side."""
        return 0 if self._last_nl is None else self._last_nl.pair_count()

    # -- pieces ------------------------------------------------------------
    def _search(self, kind: str, grid: CellGrid, q: torch.Tensor, remap: bool = False):
        """Neighbor list of the wrapped queries `q` against `grid`: the
        one-pass fixed-width list while its rows fit, else the two-pass one
        (and the next search of this kind is wider).  `remap`: entries are
        the reference particles' indices in the grid's input order."""
        w = self._widths.get(kind, max(int(self.cfg.neighbor_width), 0))
        if self._deferred is not None:
            if w < 1:
                raise DsphError("captured steps need one-pass neighbor lists (neighbor_width > 0)")
            return search(grid, q, None, "scatter", qpos_wrapped=True, layout="fixed", width=w,
                          deferred=self._deferred, remap=remap)
        nl = search(grid, q, None, "scatter", qpos_wrapped=True, layout="fixed" if w > 0 else "sell", width=w,
                    remap=remap)
        if w > 0 and nl.width == 0:
            self._widths[kind] = (int(nl.counts.max()) + 16) | 1
        return nl

    def neighbors(self, pos: torch.Tensor):
        """Neighbor list of `pos` with rows in the state's order and entries
        mapped to state indices; returns (NeighborList, pairs descriptor,
        wrapped positions, grid)."""
        cfg = self.cfg
        grid = CellGrid.build(pos, None, self.domain, cfg.support * cfg.cell_scale, h_uniform=cfg.support)
        wpos = wrap_positions(pos, self.domain)
        nl = self._search("fluid", grid, wpos, remap=True)
        pr = nl.pairs()
        pr.rpos = ptr(wpos)
        pr.rh = None
        return nl, pr, wpos, grid

    def wall_pairs(self, wpos, grid, rho):
        """Fluid->wall list (rows fluid, entries wall indices), the ghost->
        fluid list, and the mDBC wall densities at the fluid state `rho`.
        Returns (boundary descriptor, objects to keep alive)."""
        W = self.walls
        st = _lib.stream_handle(wpos.device)
        nl_b = self._search("wall", W.grid, wpos)
        pr_b = nl_b.pairs()
        nl_g = self._search("ghost", grid, W.ghost, remap=True)
        pr_g = nl_g.pairs()
        pr_g.rpos = ptr(wpos)
        pr_g.rh = None
        check(lib().dsph_mdbc(C.byref(pr_g), ptr(self.mass), ptr(rho), ptr(W.pos), C.c_float(self.cfg.rho0),
                              C.c_float(self.cfg.support), MDBC[self.cfg.mdbc], ptr(W.rho), ptr(W.shepard), st),
              "mdbc")
        b = _lib.Boundary()
        b.pairs = C.pointer(pr_b)
        b.mass, b.rho, b.vel = ptr(W.mass), ptr(W.rho), ptr(W.vel)
        b.visc_coeff, b.visc_eps = W.visc_coeff, W.eps
        b.viscosity, b.pi_switch = W.viscosity, W.pi_switch
        return b, (nl_b, pr_b, nl_g, pr_g)

    def rhs(self, pos, vel, rho, dudt, drhodt, terms=None, L=None, eigvals=None) -> RHS:
        """One right-hand side at (pos, vel, rho) into dudt [n,2], drhodt [n]."""
        st = _lib.stream_handle(pos.device)
        nl, pr, wpos, grid = self.neighbors(pos)
        bnd, keep = (None, None)
        if self.walls is not None:
            bnd, keep = self.wall_pairs(wpos, grid, rho)
        bp = C.byref(bnd) if bnd is not None else None
        L_ = lib()
        if self.cfg.diffusion or L is not None or eigvals is not None:
            # V_j = m_j / rho_j once per particle for the normalization pass
            check(L_.dsph_volume(self.n, ptr(self.mass), ptr(rho), ptr(self._vol), st), "volume")
            check(L_.dsph_deltasph_normalization_v(C.byref(pr), bp, ptr(self.mass), ptr(rho), ptr(self._vol),
                                                   ptr(L), ptr(eigvals), ptr(self._grad), st),
                  "deltasph_normalization")
        check(L_.dsph_deltasph_rhs(C.byref(pr), bp, C.byref(self.prm), ptr(vel), ptr(self.mass), ptr(rho),
                                   ptr(self._grad), ptr(dudt), ptr(drhodt), ptr(terms), st), "deltasph_rhs")
        self._keep = (nl, wpos, keep)
        return RHS(dudt, drhodt, self._grad, nl.num_pairs)

    def _sort(self):
        grid = CellGrid.build(self.pos, None, self.domain, self.cfg.support * self.cfg.cell_scale,
                              h_uniform=self.cfg.support)
        si = grid.sort_index
        self.last_sort = si
        st = _lib.stream_handle(self.pos.device)

        def perm(t):
            out = torch.empty_like(t)
            if t.dtype == torch.float32:
                w = 1 if t.dim() == 1 else t.shape[1]
                check(lib().dsph_gather_f32(ptr(t), ptr(si), self.n, w, ptr(out), st), "gather")
                return out
            return t[si.long()]

        self.pos, self.vel, self.mass, self.rho, self.uid = (perm(t) for t in
                                                             (self.pos, self.vel, self.mass, self.rho, self.uid))

    # -- one step ------------------------------------------------------------
    def step(self) -> float:
        """One RK4 step with the current dt; returns the dt used."""
        L_ = lib()
        st = _lib.stream_handle(self.pos.device)
        self._sort()
        if self.on_sorted is not None:
            self.on_sorted(self)
        dt = float(np.float32(self.dt))
        half = float(np.float32(dt) * np.float32(0.5))
        x0, u0, r0 = self.pos, self.vel, self.rho
        xs, us, rs = self._stage
        x, u, r = x0, u0, r0
        for s, c in enumerate([half, half, dt, None]):
            kx, ku, kr = self._k[s]
            kx.copy_(u)                     # dx/dt of this stage = its velocity
            self.rhs(x, u, r, ku, kr)
            self._last_nl = self._keep[0]
            if c is not None:
                check(L_.dsph_rk_update(self.n, ptr(x0), ptr(kx), ptr(xs), ptr(u0), ptr(ku), ptr(us), ptr(r0),
                                        ptr(kr), ptr(rs), C.c_float(c), st), "rk_update")
                x, u, r = xs, us, rs
        stages = _lib.RK4Stages()
        for s in range(4):
            stages.dxdt[s] = ptr(self._k[s][0])
            stages.dudt[s] = ptr(self._k[s][1])
            stages.drhodt[s] = ptr(self._k[s][2])
        if self.dudt is None:
            self.dudt = torch.empty_like(self.vel)
            self.drhodt = torch.empty_like(self.rho)
        check(L_.dsph_rk4_combine(self.n, C.byref(stages), ptr(self.pos), ptr(self.vel), ptr(self.rho),
                                  C.c_float(dt), None, ptr(self.dudt), ptr(self.drhodt), st), "rk4_combine")
        if self.cfg.shifting:
            self.shift()
        self.time += dt
        self.step_count += 1
        if self.cfg.adaptive:
            self.dt = compute_timestep(self.cfg, self.dt, self.max_accel())
        return dt

    # -- steps without host synchronisation (captured into a graph) ---------
    def _device_clock(self):
        """Device copies of the step and the time (created on first use from
        the host values)."""
        if getattr(self, "_dt_dev", None) is None:
            dev = self.pos.device
            self._dt_dev = torch.tensor([np.float32(self.dt)], dtype=torch.float32, device=dev)
            self._time_dev = torch.tensor([self.time], dtype=torch.float64, device=dev)
            self._tp = timestep_params(self.cfg)
        return self._dt_dev, self._time_dev

    def _step_device(self):
        """One RK4 step as step() computes it, with the time step and the time
        read and updated on the device (dsph_rk_update_dev,
        dsph_rk4_combine_dev, dsph_timestep_dev) and the searches' overflow
        flags OR-ed into self._deferred: nothing is read back."""
        L_ = lib()
        st = _lib.stream_handle(self.pos.device)
        dt_dev, time_dev = self._device_clock()
        self._sort()
        x0, u0, r0 = self.pos, self.vel, self.rho
        xs, us, rs = self._stage
        x, u, r = x0, u0, r0
        for k, half in enumerate((1, 1, 0, None)):
            kx, ku, kr = self._k[k]
            kx.copy_(u)
            self.rhs(x, u, r, ku, kr)
            self._last_nl = self._keep[0]
            if half is not None:
                check(L_.dsph_rk_update_dev(self.n, ptr(x0), ptr(kx), ptr(xs), ptr(u0), ptr(ku), ptr(us), ptr(r0),
                                            ptr(kr), ptr(rs), ptr(dt_dev), half, st), "rk_update_dev")
                x, u, r = xs, us, rs
        stages = _lib.RK4Stages()
        for k in range(4):
            stages.dxdt[k] = ptr(self._k[k][0])
            stages.dudt[k] = ptr(self._k[k][1])
            stages.drhodt[k] = ptr(self._k[k][2])
        check(L_.dsph_rk4_combine_dev(self.n, C.byref(stages), ptr(self.pos), ptr(self.vel), ptr(self.rho),
                                      ptr(dt_dev), None, ptr(self.dudt), ptr(self.drhodt), st), "rk4_combine_dev")
        if self.cfg.shifting:
            self.shift()
        ws = WORKSPACE.get(L_.dsph_max_norm_workspace(), self.pos.device)
        check(L_.dsph_timestep_dev(ptr(self.dudt), self.n, C.byref(self._tp),
                                   0 if self.cfg.accel_norm == "global" else 1, ptr(dt_dev), ptr(time_dev),
                                   ptr(ws), ws.numel(), st), "timestep_dev")

    def _state(self):
        return (self.pos, self.vel, self.mass, self.rho, self.uid)

    def _snapshot(self):
        dt_dev, time_dev = self._device_clock()
        return [t.clone() for t in self._state()] + [dt_dev.clone(), time_dev.clone()]

    def _restore(self, snap):
        for dst, src in zip(self._state() + (self._dt_dev, self._time_dev), snap):
            dst.copy_(src)

    def capture(self, steps: int) -> "CapturedSteps":
        """Record `steps` RK4 steps into one HIP graph (torch.cuda.CUDAGraph
        over the library's launches on the current stream).  Inside the graph
        nothing is read back: the time step is formed on the device, and the
        neighbor searches use the one-pass fixed-width lists with their
        overflow flags collected on the device.  :meth:`CapturedSteps.launch`
        replays the graph, reads that flag and the clock once, and when a
        row did not fit redoes the steps with :meth:`step` (which widens the
        lists) from a copy of the state taken before the replay.  The steps
        are bit-identical to :meth:`step`.  The state keeps its buffers (the
        graph writes the final state of its steps back into them)."""
        if steps < 1:
            raise ValueError("steps must be >= 1")
        if self.on_sorted is not None:
            raise DsphError("capture: the per-step sort callback cannot run inside a graph")
        for kind in ("fluid",) + (("wall", "ghost") if self.walls is not None else ()):
            if self._widths.get(kind, int(self.cfg.neighbor_width)) < 1:
                raise DsphError("capture: needs one-pass neighbor lists (neighbor_width > 0)")
        dev = self.pos.device
        if self.dudt is None:
            self.dudt = torch.empty_like(self.vel)
            self.drhodt = torch.empty_like(self.rho)
        self._device_clock()
        self._dt_dev.fill_(float(np.float32(self.dt)))
        self._time_dev.fill_(float(self.time))
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        bufs = self._state()
        # one step outside the graph on a copy of the state: sizes the
        # workspaces the captured launches use (nothing grows inside)
        snap = self._snapshot()
        derivs = (self.dudt.clone(), self.drhodt.clone())
        self._deferred = flag
        try:
            self._step_device()
            torch.cuda.current_stream(dev).synchronize()
            self.pos, self.vel, self.mass, self.rho, self.uid = bufs
            self._restore(snap)
            self.dudt.copy_(derivs[0])
            self.drhodt.copy_(derivs[1])
            graph = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.graph(graph, stream=side):
                flag.zero_()
                for _ in range(steps):
                    self._step_device()
                for dst, src in zip(bufs, self._state()):
                    if dst.data_ptr() != src.data_ptr():
                        dst.copy_(src)
            torch.cuda.current_stream(dev).wait_stream(side)
        finally:
            self._deferred = None
            self.pos, self.vel, self.mass, self.rho, self.uid = bufs
        cs = CapturedSteps(self, graph, steps, flag, self._widths.copy(), bufs)
        # the graph holds raw pointers into the shared scratch buffer: keep
        # that tensor alive even if the buffer is regrown later
        cs.scratch = WORKSPACE.get(0, dev)
        return cs

    def shift(self, shift_out: torch.Tensor | None = None):
        """delta+ particle shifting of the current state (solveShifting,
        v2/modules/shifting.py:880-1018 with scheme 'deltaSPH'): a fresh
        neighbor search, fluid and wall neighbors, positions moved in place;
        `shift_out` [n,2] receives the displacement."""
        cfg = self.cfg
        st = _lib.stream_handle(self.pos.device)
        nl, pr, wpos, grid = self.neighbors(self.pos)
        bnd, keep = None, None
        if self.walls is not None:
            W = self.walls
            nl_b = self._search("wall", W.grid, wpos)
            pr_b = nl_b.pairs()
            bnd = _lib.Boundary()
            bnd.pairs = C.pointer(pr_b)
            bnd.mass, bnd.rho, bnd.vel = ptr(W.mass), ptr(W.rho), ptr(W.vel)
            bnd.visc_coeff, bnd.visc_eps, bnd.viscosity, bnd.pi_switch = 0.0, 0.0, -1, 0
            keep = (nl_b, pr_b)
        prm = _lib.ShiftParams()
        prm.dx, prm.support, prm.kernel_scale = cfg.dx, cfg.support, cfg.kernel_scale
        prm.R, prm.n, prm.cfl, prm.cs = cfg.shift_R, cfg.shift_n, cfg.shift_cfl, cfg.cs
        prm.threshold = cfg.shift_threshold
        prm.mach = -1.0 if cfg.shift_compute_mach else 0.1
        ws = WORKSPACE.get(lib().dsph_shift_workspace(self.n), self.pos.device)
        check(lib().dsph_deltaplus_shift(C.byref(pr), C.byref(bnd) if bnd is not None else None, C.byref(prm),
                                         ptr(self.mass), ptr(self.rho), ptr(self.vel), ptr(self.pos), ptr(shift_out),
                                         ptr(ws), ws.numel(), st), "deltaplus_shift")
        self._keep_shift = (nl, wpos, keep)

    def max_accel(self) -> float:
        """The acceleration of the time-step condition (see accel_norm)."""
        ws = WORKSPACE.get(lib().dsph_max_norm_workspace(), self.vel.device)
        st = _lib.stream_handle(self.vel.device)
        if self.cfg.accel_norm == "global":
            out = C.c_double(0.0)
            check(lib().dsph_sum_squares(ptr(self.dudt), 2 * self.n, C.byref(out), ptr(ws), ws.numel(), st),
                  "sum_squares")
            return float(np.float32(math.sqrt(out.value)))
        out = C.c_float(0.0)
        check(lib().dsph_max_norm2(ptr(self.dudt), self.n, C.byref(out), ptr(ws), ws.numel(), st), "max_norm2")
        return float(out.value)

    def kinetic_energy(self) -> float:
        """sum 0.5 m |u|^2 (runner.py:44-45 uses areas * densities = m)."""
        return float(0.5 * (self.mass.double() * (self.vel.double() ** 2).sum(1)).sum())

    def state_by_uid(self):
        order = torch.argsort(self.uid.long())
        return self.pos[order], self.vel[order], self.rho[order]


def taylor_green_state(n_side: int, device, u_max: float = 1.0, target_neighbors: float = 50.0):
    """Config of the reference's Taylor-Green benchmark: n_side^2 cell-centred
    lattice on [-1,1]^2 (sampling.py:55-74), area dx^2, m = area rho0,
    h = sqrt(n dx^2 / pi) (math.py:96-105), rho = rho0, u = sin(pi x) cos(pi y),
    v = -cos(pi x) sin(pi y), c_s = 10 u_max."""
    from . import sampling
    pos, dx = sampling.sample_regular(n_side)
    area = dx * dx
    h = sampling.volume_to_support(area, target_neighbors)
    vel = sampling.taylor_green(pos, u_max)
    n = pos.shape[0]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device)  # noqa: E731
    cfg = DeltaSPHConfig(support=h, dx=dx, cs=10.0 * u_max)
    return (T(pos), T(vel), T(np.full(n, area * cfg.rho0)), T(np.full(n, cfg.rho0)), cfg)


def analytic_tgv_energy_ratio(t: float, nu: float) -> float:
    """E_k(t)/E_k(0) of the Taylor-Green vortex with wavenumber pi per axis:
    exp(-4 nu pi^2 t)."""
    return math.exp(-4.0 * nu * math.pi ** 2 * t)


def dam_break_state(res: int, device, g: float = 9.81, column_height: float = 2.0):
    """The reference's dam-break benchmark: a 1 x 2 water column in the
    closed 4 x 2 box, dx = 1/res, three wall layers, gravity -g, c_s =
    20 sqrt(2 g H), delta-SPH inviscid viscosity alpha = 0.01 for fluid and
    wall, density diffusion delta = 0.1, RK4 with the CFL time step.  The
    minimum time step is lowered below the acoustic one (the reference's
    default floor of 1e-3 would exceed it).  Wall densities follow the
    reference's C++ mDBC op (clamped at rho0): the column reaches the lid,
    and once the water falls away from it the unclamped extrapolation of
    mDBCDensity drives the lid densities to ~0 and the run diverges.
    Returns (positions, velocities, masses, densities, config, walls)."""
    from . import sampling
    d = sampling.dam_break(res)
    n = d["fluid_pos"].shape[0]
    dx = d["dx"]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device)  # noqa: E731
    cs = 20.0 * math.sqrt(2.0 * g * column_height)
    cfg = DeltaSPHConfig(support=d["h"], dx=dx, cs=cs, domain_min=(-4 * dx, -4 * dx),
                         domain_max=(4.0 + 4 * dx, 2.0 + 4 * dx), periodic=(False, False),
                         viscosity="deltaSPH_inviscid", alpha=0.01, gravity=(0.0, -g), min_dt=1e-7,
                         mdbc="mDBC_op")
    walls = dict(pos=T(d["wall_pos"]), dist=T(d["wall_dist"]), normal=T(d["wall_normal"]),
                 area=torch.tensor(d["area"], dtype=torch.float32, device=device))
    return (T(d["fluid_pos"]), torch.zeros(n, 2, device=device), T(np.full(n, d["area"] * cfg.rho0)),
            T(np.full(n, cfg.rho0)), cfg, walls)
