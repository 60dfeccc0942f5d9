"""Batched ensemble: many independent periodic delta-SPH domains stepped
together on one GPU.

The reference has no batched step; several domains run as a Python loop over
batch ids (v2/neighborhoodFallback/neighborhood.py:121-183), each stepped by
deltaPlus.simulationStep (v2/simulationSchemes/deltaPlus.py:176-422) with
RK4 (v2/modules/integration.py:231-307) and computeTimestep
(v2/modules/timestep.py:8-39).  Here the domains are concatenated into one
particle state and every pass of a step is one kernel launch over all of
them (``dsph_batch_step``); ``capture`` records any number of steps into one
HIP graph that replays without host involvement.  Each domain's trajectory
is bit-identical to :class:`~diffsph_amd.deltasph.DeltaSPHSolver` run on that
domain alone with the same ``cell_scale``.

:func:`autotune` picks the grid cell size and the workgroup size by timing
candidate configurations on the actual ensemble.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr
from .deltasph import DeltaSPHConfig, compute_timestep
from .deltasph import timestep_params as _timestep_params


def timestep_params(cfg: DeltaSPHConfig) -> _lib.TimestepParams:
    if cfg.accel_norm != "global":
        raise ValueError("the batched step implements the reference's global acceleration norm only")
    return _timestep_params(cfg)


class HipGraph:
    """An executable HIP graph of captured batch steps."""

    def __init__(self, handle: int, steps: int, owner):
        self.handle = handle
        self.steps = steps
        self._owner = owner      # keeps the captured buffers alive

    def launch(self, stream=None):
        if not self.handle:
            raise RuntimeError("graph was destroyed")
        s = stream if stream is not None else _lib.stream_handle(self._owner.device)
        check(lib().dsph_graph_launch(C.c_void_p(self.handle), s), "graph_launch")

    def close(self):
        if self.handle:
            check(lib().dsph_graph_destroy(C.c_void_p(self.handle)), "graph_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedDeltaSPH:
    """Concatenated state of `len(boxes)` periodic domains.

    boxes: list of ((xmin, ymin), (xmax, ymax)); counts: particles per
    domain; positions, velocities [n,2], masses, densities [n] on the GPU,
    domain d occupying rows [sum(counts[:d]), sum(counts[:d+1])).  All
    domains share the physics of `cfg` (support, spacing, speed of sound,
    viscosity, ...); each has its own box, grid and time step."""

    def __init__(self, boxes, counts, positions, velocities, masses, densities, cfg: DeltaSPHConfig, *,
                 width: int | None = None, block: int = 256, width_margin: float = 1.3):
        if not positions.is_cuda:
            raise RuntimeError("BatchedDeltaSPH needs GPU tensors")
        self.cfg = cfg
        self.device = positions.device
        self.ndom = len(boxes)
        self.counts = [int(c) for c in counts]
        if len(self.counts) != self.ndom:
            raise ValueError("one count per box")
        self.n = sum(self.counts)
        if positions.shape[0] != self.n:
            raise ValueError("positions do not match the counts")
        self.first = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        f = dict(dtype=torch.float32, device=self.device)
        self.pos = positions.detach().to(**f).contiguous().clone()
        self.vel = velocities.detach().to(**f).contiguous().clone()
        self.mass = masses.detach().to(**f).contiguous().clone()
        self.rho = densities.detach().to(**f).contiguous().clone()
        self.uid = torch.arange(self.n, dtype=torch.int32, device=self.device)
        self.domains = (_lib.Domain * self.ndom)()
        for k, (mn, mx) in enumerate(boxes):
            self.domains[k] = _lib.make_domain(mn, mx, (True, True))
        self._counts = (C.c_int64 * self.ndom)(*self.counts)
        self.prm = cfg.params()
        self.tp = timestep_params(cfg)
        dt0 = cfg.dt if cfg.dt is not None else compute_timestep(cfg, cfg.max_dt, None)
        self.dt = torch.full((self.ndom,), float(np.float32(dt0)), **f)
        self.time = torch.zeros(self.ndom, dtype=torch.float64, device=self.device)
        self.block = int(block)
        self.hsearch = C.c_float(cfg.support * cfg.cell_scale).value
        self.batch = None
        if width is None:
            self._init(8)
            width = self.suggest_width(width_margin)
        self._init(int(width))

    # -- set-up --------------------------------------------------------------
    def _init(self, width: int):
        L = lib()
        nbytes = L.dsph_batch_workspace(self.domains, self._counts, self.ndom, C.c_float(self.hsearch), width)
        if nbytes == 0:
            raise _lib.DsphError("dsph_batch_workspace: invalid batch (boxes, counts or grid)")
        self.batch = None
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        b = _lib.Batch()
        check(L.dsph_batch_init(self.domains, self._counts, self.ndom, C.c_float(self.cfg.support),
                                C.c_float(self.hsearch), width, self.block, ptr(self.ws), nbytes, C.byref(b)),
              "batch_init")
        self.batch = b
        self.width = width

    def probe(self):
        """(longest neighbor row, number of pairs) of the current positions;
        leaves these lists in the workspace."""
        m = C.c_int32(0)
        p = C.c_int64(0)
        check(lib().dsph_batch_probe(C.byref(self.batch), ptr(self.pos), C.byref(m), C.byref(p),
                                     _lib.stream_handle(self.device)), "batch_probe")
        return int(m.value), int(p.value)

    def max_neighbors(self) -> int:
        """Longest neighbor row of the current positions."""
        return self.probe()[0]

    def suggest_width(self, margin: float = 1.3) -> int:
        return int(math.ceil(self.max_neighbors() * margin / 8.0) * 8)

    # -- stepping --------------------------------------------------------------
    def _args(self):
        return (C.byref(self.batch), C.byref(self.prm), C.byref(self.tp), ptr(self.pos), ptr(self.vel),
                ptr(self.mass), ptr(self.rho), ptr(self.uid), ptr(self.dt), ptr(self.time))

    def step(self, steps: int = 1):
        """`steps` steps of every domain, launched eagerly (no host sync)."""
        check(lib().dsph_batch_step(*self._args(), int(steps), _lib.stream_handle(self.device)), "batch_step")

    def capture(self, steps: int) -> HipGraph:
        """Record `steps` steps into one executable HIP graph."""
        h = C.c_void_p(0)
        torch.cuda.synchronize(self.device)
        check(lib().dsph_batch_graph(*self._args(), int(steps), C.byref(h)), "batch_graph")
        return HipGraph(h.value, steps, self)

    def status(self, reset: bool = False):
        """(rows truncated to the width since the last reset, longest row seen)."""
        out = (C.c_int32 * 2)()
        check(lib().dsph_batch_status(C.byref(self.batch), out, 1 if reset else 0, _lib.stream_handle(self.device)),
              "batch_status")
        return int(out[0]), int(out[1])

    def check_rows(self):
        """Raise if a neighbor row exceeded the list width (results invalid)."""
        over, longest = self.status()
        if over:
            raise _lib.DsphError(f"{over} neighbor rows exceeded the list width {self.width} "
                                 f"(longest {longest}); rebuild with a larger width")

    # -- views -------------------------------------------------------------------
    def domain(self, d: int) -> slice:
        return slice(int(self.first[d]), int(self.first[d + 1]))

    def state_by_uid(self):
        order = torch.argsort(self.uid.long())
        return self.pos[order], self.vel[order], self.rho[order]

    def momentum(self) -> torch.Tensor:
        """Per-domain total momentum sum m u (float64) [ndom, 2]."""
        mu = (self.mass.double()[:, None] * self.vel.double())
        idx = torch.repeat_interleave(torch.arange(self.ndom, device=self.device),
                                      torch.tensor(self.counts, device=self.device))
        out = torch.zeros(self.ndom, 2, dtype=torch.float64, device=self.device)
        return out.index_add_(0, idx, mu)


@dataclass
class Ensemble:
    boxes: list
    counts: list
    pos: torch.Tensor
    vel: torch.Tensor
    mass: torch.Tensor
    rho: torch.Tensor
    cfg: DeltaSPHConfig


def make_ensemble(ndom: int = 512, nmin: int = 1000, nmax: int = 65536, seed: int = 0, dx: float = 2.0 / 256,
                  kmax: int = 4, umax: float = 1.0, target_neighbors: float = 50.0, device="cuda",
                  sizes=None) -> Ensemble:
    """`ndom` periodic square boxes with N_d = round(exp(U(ln nmin, ln nmax)))
    target particles (seeded), each a cell-centred lattice of side
    round(sqrt(N_d)) at the common spacing dx (so the box side follows the
    count), with a seeded divergence-free spectral-noise velocity of that
    box's period (max |u| = umax).  Physics of the Taylor-Green set-up:
    h for ~50 neighbors, rho0 = 1, c_s = 10 umax, viscid nu = 0.01,
    delta = 0.1, RK4 with the CFL time step.  `sizes` overrides the drawn
    lattice sides."""
    from . import sampling
    rng = np.random.default_rng(seed)
    if sizes is None:
        target = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), ndom)))
        sizes = [max(int(round(math.sqrt(t))), 1) for t in target]
    area = dx * dx
    h = sampling.volume_to_support(area, target_neighbors)
    cfg = DeltaSPHConfig(support=h, dx=dx, cs=10.0 * umax)
    boxes, counts, P, U = [], [], [], []
    for d, s in enumerate(sizes):
        side = s * dx
        lo, hi = (-0.5 * side, -0.5 * side), (0.5 * side, 0.5 * side)
        pos, _ = sampling.sample_regular(s, s, lo, hi)
        pt = torch.from_numpy(pos).to(device)
        U.append(sampling.spectral_velocity_torch(pt, kmax=kmax, seed=seed * 100003 + d, umax=umax, period=side))
        P.append(pt)
        boxes.append((lo, hi))
        counts.append(s * s)
    pos = torch.cat(P)
    vel = torch.cat(U)
    n = pos.shape[0]
    f = dict(dtype=torch.float32, device=pos.device)
    return Ensemble(boxes, counts, pos, vel, torch.full((n,), area * cfg.rho0, **f), torch.full((n,), cfg.rho0, **f),
                    cfg)


def autotune(ens: Ensemble, cell_scales=(1.0, 1.15, 1.3), blocks=(64, 128, 256), steps: int = 2,
             warmup: int = 1, width: int | None = None, log=None):
    """Time `steps` eager steps (after `warmup`) of the ensemble for every
    (cell_scale, block) candidate, each from the same initial state, and
    return (best (cell_scale, block), {candidate: ms per step}).

    The cell scale changes the order in which neighbor entries are summed
    (results differ by float32 rounding, not in the neighbor sets); the
    workgroup size changes nothing in the results."""
    times = {}
    for cs in cell_scales:
        cfg = replace(ens.cfg, cell_scale=float(cs))
        for blk in blocks:
            b = BatchedDeltaSPH(ens.boxes, ens.counts, ens.pos, ens.vel, ens.mass, ens.rho, cfg, width=width,
                                block=blk)
            width = b.width if width is None else width
            b.step(warmup)
            torch.cuda.synchronize(b.device)
            t0 = time.perf_counter()
            b.step(steps)
            torch.cuda.synchronize(b.device)
            ms = (time.perf_counter() - t0) * 1e3 / steps
            b.check_rows()
            times[(float(cs), int(blk))] = ms
            if log:
                log(f"autotune cell_scale {cs:.2f} block {blk:3d}: {ms:.3f} ms/step")
            del b
    best = min(times, key=times.get)
    return best, times
