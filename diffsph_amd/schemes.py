"""The reference's scheme-level Python interface over the fused HIP path.

Drop-ins with the reference's names, arguments, state-dict keys and config
keys:

* ``callModule(state, module, config, direction, reduce)``
  (v2/simulationSchemes/deltaPlus.py:30-96): pure dispatch of a module over
  the fluid / boundary sets, kept as written;
* ``computeDensity`` (v2/modules/density.py:6-15) and ``computeGravity``
  (v2/modules/gravity.py:8-29) on state dicts;
* ``dfsphSimulationStep(state, config)`` = the DFSPH scheme's
  ``simulationStep`` (v2/simulationSchemes/dfsph.py:148-227) and
  ``deltaPlusSimulationStep(state, config)`` = the delta-SPH scheme's
  (v2/simulationSchemes/deltaPlus.py:176-422).  Both return the reference's
  ``((dxdt, dudt, drhodt), (boundary...))`` tuple that ``integrate``
  consumes and write the same intermediate fields into ``state['fluid']``.

The scheme steps run the fused kernels (neighbor search, density,
alpha/source/viscosity sweep, the device-controlled pressure solves; the
delta-SPH normalization and right-hand-side kernels) instead of composing
per-pair tensors.  ``state['fluid']['neighborhood']`` is then a
:class:`LazyNeighborDict`: the per-pair arrays of the reference's
neighborDict are built (by :func:`neighborhood.neighborSearch`) only if a
caller reads them.

Scope: fluid-only states (``config['boundary']['active']`` False) with a
uniform support; the DFSPH solver has no wall model (as the reference's
DFSPH scheme solves the fluid only) and the delta-SPH walls are driven by
:class:`deltasph.DeltaSPHSolver` directly.
"""
from __future__ import annotations

import torch

from ._lib import DsphError
from .sphops import sphOperationStates


def callModule(state, module, config, direction="all", reduce="sum", **kwargs):
    """v2/simulationSchemes/deltaPlus.py:30-96."""
    def red(a, b):
        if reduce == "sum":
            return a + b
        if reduce == "min":
            return torch.min(a, b)
        if reduce == "max":
            return torch.max(a, b)
        raise ValueError(f"Invalid reduce {reduce}")

    f, b = state["fluid"], state.get("boundary")
    if not config["boundary"]["active"]:
        return module(f, f, f["neighborhood"], config, **kwargs), None
    if direction == "all":
        ftf = module(f, f, f["neighborhood"], config, **kwargs)
        btb = module(b, b, b["neighborhood"], config, **kwargs)
        btf = module(f, b, state["boundaryToFluidNeighborhood"], config, **kwargs)
        ftb = module(b, f, state["fluidToBoundaryNeighborhood"], config, **kwargs)
        return red(ftf, btf), red(btb, ftb)
    if direction == "fluid":
        return module(f, f, f["neighborhood"], config, **kwargs), None
    if direction in ("fluidwBoundary", "oneWay"):
        ftf = module(f, f, f["neighborhood"], config, **kwargs)
        btf = module(f, b, state["boundaryToFluidNeighborhood"], config, **kwargs)
        return red(ftf, btf), None
    if direction == "boundary":
        return None, module(b, b, b["neighborhood"], config, **kwargs)
    if direction == "boundarywFluid":
        btb = module(b, b, b["neighborhood"], config, **kwargs)
        ftb = module(b, f, state["fluidToBoundaryNeighborhood"], config, **kwargs)
        return None, red(btb, ftb)
    if direction == "fluidToBoundary":
        return None, module(b, f, state["fluidToBoundaryNeighborhood"], config, **kwargs)
    if direction == "boundaryToFluid":
        return module(f, b, state["boundaryToFluidNeighborhood"], config, **kwargs), None
    raise ValueError(f"Invalid direction {direction}")


def computeDensity(stateA, stateB, neighborhood, simConfig):
    """v2/modules/density.py:6-15."""
    return sphOperationStates(stateA, stateB, quantities=None, operation="density", neighborhood=neighborhood)


def computeGravity(fluidState, config):
    """v2/modules/gravity.py:8-29 (constant field or potential field)."""
    g = config.get("gravity", {})
    v = fluidState["velocities"]
    if not g.get("active", False):
        return torch.zeros_like(v)
    if g.get("gravityMode", "constant") == "potentialField":
        x = fluidState["positions"]
        d = config["domain"]
        mn, mx, per = d["minExtent"], d["maxExtent"], d["periodicity"]
        center = torch.tensor(g["potentialOrigin"], dtype=x.dtype, device=x.device)[:x.shape[-1]]
        cols = []
        for k in range(x.shape[1]):
            c = x[:, k] - center[k]
            if bool(per[k]):
                lo, hi = float(mn[k]), float(mx[k])
                c = torch.where(torch.abs(c) > (hi - lo) / 2, torch.sgn(c) * ((torch.abs(c) + lo) % (hi - lo) + lo), c)
            cols.append(c)
        xij = torch.stack(cols, -1)
        rij = torch.linalg.norm(xij, dim=-1)
        m = rij > 1e-7
        xij[m] = xij[m] / rij[m, None]
        return -g["magnitude"] ** 2 * xij * rij[:, None]
    direction = torch.tensor(g["direction"], dtype=fluidState["positions"].dtype, device=v.device)
    return (direction[:v.shape[1]] * g["magnitude"]).repeat(v.shape[0], 1)


class LazyNeighborDict(dict):
    """A neighborDict whose per-pair arrays are built on first access (by the
    reference-compatible :func:`neighborhood.neighborSearch` at the state's
    positions); ``numNeighbors`` of the fused search is available eagerly."""

    def __init__(self, materialize, **eager):
        super().__init__(**eager)
        self._materialize = materialize
        self._done = False

    def _fill(self):
        if not self._done:
            self._done = True
            full = self._materialize()
            for k, v in full.items():
                if not dict.__contains__(self, k):
                    dict.__setitem__(self, k, v)

    def __getitem__(self, key):
        if not dict.__contains__(self, key):
            self._fill()
        return dict.__getitem__(self, key)

    def __contains__(self, key):
        if not dict.__contains__(self, key):
            self._fill()
        return dict.__contains__(self, key)

    def get(self, key, default=None):
        return self[key] if key in self else default

    def keys(self):
        self._fill()
        return dict.keys(self)

    def items(self):
        self._fill()
        return dict.items(self)

    def values(self):
        self._fill()
        return dict.values(self)


def _domain(config):
    d = config["domain"]
    mn = [float(v) for v in (d["minExtent"].tolist() if isinstance(d["minExtent"], torch.Tensor) else d["minExtent"])]
    mx = [float(v) for v in (d["maxExtent"].tolist() if isinstance(d["maxExtent"], torch.Tensor) else d["maxExtent"])]
    pr = d.get("periodicity", d.get("periodic"))
    pr = pr.tolist() if isinstance(pr, torch.Tensor) else pr
    if not isinstance(pr, (list, tuple)):
        pr = [pr, pr]
    return tuple(mn[:2]), tuple(mx[:2]), tuple(bool(v) for v in pr[:2])


def _uniform_support(fluid):
    h = fluid["supports"]
    hmin, hmax = float(h.min()), float(h.max())
    if hmin != hmax:
        raise DsphError("the fused scheme step needs a uniform support (supports differ: "
                        f"{hmin:g} .. {hmax:g})")
    return hmin


def _lazy_neighborhood(fluid, config, counts=None):
    """The neighborDict of the positions the step ran at (snapshotted now,
    so a later integration of the state does not change it)."""
    from .neighborhood import neighborSearch
    snap = {"positions": fluid["positions"].detach().clone(), "supports": fluid["supports"].detach().clone(),
            "numParticles": fluid["numParticles"]}

    def mat():
        ds, nd = neighborSearch(snap, snap, config, computeKernels=True)
        fluid["datastructure"] = ds
        return nd
    eager = {} if counts is None else {"numNeighbors": counts}
    return LazyNeighborDict(mat, **eager)


def _to_caller(t, uid):
    out = torch.empty_like(t)
    out.index_copy_(0, uid.long(), t)
    return out


def _from_caller(t, uid):
    return None if t is None else t.detach().to(torch.float32).contiguous()[uid.long()]


def dfsphSimulationStep(state, config):
    """The DFSPH scheme's ``simulationStep(state, config)``
    (v2/simulationSchemes/dfsph.py:148-227): density, viscosity, gravity,
    divergence solve (omega 0.3, 1e-2 / rho0), density solve (omega 0.3,
    5e-4 / rho0), both warm-started with 0.75 of the previous step's
    ``pressureDivergence`` / ``pressureIncompressible``; returns
    ``((v, dudt, None), (None, None, None))``."""
    from .dfsph import DFSPHConfig, DFSPHSolver
    if config["boundary"]["active"]:
        raise DsphError("dfsphSimulationStep: fluid-only states (the DFSPH solver has no wall model)")
    f = state["fluid"]
    h = _uniform_support(f)
    mn, mx, per = _domain(config)
    rho0 = float(config["fluid"]["rho0"])
    diff = config.get("diffusion", {})
    scheme = diff.get("velocityScheme", "deltaSPH_inviscid")
    if scheme not in ("deltaSPH_inviscid", "deltaSPH_viscid"):
        raise DsphError(f"dfsphSimulationStep: velocityScheme {scheme} is not built (deltaSPH_inviscid/viscid)")
    grav = computeGravity(f, config)
    gtuple = None
    if config.get("gravity", {}).get("active", False):
        if config["gravity"].get("gravityMode", "constant") != "constant":
            raise DsphError("dfsphSimulationStep: only a constant gravity field is fused")
        gtuple = tuple(float(v) for v in grav[0].tolist()) if grav.shape[0] else (0.0, 0.0)
    dfc = config["dfsph"]
    cfg = DFSPHConfig(dt=float(config["timestep"]["dt"]), rho0=rho0, support=h, domain_min=mn, domain_max=mx,
                      periodic=per, omega_div=0.3, omega_den=0.3, threshold_div=1e-2, threshold_den=5e-4,
                      min_iters=int(dfc.get("minIters", 2)), max_iters=int(dfc.get("maxIters", 256)),
                      clamp_pressure=bool(dfc.get("clampPressure", True)), warm_start=0.75, viscosity=scheme,
                      visc_alpha=float(diff.get("alpha", 0.01)), visc_nu=float(diff.get("nu", 0.01)),
                      cs=float(config["fluid"]["cs"]), visc_eps=float(diff.get("eps", 1e-6)),
                      pi_switch=bool(diff.get("pi-switch", False)), gravity=gtuple)
    s = DFSPHSolver(f["positions"], f["velocities"], f["masses"], f["areas"], cfg)
    if "pressureDivergence" in f:
        s.p_div = f["pressureDivergence"].detach().to(torch.float32).contiguous().clone()
    if "pressureIncompressible" in f:
        s.p_den = f["pressureIncompressible"].detach().to(torch.float32).contiguous().clone()
    stats = s.derivatives()
    uid = s.uid
    back = lambda t: _to_caller(t, uid)  # noqa: E731
    f["densities"] = back(s.rho)
    f["velocityDiffusion"] = back(s.visc)
    f["gravityAccel"] = grav
    f["divergenceAccel"] = back(s.a_div)
    f["pressureDivergence"] = back(s.p_div)
    f["advection"] = back(s.adv)
    f["pressureAccel"] = back(s.a_den)
    f["pressureIncompressible"] = back(s.p_den)
    f["pressureB"] = f["pressureIncompressible"]
    # the fused solves keep their loop on the device: the convergence lists
    # hold the final value of the metric (the iteration counts beside them)
    f["convergence_divergence"] = [stats.div_error]
    f["convergence_density"] = [stats.den_error]
    f["convergence_iterations"] = {"divergence": stats.div_iters, "density": stats.den_iters}
    f["numNeighbors"] = back(s.nl.counts)
    f["neighborhood"] = _lazy_neighborhood(f, config, f["numNeighbors"])
    dfc["sourceTerm"], dfc["errorThreshold"], dfc["omega"] = "density", 5e-4 / rho0, 0.3
    dudt = back(s.dudt)
    return (f["velocities"].clone(), dudt, None), (None, None, None)


def deltaPlusSimulationStep(state, config):
    """The delta-SPH scheme's ``simulationStep(state, config)``
    (v2/simulationSchemes/deltaPlus.py:176-422) for a fluid-only state with
    the continuity density: normalization matrices and L, the renormalized
    density gradient, density diffusion, viscosity, continuity, EOS,
    pressure acceleration and gravity; returns ``((v, dudt, drhodt),
    (None, None, None))``."""
    from .deltasph import DeltaSPHConfig, DeltaSPHSolver
    if config["boundary"]["active"]:
        raise DsphError("deltaPlusSimulationStep: fluid-only states (walls: deltasph.DeltaSPHSolver(walls=...))")
    if config.get("simulation", {}).get("densityScheme", "continuity") == "summation":
        raise DsphError("deltaPlusSimulationStep: the fused step integrates the continuity density")
    f = state["fluid"]
    h = _uniform_support(f)
    mn, mx, per = _domain(config)
    diff = config.get("diffusion", {})
    scheme = diff.get("velocityScheme", "deltaSPH_viscid")
    grav = computeGravity(f, config)
    gtuple = None
    if config.get("gravity", {}).get("active", False):
        if config["gravity"].get("gravityMode", "constant") != "constant":
            raise DsphError("deltaPlusSimulationStep: only a constant gravity field is fused")
        gtuple = tuple(float(v) for v in grav[0].tolist()) if grav.shape[0] else (0.0, 0.0)
    ks = config["kernel"]["kernelScale"]
    ks = float(ks) if not isinstance(ks, torch.Tensor) else float(ks.item())
    dx = float(config["particle"]["dx"]) if "dx" in config.get("particle", {}) else h / (2.0 * ks)
    eos = config.get("EOS", {}).get("type", "isoThermal")
    cfg = DeltaSPHConfig(support=h, dx=dx, rho0=float(config["fluid"]["rho0"]), cs=float(config["fluid"]["cs"]),
                         domain_min=mn, domain_max=mx, periodic=per, delta=float(diff.get("delta", 0.1)),
                         diffusion=diff.get("densityScheme", "deltaSPH") == "deltaSPH", viscosity=scheme,
                         nu=float(diff.get("nu", 0.01)), alpha=float(diff.get("alpha", 0.01)),
                         eps=float(diff.get("eps", 1e-6)), pi_switch=bool(diff.get("pi-switch", False)), eos=eos,
                         kappa=float(config.get("EOS", {}).get("kappa", 1.3)), gravity=gtuple, dt=1e-3)
    pos = f["positions"].detach().to(torch.float32).contiguous()
    vel = f["velocities"].detach().to(torch.float32).contiguous()
    rho = f["densities"].detach().to(torch.float32).contiguous()
    s = DeltaSPHSolver(pos, vel, f["masses"], rho, cfg)
    n = pos.shape[0]
    dev = pos.device
    dudt = torch.empty(n, 2, device=dev)
    drhodt = torch.empty(n, device=dev)
    terms = torch.empty(n, 8, device=dev)  # [n,8]: a_p.xy, visc.xy, continuity, diffusion, p, 0
    L = torch.empty(n, 2, 2, device=dev)
    ev = torch.empty(n, 2, device=dev)
    r = s.rhs(pos, vel, rho, dudt, drhodt, terms=terms, L=L, eigvals=ev)
    f["L"], f["L.EVs"] = L, ev
    f["gradRho^L"] = r.grad_rho.clone()
    f["densityDiffusion"] = terms[:, 5].contiguous()
    f["velocityDiffusion"] = terms[:, 2:4].contiguous()
    f["momentumEquation"] = terms[:, 4].contiguous()
    f["pressures"] = terms[:, 6].contiguous()
    f["pressureAccel"] = terms[:, 0:2].contiguous()
    f["gravityAccel"] = grav
    f["numNeighbors"] = s._keep[0].counts.clone()
    f["neighborhood"] = _lazy_neighborhood(f, config, f["numNeighbors"])
    return (f["velocities"].clone(), dudt, drhodt), (None, None, None)
