"""Simulation driver over the fused solvers (v2/runner.py:25-92).

The reference's ``runSimulation`` loops integrate -> solveShifting ->
kinetic energy -> ``computeStatistics`` -> ``writeFrame`` every
``export.interval`` steps -> time / time-step update, with host reads of
several statistics after every step.  :func:`runSimulation` here drives a
:class:`deltasph.DeltaSPHSolver` (RK4 with the CFL time step and the
optional delta+ shifting inside its step) or a :class:`dfsph.DFSPHSolver`:

* delta-SPH steps are replayed ``capture`` at a time from one HIP graph
  (:meth:`DeltaSPHSolver.capture`, one host read per launch, recaptured when
  a list widens), and run eagerly up to the next frame when fewer remain;
* the frame statistics (the keys of the reference's ``computeStatistics``,
  v2/util.py:65-139, plus ``E_k``) are formed on the device and read in one
  transfer, every ``statsInterval`` steps and at every exported frame;
* frames go through :func:`export.writeFrame` (asynchronous writer) in the
  reference's io.py layout.

Plotting (``setupInitialPlot``, ``updatePlots``) is not part of this
package.
"""
from __future__ import annotations

import math

import torch

from .export import createOutputFile, writeFrame


def _counts(solver) -> torch.Tensor | None:
    nl = getattr(solver, "nl", None)
    if nl is None:
        nl = getattr(solver, "_last_nl", None)
    return None if nl is None else nl.counts


def _time(solver) -> float:
    t = getattr(solver, "time", None)
    return float(t) if t is not None else solver.step_count * float(solver.cfg.dt)


def _dt(solver) -> float:
    dt = getattr(solver, "dt", None)
    return float(dt) if dt is not None else float(solver.cfg.dt)


def computeStatistics(solver, config: dict | None = None) -> dict:
    """The frame statistics of v2/util.py:65-139 for the solver's current
    state: compression extremes and mean, neighbor-count extremes and
    median (of the last list), the viscous / acoustic time-step bounds and
    the CFL number (from `config` when given, in the reference's keys),
    time, step and dt, and E_k = sum 0.5 m |u|^2 (runner.py:44-45).  One
    device-to-host transfer."""
    rho0 = float(solver.cfg.rho0)
    c = solver.rho / rho0 - 1.0
    v = [c.max(), c.min(), solver.rho.mean() / rho0,
         0.5 * (solver.mass.double() * (solver.vel.double() ** 2).sum(1)).sum()]
    cnt = _counts(solver)
    if cnt is not None:
        v += [cnt.min(), cnt.max(), cnt.median()]
    host = [float(x) for x in torch.stack([x.double() for x in v]).cpu()]
    st = {"maxCompression": host[0], "minCompression": host[1], "averageDensity": host[2],
          "averageCompression": host[2] - 1.0, "maxShift": 0.0, "E_k": host[3]}
    shift = getattr(solver, "last_shift", None)
    if shift is not None and config is not None:
        st["maxShift"] = float(torch.linalg.norm(shift, dim=-1).max()) / float(config["particle"]["dx"])
    if cnt is not None:
        st.update(minNeighborCount=int(host[4]), maxNeighborCount=int(host[5]), medianNeighborCount=int(host[6]))
    if config is not None:
        h = float(config["particle"]["support"])
        ks = float(config["kernel"]["kernelScale"])
        cs = float(config["fluid"]["cs"])
        nu = config.get("diffusion", {}).get("nu_sph")
        if nu:
            st["dt_v"] = 0.125 * h * h / float(nu) / ks ** 2
        st["dt_c"] = float(config["timestep"]["CFL"]) * h / cs / ks
        st["CFLNumber"] = _dt(solver) / (h / cs / ks)
    st.update(time=_time(solver), timestep=int(solver.step_count), dt=_dt(solver))
    return st


def _frame_states(solver):
    """(perennialState, priorState) in the keys writeFrame reads; particles
    in the solver's (cell) order, with their UIDs."""
    fl = {"positions": solver.pos, "velocities": solver.vel, "densities": solver.rho}
    shift = getattr(solver, "last_shift", None)
    per = {"timestep": int(solver.step_count),
           "fluid": {"shiftAmount": shift if shift is not None else torch.zeros_like(solver.pos),
                     "index": solver.uid}}
    return per, {"fluid": fl}


def runSimulation(solver, stepLimit: int = 1000, timeLimit: float = -1, capture: int = 0, config: dict | None = None,
                  statsInterval: int = 1, callBack=None):
    """Advance `solver` until `stepLimit` steps (> 0) or simulated time
    `timeLimit` (> 0) is passed (runner.py:25-92).  `capture` > 0 replays
    delta-SPH steps from a graph of that many steps.  With `config` (the
    reference's config dict) and ``config['export']['active']``, frames are
    written every ``config['export']['interval']`` steps to
    ``config['export']['exportPath']``.  Statistics are gathered every
    `statsInterval` steps (0: only at frames) and passed to
    ``callBack(solver, stats)``.  Returns (list of statistics, solver)."""
    if stepLimit <= 0 and timeLimit <= 0:
        raise ValueError("runSimulation needs a step or time limit")
    export = config is not None and config.get("export", {}).get("active", False)
    interval = int(config["export"]["interval"]) if export else 0
    f, group = createOutputFile(config) if export else (None, None)
    stats = [computeStatistics(solver, config)]
    cap = None
    try:
        while True:
            if stepLimit > 0 and solver.step_count >= stepLimit:
                break
            if timeLimit > 0 and _time(solver) > timeLimit:
                break
            # steps until the next frame, statistics point or the step limit
            left = [stepLimit - solver.step_count] if stepLimit > 0 else []
            for every in (interval, statsInterval):
                if every > 0:
                    left.append(every - solver.step_count % every)
            ahead = min(left) if left else math.inf
            if capture > 0 and ahead >= capture and timeLimit <= 0 and hasattr(solver, "capture"):
                if cap is None or not cap.valid:
                    cap = solver.capture(capture)
                cap.launch()
            else:
                solver.step()
            k = solver.step_count
            frame = export and k % interval == 0
            if frame or (statsInterval > 0 and k % statsInterval == 0):
                st = computeStatistics(solver, config)
                stats.append(st)
                if frame:
                    per, prior = _frame_states(solver)
                    writeFrame(group, per, prior, st, config)
                if callBack is not None:
                    callBack(solver, st)
    finally:
        if f is not None:
            f.close()
    return stats, solver


def setupSimulation(solver, config: dict | None = None):
    """The statistics of the initial state (runner.py:11-23 without the
    plots): the first entry runSimulation records."""
    return computeStatistics(solver, config)


__all__ = ["computeStatistics", "runSimulation", "setupSimulation"]
