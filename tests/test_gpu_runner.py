"""The simulation driver (runSimulation, v2/runner.py:25-92) over the fused
solvers: delta-SPH steps replayed from captured graphs between frames must
leave the state bit-identical to the same number of eager steps; frames are
written every export interval with the state of that step; the statistics
carry the reference's keys (v2/util.py:122-139) with values computed here
from the same state.  A DFSPH solver runs through the same loop eagerly."""
import math
import os

import numpy as np
import pytest
import torch

from diffsph_amd import sampling
from diffsph_amd.deltasph import DeltaSPHSolver, dam_break_state
from diffsph_amd.dfsph import DFSPHConfig, DFSPHSolver
from diffsph_amd.export import readFrame
from diffsph_amd.runner import computeStatistics, runSimulation

pytestmark = pytest.mark.gpu

KEYS = ["maxCompression", "minCompression", "averageDensity", "averageCompression", "maxShift", "dt_c",
        "CFLNumber", "minNeighborCount", "maxNeighborCount", "medianNeighborCount", "time", "timestep", "dt"]


def _config(tmp, cfg, interval):
    return {"plot": {"namingScheme": "runner"},
            "export": {"active": True, "interval": interval, "exportPath": str(tmp)},
            "particle": {"support": cfg.support, "dx": cfg.dx},
            "kernel": {"targetNeighbors": 50.0, "kernelScale": cfg.kernel_scale},
            "fluid": {"rho0": cfg.rho0, "cs": cfg.cs}, "timestep": {"CFL": cfg.cfl},
            "simulation": {"timestamp": "x"}}


def test_runner_captured_steps_and_frames(tmp_path):
    dev = torch.device("cuda:0")
    pos, vel, mass, rho, cfg, walls = dam_break_state(32, dev)
    a = DeltaSPHSolver(pos, vel, mass, rho, cfg, walls=walls)
    b = DeltaSPHSolver(pos, vel, mass, rho, cfg, walls=walls)
    print(f"dam break {a.n} particles", flush=True)
    conf = _config(tmp_path, cfg, 10)
    seen = []
    stats, _ = runSimulation(b, stepLimit=30, capture=5, config=conf, statsInterval=0,
                             callBack=lambda s, st: seen.append(st["timestep"]))
    # frames at steps 10, 20, 30 only; statistics of those frames + the start
    assert seen == [10, 20, 30]
    assert [s["timestep"] for s in stats] == [0, 10, 20, 30]
    frames = {}
    for _ in range(30):
        a.step()
        if a.step_count % 10 == 0:
            frames[a.step_count] = (a.pos.cpu().numpy().copy(), a.vel.cpu().numpy().copy(),
                                    a.rho.cpu().numpy().copy(), a.uid.cpu().numpy().copy(), a.time)
    for name in ("pos", "vel", "rho", "uid"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    assert a.time == b.time and a.dt == b.dt
    root = os.path.join(tmp_path, "runner.hdf5.d")
    for k, (p, v, r, u, t) in frames.items():
        fr = readFrame(root, k)
        np.testing.assert_array_equal(fr["fluidPosition"], p)
        np.testing.assert_array_equal(fr["fluidVelocity"], v)
        np.testing.assert_array_equal(fr["fluidDensity"], r / np.float32(cfg.rho0))
        np.testing.assert_array_equal(fr["UID"], u)
        assert fr["attrs"]["time"] == pytest.approx(t, rel=0, abs=0)
    st = stats[-1]
    for key in KEYS:
        assert key in st, key
    c = b.rho.double() / cfg.rho0 - 1.0
    assert st["maxCompression"] == float(c.max()) and st["minCompression"] == float(c.min())
    nn = b._last_nl.counts
    assert st["minNeighborCount"] == int(nn.min()) and st["maxNeighborCount"] == int(nn.max())
    assert math.isclose(st["CFLNumber"], b.dt / (cfg.support / cfg.cs / cfg.kernel_scale), rel_tol=1e-12)
    assert math.isclose(st["E_k"], b.kinetic_energy(), rel_tol=1e-12)


def test_runner_dfsph_eager():
    dev = torch.device("cuda:0")
    d = sampling.dfsph_random_flow(64, seed=0, jitter_width=0.1)
    dt = 3.0 * d["dx"] / 10.0
    T = lambda x: torch.from_numpy(x).to(dev)  # noqa: E731
    mk = lambda: DFSPHSolver(T(d["pos"]), T(d["vel"]), T(d["mass"]), T(d["area"]),  # noqa: E731
                             DFSPHConfig(dt=dt, support=d["h"]))
    a, b = mk(), mk()
    stats, _ = runSimulation(b, stepLimit=4, statsInterval=2)
    for _ in range(4):
        a.step()
    for name in ("pos", "vel", "rho", "uid"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    assert [s["timestep"] for s in stats] == [0, 2, 4]
    assert stats[-1]["time"] == pytest.approx(4 * dt)
    assert stats[-1] == computeStatistics(b)
