"""Whole single-domain delta-SPH steps captured into one HIP graph
(DeltaSPHSolver.capture): the time step formed on the device, the neighbor
lists' overflow flags collected on the device, nothing read back inside the
graph.  Every replay is compared bit for bit with the same number of eager
steps (DeltaSPHSolver.step, which reads the lists' flags and the time-step
norm back on the host), with eager work and synchronisations between the
replays, as the reference's runner interleaves its own work between steps
(runner.py:25-92)."""
from dataclasses import replace

import numpy as np
import pytest
import torch

from diffsph_amd.deltasph import DeltaSPHSolver, dam_break_state, taylor_green_state

pytestmark = pytest.mark.gpu


def _pair(case):
    dev = torch.device("cuda:0")
    if case.startswith("dam_break"):
        pos, vel, mass, rho, cfg, walls = dam_break_state(32, dev)
        if case == "dam_break_shifting":
            cfg = replace(cfg, shifting=True)
    else:
        pos, vel, mass, rho, cfg = taylor_green_state(48, dev)
        cfg = replace(cfg, accel_norm="max")
        walls = None
    return [DeltaSPHSolver(pos, vel, mass, rho, cfg, walls=walls) for _ in range(2)]


def _same(a, b, what):
    for name in ("pos", "vel", "rho", "mass", "uid"):
        ta, tb = getattr(a, name), getattr(b, name)
        assert torch.equal(ta, tb), f"{what}: {name} differs (max {float((ta.double() - tb.double()).abs().max())})"
    assert a.dt == b.dt and a.time == b.time and a.step_count == b.step_count, what


@pytest.mark.parametrize("case", ["dam_break_walls", "dam_break_shifting", "tgv_max_norm"])
def test_captured_steps_match_eager_steps(case):
    a, b = _pair(case)
    print(f"case {case}: {a.n} particles", flush=True)
    k = 5
    a.step()
    b.step()
    cs = b.capture(k)
    _same(a, b, "after capture (the capture leaves the state as it was)")
    for launch in range(3):
        for _ in range(k):
            a.step()
        cs.launch()
        _same(a, b, f"launch {launch}")
        # eager work between replays: reads, a synchronisation
        assert np.isfinite(b.kinetic_energy())
        torch.cuda.synchronize()
    assert cs.fallbacks == 0 and cs.valid
    # an eager step moves the state to new tensors; the next launch copies
    # it back into the graph's buffers
    a.step()
    b.step()
    for _ in range(k):
        a.step()
    cs.launch()
    _same(a, b, "launch after an eager step")
    print(f"dt {b.dt:.6g}, t {b.time:.6g}, pairs {b.last_pairs}", flush=True)


def test_row_overflow_falls_back_to_the_checked_steps():
    """Lists too narrow for the state: the replay flags the overflow, the
    launch redoes its steps with the host-checked path from the saved state
    (widening the lists), and the result still equals the eager steps."""
    a, b = _pair("dam_break_walls")
    a.step()
    b.step()
    b._widths = {"fluid": 16, "wall": 4, "ghost": 16}
    cs = b.capture(4)
    for _ in range(4):
        a.step()
    cs.launch()
    _same(a, b, "after the fallback")
    assert cs.fallbacks == 1 and not cs.valid
    assert b._widths["fluid"] > 16
    with pytest.raises(Exception, match="stale"):
        cs.launch()
    cs2 = b.capture(4)
    for _ in range(4):
        a.step()
    cs2.launch()
    _same(a, b, "recaptured")
    assert cs2.fallbacks == 0
