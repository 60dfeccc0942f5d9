"""Batched ensemble of periodic delta-SPH domains on the GPU.

Parity: every domain of a batch follows, bit for bit, the single-domain
DeltaSPHSolver run on that domain alone (whose steps are checked bit-exact
against the oracle in test_gpu_deltasph.py), including its own adaptive
time step; captured HIP graphs replay the eager steps bit for bit.  At the
full ensemble size (512 domains, ~8.8M particles) the checks are
size-independent: per-domain momentum conservation, no truncated neighbor
rows, and bit-exact parity of the smallest and largest domains."""
from dataclasses import replace

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DX = 2.0 / 256


def _batched():
    from diffsph_amd import batched
    return batched


def _single_run(ens, d, cfg, steps):
    from diffsph_amd.deltasph import DeltaSPHSolver
    b0 = int(np.sum(ens.counts[:d]))
    sl = slice(b0, b0 + ens.counts[d])
    lo, hi = ens.boxes[d]
    c = replace(cfg, domain_min=lo, domain_max=hi, periodic=(True, True))
    s = DeltaSPHSolver(ens.pos[sl], ens.vel[sl], ens.mass[sl], ens.rho[sl], c)
    for _ in range(steps):
        s.step()
    return s


def _assert_domain_equal(b, s, d, tag=""):
    pos, vel, rho = b.state_by_uid()
    sl = b.domain(d)
    p1, v1, r1 = s.state_by_uid()
    for name, x, y in (("pos", pos[sl], p1), ("vel", vel[sl], v1), ("rho", rho[sl], r1)):
        diff = (x - y).abs().max().item()
        assert torch.equal(x, y), f"{tag} domain {d} {name} differs by {diff:.3e}"
    assert float(b.dt[d].item()) == float(np.float32(s.dt)), f"{tag} domain {d} dt"
    assert float(b.time[d].item()) == pytest.approx(s.time, rel=0, abs=0)


@pytest.mark.parametrize("cell_scale,block", [(1.0, 256), (1.3, 64), (1.15, 128)])
def test_batched_domains_match_single_domain_bitexact(cell_scale, block):
    B = _batched()
    ens = B.make_ensemble(sizes=[16, 23, 30, 37, 45], dx=DX, seed=7, device="cuda")
    cfg = replace(ens.cfg, cell_scale=cell_scale)
    steps = 4
    print(f"batched parity: cell_scale {cell_scale} block {block} sizes {ens.counts} steps {steps}", flush=True)
    b = B.BatchedDeltaSPH(ens.boxes, ens.counts, ens.pos, ens.vel, ens.mass, ens.rho, cfg, block=block)
    b.step(steps)
    torch.cuda.synchronize()
    b.check_rows()
    for d in range(len(ens.counts)):
        _assert_domain_equal(b, _single_run(ens, d, cfg, steps), d, f"cs={cell_scale}")


def test_batched_graph_replays_eager_steps_bitexact():
    B = _batched()
    ens = B.make_ensemble(sizes=[18, 33, 26], dx=DX, seed=3, device="cuda")
    print("batched graph vs eager: sizes", ens.counts, flush=True)
    a = B.BatchedDeltaSPH(ens.boxes, ens.counts, ens.pos, ens.vel, ens.mass, ens.rho, ens.cfg)
    g = B.BatchedDeltaSPH(ens.boxes, ens.counts, ens.pos, ens.vel, ens.mass, ens.rho, ens.cfg, width=a.width)
    graph = g.capture(3)
    # capture records work but runs none of it
    assert torch.equal(g.vel, a.vel) and float(g.time.sum()) == 0.0
    for rep in range(3):
        a.step(3)
        graph.launch()
        # eager work of the program between replays
        junk = torch.randn(4096, device="cuda").sum()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for name in ("pos", "vel", "rho", "uid", "dt", "time"):
            x, y = getattr(a, name), getattr(g, name)
            assert torch.equal(x, y), f"replay {rep}: {name} differs"
    graph.close()
    a.check_rows()
    g.check_rows()


def test_batched_cell_scales_agree_to_rounding():
    """Other cell sizes sum the same neighbor sets in another order."""
    B = _batched()
    ens = B.make_ensemble(sizes=[24, 40], dx=DX, seed=11, device="cuda")
    out = []
    for cs in (1.0, 1.3):
        b = B.BatchedDeltaSPH(ens.boxes, ens.counts, ens.pos, ens.vel, ens.mass, ens.rho,
                              replace(ens.cfg, cell_scale=cs))
        b.step(3)
        out.append(b.state_by_uid())
    for k, name in enumerate(("pos", "vel", "rho")):
        x, y = out[0][k], out[1][k]
        rel = ((x - y).abs().max() / y.abs().max()).item()
        print(f"cell scale 1.0 vs 1.3 after 3 steps: {name} rel diff {rel:.2e}")
        assert rel < 2e-5


def test_batched_width_overflow_is_reported():
    B = _batched()
    ens = B.make_ensemble(sizes=[20, 25], dx=DX, seed=5, device="cuda")
    b = B.BatchedDeltaSPH(ens.boxes, ens.counts, ens.pos, ens.vel, ens.mass, ens.rho, ens.cfg, width=16)
    assert b.max_neighbors() > 16
    b.step(1)
    over, longest = b.status()
    assert over >= b.n and longest > 16  # every row of every stage's list
    with pytest.raises(Exception, match="exceeded the list width"):
        b.check_rows()
    b.status(reset=True)
    assert b.status() == (0, 0)


def test_batched_autotune_returns_measured_candidate():
    B = _batched()
    ens = B.make_ensemble(sizes=[30, 45, 60], dx=DX, seed=2, device="cuda")
    best, times = B.autotune(ens, cell_scales=(1.0, 1.3), blocks=(64, 256), steps=1, warmup=1)
    assert best in times and len(times) == 4
    assert all(t > 0 for t in times.values())
    assert times[best] == min(times.values())


def test_ensemble_512_domains_properties():
    """The full ensemble: 512 periodic boxes, N_d log-uniform in [1k, 64k]
    (~8.8M particles), a captured graph of steps.  Checks: no truncated row,
    finite state, per-domain momentum conserved to float32 rounding,
    per-domain dt within the computeTimestep bounds, and the smallest and
    largest domains bit-identical to single-domain runs."""
    B = _batched()
    ens = B.make_ensemble(ndom=512, seed=0, dx=DX, device="cuda")
    n = sum(ens.counts)
    print(f"ensemble: {len(ens.counts)} domains, {n} particles, min {min(ens.counts)} max {max(ens.counts)}",
          flush=True)
    assert 7_000_000 < n < 10_000_000
    b = B.BatchedDeltaSPH(ens.boxes, ens.counts, ens.pos, ens.vel, ens.mass, ens.rho, ens.cfg)
    print(f"width {b.width}", flush=True)
    P0 = b.momentum()
    scale = (b.mass.double()[:, None] * b.vel.double().abs()).sum().item() / len(ens.counts)
    steps = 4
    graph = b.capture(steps)
    graph.launch()
    torch.cuda.synchronize()
    b.check_rows()
    assert torch.isfinite(b.pos).all() and torch.isfinite(b.vel).all() and torch.isfinite(b.rho).all()
    dP = (b.momentum() - P0).abs().max().item()
    print(f"max per-domain |dP| {dP:.3e} (mean sum m|u| per domain {scale:.3e})")
    assert dP < 1e-5 * scale
    from diffsph_amd.deltasph import timestep_constants
    dt_v, dt_c = timestep_constants(ens.cfg)
    assert (b.dt <= float(min(dt_v, dt_c)) + 1e-12).all() and (b.dt >= ens.cfg.min_dt * (1 - 1e-6)).all()
    for d in (int(np.argmin(ens.counts)), int(np.argmax(ens.counts))):
        _assert_domain_equal(b, _single_run(ens, d, ens.cfg, steps), d, "ensemble")
    graph.close()
