"""Benchmark: DFSPH particle-updates/s on the 1M-particle random-flow workload,
plain (the reference's 0.75 p_prev warm start) vs the learned warm start.

Workload: 1024x1024 lattice on [-1,1]^2 (periodic) + U(-0.05dx, 0.05dx)
jitter, divergence-free spectral-noise velocity (|k| <= 8, max |u| = 1),
Wendland C2 with ~50 neighbors, rho0 = 1, delta-SPH inviscid viscosity
(alpha 0.01, c_s 10), dt = 1.5 CFL acoustic (3 dx / c_s).  One step = neighbor
search + density + divergence solve + density solve + semi-implicit Euler,
exactly the reference's DFSPH scheme (v2/simulationSchemes/dfsph.py).

The warm-start network is trained first (untimed, on trajectories of other
seeds, on the GPU); then the plain and the learned solver each run `warmup`
untimed and `steps` timed steps from the same initial state.

Prints ONE JSON line on stdout (progress on stderr):
  value = particles * timed steps / seconds of the learned solver (inputs
  resident in HBM); the plain solver's figures sit beside it,
  roofline = the dominant pair kernel (pressure sweeps) measured with HIP
  events on the stream it runs on, algorithmic bytes per launch / time,
  cpu_baseline = the CPU oracle (C++/OpenMP restatement, host cores) running
  the plain solver's first timed step from the same state.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--train-rounds", type=int, default=3)
    ap.add_argument("--train-steps", type=int, default=24, help="trajectory steps per training round")
    ap.add_argument("--no-ablation", action="store_true", help="skip the p_prev (warm start 1.0) run")
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device copy (reference kernel for PMC byte counters)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the workload runs on a single MI355X); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)
    import torch

    from diffsph_amd import _lib, sampling
    from diffsph_amd.dfsph import DFSPHConfig, DFSPHSolver
    from diffsph_amd.learned import train_warm_start

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    d = sampling.dfsph_random_flow(args.nside, seed=args.seed, jitter_width=0.1)
    n = d["pos"].shape[0]
    dt = 3.0 * d["dx"] / 10.0
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731

    def make(seed, warm=0.75):
        dd = d if seed == args.seed else sampling.dfsph_random_flow(args.nside, seed=seed, jitter_width=0.1)
        return DFSPHSolver(T(dd["pos"]), T(dd["vel"]), T(dd["mass"]), T(dd["area"]),
                           DFSPHConfig(dt=dt, support=d["h"], warm_start=warm))

    log(f"workload: {n} particles, h={d['h']:.6g}, dt={dt:.6g}")
    t0 = time.perf_counter()
    net = train_warm_start(lambda r: make(args.seed + 100 + r), rounds=args.train_rounds, steps=args.train_steps,
                           sample=0.1, train_steps=400, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    log(f"warm-start network trained in {t_train:.1f} s")

    def run(label, solver, snapshot=None):
        for k in range(args.warmup):
            st = solver.step()
            log(f"{label} warmup {k}: div it {st.div_iters} den it {st.den_iters} pairs {st.num_pairs}")
        if snapshot is not None:
            snapshot.update(pos=solver.pos.cpu().numpy(), vel=solver.vel.cpu().numpy(),
                            mass=solver.mass.cpu().numpy(), area=solver.areas.cpu().numpy(),
                            uid=solver.uid.cpu().numpy(), p_div=solver.p_div.cpu().numpy(),
                            p_den=solver.p_den.cpu().numpy())
        torch.cuda.synchronize()
        it_div, it_den, errs = [], [], []
        c0 = time.perf_counter()
        for _ in range(args.steps):
            st = solver.step()
            it_div.append(st.div_iters)
            it_den.append(st.den_iters)
            errs.append(st.den_error)
        torch.cuda.synchronize()
        secs = time.perf_counter() - c0
        me, mx = solver.density_error()
        r = {"value": n * args.steps / secs, "ms_per_step": secs * 1e3 / args.steps,
             "iters_div_mean": float(np.mean(it_div)), "iters_den_mean": float(np.mean(it_den)),
             "den_residual_mean": float(np.mean(errs)), "rho_err_mean": me, "rho_err_max": mx,
             "first_step_iters_div": it_div[0]}
        log(f"{label}: {r['ms_per_step']:.3f} ms/step, {r['value']:.4g} particle-updates/s; iters div "
            f"{r['iters_div_mean']:.1f} den {r['iters_den_mean']:.1f}; rho err mean {me:.2e} max {mx:.2e}")
        return r

    snap = {}
    plain = run("plain", make(args.seed), snapshot=snap)
    prev = None if args.no_ablation else run("p_prev", make(args.seed, warm=1.0))
    solver = make(args.seed)
    net.install(solver)
    learned = run("learned", solver)
    log(f"judge: {net.judge.judged} judged steps, net used {net.judge.used_net}, plain used {net.judge.used_plain}")
    value = learned["value"]
    secs = learned["ms_per_step"] * args.steps * 1e-3

    if args.copy_calibration:
        src = torch.empty(64 << 20, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- roofline of the pressure sweeps (HIP events on the solver's stream)
    L = _lib.lib()
    pairs = solver.nl.pairs()
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    p = solver.p_den.clone()
    acc = torch.empty(n, 2, device=dev)
    reps = args.kernel_reps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(3):
        _lib.check(L.dsph_pressure_accel(C.byref(pairs), _lib.ptr(solver.kgrad), _lib.ptr(solver.vol),
                                         _lib.ptr(solver.rho), _lib.ptr(p), _lib.ptr(acc), sh))
    e0.record(stream)
    for _ in range(reps):
        _lib.check(L.dsph_pressure_accel(C.byref(pairs), _lib.ptr(solver.kgrad), _lib.ptr(solver.vol),
                                         _lib.ptr(solver.rho), _lib.ptr(p), _lib.ptr(acc), sh))
    e1.record(stream)
    e1.synchronize()
    t_a = e0.elapsed_time(e1) / reps * 1e-3
    prm = _lib.DfsphParams()
    prm.dt, prm.rho0, prm.omega, prm.error_threshold = dt, 1.0, 0.3, 5e-4
    prm.min_iters, prm.max_iters, prm.clamp_pressure, prm.source_density = 1, 1, 1, 1
    res = torch.empty(n, device=dev)
    parts = torch.empty((n + 255) // 256, dtype=torch.float64, device=dev)

    def sweep_b():
        _lib.check(L.dsph_pressure_update(C.byref(pairs), C.byref(prm), _lib.ptr(solver.kgrad), _lib.ptr(solver.vol),
                                          _lib.ptr(solver.div_src), _lib.ptr(solver.alpha_den), _lib.ptr(acc),
                                          _lib.ptr(p), _lib.ptr(res), _lib.ptr(parts), sh))

    for _ in range(3):
        sweep_b()
    e0.record(stream)
    for _ in range(reps):
        sweep_b()
    e1.record(stream)
    e1.synchronize()
    t_b = e0.elapsed_time(e1) / reps * 1e-3
    P = solver.nl.num_pairs
    # Algorithmic bytes (each per-particle array once per launch, each pair
    # entry once; gathered neighbor fields and list padding not counted):
    # sweep A: p 4 + rho 4 + count 4 + a(write) 8 = 20 B/particle;
    #          j 4 + V_j gradW_ij 8 = 12 B/pair
    # sweep B: a 8 + p 4(r)+4(w) + s 4 + alpha 4 + count 4 + residual 4(w) = 32 B/particle; 12 B/pair
    bytes_a = 20 * n + 12 * P
    bytes_b = 32 * n + 12 * P
    gbs_a = bytes_a / t_a / 1e9
    gbs_b = bytes_b / t_b / 1e9
    log(f"sweep A (dsph_pressure_accel): {t_a * 1e6:.1f} us, {bytes_a / 1e6:.1f} MB -> {gbs_a:.0f} GB/s; "
        f"sweep B (dsph_pressure_update): {t_b * 1e6:.1f} us, {bytes_b / 1e6:.1f} MB -> {gbs_b:.0f} GB/s; "
        f"pairs {P} ({P / n:.1f}/particle)")
    if t_b >= t_a:
        kname, ach, tb, byt = "k_dfsph_sweep_b", gbs_b, t_b, bytes_b
    else:
        kname, ach, tb, byt = "k_dfsph_sweep_a", gbs_a, t_a, bytes_a
    # HBM bytes per launch from the committed PMC profile of the same kernel
    # (FETCH_SIZE / WRITE_SIZE, scaled by a 256 MiB copy): its ratio to the
    # algorithmic bytes of the profiled launch, applied to this launch.
    traffic, ratio = None, None
    pmc_path = os.path.join(ROOT, "profiles", "dfsph_sweeps_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            pmc = json.load(f)
        prof_bytes = pmc["traffic_bytes"].get(kname)
        prof_alg = pmc.get("algorithmic_bytes", {}).get(kname)
        if prof_bytes and prof_alg:
            ratio = prof_bytes / prof_alg
            traffic = int(round(ratio * byt))
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
                "kernel_us": round(tb * 1e6, 2), "algorithmic_bytes": int(byt),
                "traffic_over_algorithmic": None if ratio is None else round(ratio, 4),
                "other_sweep": {"kernel": "k_dfsph_sweep_a" if kname == "k_dfsph_sweep_b" else "k_dfsph_sweep_b",
                                "achieved": round(gbs_a if kname == "k_dfsph_sweep_b" else gbs_b, 1),
                                "kernel_us": round((t_a if kname == "k_dfsph_sweep_b" else t_b) * 1e6, 2)}}

    # ---- CPU baseline: the oracle on the host cores, continuing the plain
    # solver's trajectory from the state at the start of its timed window
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        threads = min(os.cpu_count() or 1, 16)
        orc.set_threads(threads)
        dom = orc.make_domain((-1, -1), (1, 1), (True, True))
        ost = orc.DfsphState(snap["pos"], snap["vel"], snap["mass"], snap["area"], snap["uid"].astype(np.int32),
                             snap["p_div"], snap["p_den"])
        c0 = time.perf_counter()
        its = []
        for _ in range(args.cpu_steps):
            ost, info = orc.dfsph_step(ost, dom, d["h"], dt)
            its.append(int(info["it_div"]))
        c1 = time.perf_counter()
        cval = n * args.cpu_steps / (c1 - c0)
        same = its[0] == plain["first_step_iters_div"]
        cpu = {"value": round(cval, 1), "unit": "particle-updates/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} plain DFSPH step(s) of the {n}-particle workload from the GPU plain "
                         f"solver's state at its first timed step (step {args.warmup}), C++/OpenMP oracle, "
                         f"{threads} threads; div iterations {its} (GPU plain mean over the timed steps "
                         f"{plain['iters_div_mean']:.1f}; first timed step matches the GPU: {same})"}
        log(f"cpu baseline: {cval:.4g} particle-updates/s ({(c1 - c0):.1f} s, {threads} threads, div its {its})")

    def rnd(r):
        return None if r is None else {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}

    out = {
        "metric": "particle-updates/s at 1M (DFSPH plain vs learned warm-start, rho-err<=0.1%)",
        "value": round(value, 1),
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"dfsph_random_flow_{args.nside}x{args.nside}", "particles": n,
                   "solver": "DFSPH, learned warm start (MLP 16-64-64-64-1, bf16 MFMA) with plain-Jacobi judge",
                   "dt": dt, "pairs": P, "learned": rnd(learned), "plain": rnd(plain),
                   "p_prev_warm_start": rnd(prev),
                   "speedup_vs_plain": round(learned["value"] / plain["value"], 3),
                   "train_seconds": round(t_train, 2),
                   "judge": {"judged": net.judge.judged, "net": net.judge.used_net,
                             "plain": net.judge.used_plain}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
