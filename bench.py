"""Benchmark: DFSPH particle-updates/s on the 1M-particle random-flow workload.

Workload: 1024x1024 lattice on [-1,1]^2 (periodic) + U(-0.05dx, 0.05dx)
jitter, divergence-free spectral-noise velocity (|k| <= 8, max |u| = 1),
Wendland C2 with ~50 neighbors, rho0 = 1, delta-SPH inviscid viscosity
(alpha 0.01, c_s 10), dt = 1.5 CFL acoustic (3 dx / c_s).  One step = neighbor
search + density + divergence solve + density solve + semi-implicit Euler,
exactly the reference's DFSPH scheme (v2/simulationSchemes/dfsph.py).

Prints ONE JSON line on stdout (progress on stderr):
  value = particles * timed steps / seconds (inputs resident in HBM),
  roofline = the dominant pair kernel (pressure sweeps) measured with HIP
  events on the stream it runs on, algorithmic bytes per launch / time,
  cpu_baseline = the CPU oracle (C++/OpenMP restatement, all host cores)
  on a bounded sample of the same workload.
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-nside", type=int, default=1024, help="sample size of the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--kernel-reps", type=int, default=50)
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
    from diffsph_amd.neighborhood import WORKSPACE

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    d = sampling.dfsph_random_flow(args.nside, seed=args.seed, jitter_width=0.1)
    n = d["pos"].shape[0]
    dt = 3.0 * d["dx"] / 10.0
    cfg = DFSPHConfig(dt=dt, support=d["h"])
    T = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    solver = DFSPHSolver(T(d["pos"]), T(d["vel"]), T(d["mass"]), T(d["area"]), cfg)
    log(f"workload: {n} particles, h={d['h']:.6g}, dt={dt:.6g}")
    for k in range(args.warmup):
        st = solver.step()
        log(f"warmup {k}: div it {st.div_iters} den it {st.den_iters} err {st.den_error:.3e} pairs {st.num_pairs}")
    torch.cuda.synchronize()
    it_div, it_den, errs = [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = solver.step()
        it_div.append(st.div_iters)
        it_den.append(st.den_iters)
        errs.append(st.den_error)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    secs = t1 - t0
    value = n * args.steps / secs
    mean_e, max_e = solver.density_error()
    log(f"timed {args.steps} steps: {secs * 1e3 / args.steps:.3f} ms/step, {value:.4g} particle-updates/s; "
        f"iters div {np.mean(it_div):.1f} den {np.mean(it_den):.1f}; rho err mean {mean_e:.2e} max {max_e:.2e}")

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
    ws = WORKSPACE.get(L.dsph_dfsph_workspace(n), dev)
    res = torch.empty(n, device=dev)
    it_reps = max(reps // 2, 5)
    e0.record(stream)
    _lib.check(L.dsph_dfsph_iterate(C.byref(pairs), C.byref(prm), _lib.ptr(solver.kgrad), _lib.ptr(solver.rho),
                                    _lib.ptr(solver.vol),
                                    _lib.ptr(solver.alpha_den), _lib.ptr(solver.div_src), _lib.ptr(p), _lib.ptr(acc),
                                    _lib.ptr(res), it_reps, None, _lib.ptr(ws), ws.numel(), sh))
    e1.record(stream)
    e1.synchronize()
    t_iter = e0.elapsed_time(e1) / it_reps * 1e-3
    t_b = max(t_iter - t_a, 1e-9)
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
    log(f"sweep A (pressure accel): {t_a * 1e6:.1f} us, {bytes_a / 1e6:.1f} MB -> {gbs_a:.0f} GB/s; "
        f"sweep B (pressure update): {t_b * 1e6:.1f} us, {bytes_b / 1e6:.1f} MB -> {gbs_b:.0f} GB/s; "
        f"pairs {P} ({P / n:.1f}/particle)")
    if t_b >= t_a:
        kname, ach, tb, byt = "k_dfsph_sweep_b", gbs_b, t_b, bytes_b
    else:
        kname, ach, tb, byt = "k_dfsph_sweep_a", gbs_a, t_a, bytes_a
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kname,
                "kernel_us": round(tb * 1e6, 2), "algorithmic_bytes": int(byt)}

    # ---- CPU baseline: the oracle on the host cores
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        threads = min(os.cpu_count() or 1, 16)
        orc.set_threads(threads)
        dc = sampling.dfsph_random_flow(args.cpu_nside, seed=args.seed, jitter_width=0.1)
        nc = dc["pos"].shape[0]
        dom = orc.make_domain((-1, -1), (1, 1), (True, True))
        ost = orc.DfsphState(dc["pos"].copy(), dc["vel"].copy(), dc["mass"].copy(), dc["area"].copy(),
                             np.arange(nc, dtype=np.int32))
        dtc = 3.0 * dc["dx"] / 10.0
        # one untimed step so that the timed one has the same warm start as the GPU steps
        ost, _ = orc.dfsph_step(ost, dom, dc["h"], dtc)
        c0 = time.perf_counter()
        for _ in range(args.cpu_steps):
            ost, info = orc.dfsph_step(ost, dom, dc["h"], dtc)
        c1 = time.perf_counter()
        cval = nc * args.cpu_steps / (c1 - c0)
        cpu = {"value": round(cval, 1), "unit": "particle-updates/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} DFSPH step(s) (after 1 warm-start step) of the {nc}-particle "
                         f"workload, C++/OpenMP oracle, {threads} threads; "
                         f"iters div {info['it_div']} den {info['it_den']}"}
        log(f"cpu baseline: {cval:.4g} particle-updates/s ({(c1 - c0):.1f} s, {threads} threads)")

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
                   "solver": "plain-jacobi (warm start 0.75 p_prev)", "dt": dt,
                   "iters_div_mean": float(np.mean(it_div)), "iters_den_mean": float(np.mean(it_den)),
                   "den_residual_mean": float(np.mean(errs)), "rho_err_mean": mean_e, "rho_err_max": max_e,
                   "pairs": P},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
