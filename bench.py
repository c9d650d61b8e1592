#!/usr/bin/env python3
"""Benchmark of the gfx950 Barnes-Hut force step (bh.Solver.Accelerations).

Workload: the single-galaxy model -- N = 1,048,576 (70 % exponential
Toomre-Q disk, 30 % Hernquist halo with 5x the disk mass), theta 0.5, Plummer
softening 0.02, inputs resident in HBM.  One step = one full force evaluation:
collect, Morton keys, radix sort, sorted copies, octree topology, moments and
the FP32 group walk, written back in state order.

Printed (one JSON line): particles/s of the FP32 step; interactions/s; the
FP32-vs-FP64-walk and FP32-vs-direct-sum relative force errors (p50/p99/max);
the FP64 GPU walk's throughput; the walk kernel's achieved FP32 FLOP/s against
the MI355X vector peak (HIP events around the walk on the solver's stream);
and the CPU oracle's throughput on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s at p99 force err ≤1e-3 (FP32 vs FP64 walk); batched steps/s"
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector, datasheet
HBM_PEAK_GBS = 8000.0
FLOPS_PP = 18              # bh.go:145-150 counting rule (add/sub/mul/div/sqrt = 1, fma = 2)
FLOPS_PC = 50              # bh.go:114-131


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--eps", type=float, default=0.02)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-errors", action="store_true", help="skip the error measurements")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
    return ap.parse_args()


def errors(got, want):
    from gravitas_amd.ics import error_stats
    return error_stats(got, want)


def cpu_baseline(state, theta, eps, budget):
    """The CPU oracle (C restatement of bh.Solver: serial build, ParallelFor
    walk) on the host cores: one warm-up call, then timed calls of the full
    Accelerations until the budget is spent.  Sample: the first n_s bodies of
    the workload in state order, n_s chosen so a call takes a few seconds."""
    import oracle
    workers = min(16, os.cpu_count() or 1)
    n = len(state["px"])
    # size the sample from a small probe
    probe = min(n, 65536)
    sl = lambda k: {f: state[f][:k] for f in ("px", "py", "pz", "m", "kind")}
    o = oracle.Solver(theta=theta, workers=workers)
    s = sl(probe)
    t0 = time.perf_counter()
    o.accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"], eps=eps)
    tp = time.perf_counter() - t0
    ns = int(min(n, max(probe, probe * (budget / 4.0) / max(tp, 1e-3))))
    # sample of the same model: a state-order prefix mixes disk and halo only
    # if it spans both, so take an evenly strided subset instead
    idx = np.linspace(0, n - 1, ns).astype(np.int64)
    s = {f: np.ascontiguousarray(state[f][idx]) for f in ("px", "py", "pz", "m", "kind")}
    o.accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"], eps=eps)  # warm-up
    times = []
    t_start = time.perf_counter()
    while not times or (time.perf_counter() - t_start < budget / 2 and len(times) < 5):
        t0 = time.perf_counter()
        o.accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"], eps=eps)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": ns / t, "unit": "particles/s", "cores": workers, "kind": "port",
            "sample": f"{ns} bodies evenly strided from the {n}-body workload (same theta/eps), "
                      f"full Accelerations (build + walk), median of {len(times)} after 1 warm-up; "
                      f"C restatement of bh.Solver, -O2 -ffp-contract=off, {workers} threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (every box here has one GPU; "
              "the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from gravitas_amd import DeviceState, Params, Solver, gpu_available, ics
    from gravitas_amd.hip import synchronize
    reason = gpu_available()
    if reason is not None:
        print(f"no GPU backend: {reason}", file=sys.stderr)
        sys.exit(1)

    n, theta, eps = args.n, args.theta, args.eps
    print(f"[bench] building disk+halo model n={n}", file=sys.stderr, flush=True)
    state = ics.disk_halo(n, seed=2, eps=eps)
    ds = DeviceState.from_host(state)
    params = Params(softening=eps)
    solver = Solver(theta=theta, walk="fp32", timing=True)

    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(max(1, args.warmup)):
        solver.accelerations(ds, params)
    synchronize()

    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    walk_ms, phases = [], []
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.accelerations(ds, params)
        ph = solver.phase_ms()
        walk_ms.append(ph["walk"])
        phases.append(ph)
    synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    st = solver.stats()
    acc32 = ds.acc_host()
    npp, npc = st["pp_interactions"], st["pc_interactions"]
    walk_s = float(np.mean(walk_ms)) * 1e-3
    flops = FLOPS_PP * npp + FLOPS_PC * npc
    achieved = flops / walk_s / 1e12
    phase_avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}

    out = {
        "metric": METRIC,
        "value": n / (ms_per_step * 1e-3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "disk-halo-1m", "n": n, "theta": theta, "eps": eps,
                   "model": "70% exponential disk (R_d 1, z0 0.1, Q 1.5) + 30% Hernquist halo (a 5, 5x disk mass)",
                   "walk": "fp32 group walk (64 stars per wavefront), Kahan-compensated per-star sums",
                   "parallelism": "single GPU"},
        "interactions_per_s": (npp + npc) / (ms_per_step * 1e-3),
        "interactions_per_particle": {"pp": npp / n, "pc": npc / n},
        "tree": {"nodes": st["n_nodes"], "depth": st["depth"], "max_stack": st["max_stack"]},
        "walk_lists": {"groups": (n + 63) // 64, "cell_entries_per_group": st["list_entries"] / ((n + 63) // 64),
                       "cell_lane_utilization": npc / max(1, 64 * st["list_entries"]),
                       "rounds_per_group": st["rounds"] / ((n + 63) // 64),
                       "mixed_nodes_per_group": st["mixed_nodes"] / ((n + 63) // 64),
                       "max_list": st["max_list"]},
        "phases_ms": phase_avg,
        "roofline": {"bound": "valu", "kernel": "k_walk32", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": flops, "walk_ms": walk_s * 1e3},
    }

    if not args.no_errors:
        print("[bench] FP64 walk and direct-sum sample", file=sys.stderr, flush=True)
        s64 = Solver(theta=theta, walk="fp64", timing=True)
        s64.accelerations(ds, params)  # warm-up
        t64 = []
        for _ in range(2):
            s64.accelerations(ds, params)
            t64.append(s64.phase_ms()["total"])
        acc64 = ds.acc_host()
        e64 = errors(acc32, acc64)
        targets = ics.sample_targets(n, 4096, 7)
        direct = solver.direct_sample(ds, params, targets).numpy()
        out["err_fp32_vs_fp64_walk"] = e64
        out["err_fp32_vs_direct"] = errors(acc32[targets], direct)
        out["err_fp64_walk_vs_direct"] = errors(acc64[targets], direct)
        fp64_ps = n / (float(np.median(t64)) * 1e-3)
        gpu32_ps = n / (phase_avg["total"] * 1e-3)
        out["fp64_walk"] = {"particles_per_s": fp64_ps, "step_ms": float(np.median(t64)),
                            "fp32_over_fp64": gpu32_ps / fp64_ps}
        out["gate_p99_le_1e-3"] = bool(e64["p99"] <= 1e-3)

    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(state, theta, eps, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = out["value"] / cb["value"]

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
