#!/usr/bin/env python3
"""Benchmark of the gfx950 Barnes-Hut force step (bh.Solver.Accelerations).

Default workload (the headline): the single-galaxy model -- N = 1,048,576
(70 % exponential Toomre-Q disk, 30 % Hernquist halo with 5x the disk mass),
theta 0.5, Plummer softening 0.02, inputs resident in HBM.  One step = one
full force evaluation: collect, Morton keys, radix sort, sorted copies,
octree topology, moments and the FP32 group walk, written back in state order.

Printed (one JSON line): particles/s of the FP32 step; interactions/s; the
FP32-vs-FP64-walk and FP32-vs-direct-sum relative force errors (p50/p99/max);
the FP64 GPU walk's throughput; the walk kernel's achieved FP32 FLOP/s against
the MI355X vector peak (HIP events around the walk on the solver's stream);
the CPU oracle's throughput on the host (16 threads and 1 thread, with the
CPU model); and, as "batched", the metric's second half: steps/s of the
256-system batched workload (hipGraph KDK replays) against 256 separate
per-system steps, with the autotuner's table.

Other workloads (--config): plummer-16k (force + one KDK step, FP64 and FP32),
clumps-16m (precision ladder FP64 / FP32 / FP32 with FP16-stored quadrupoles),
batched-256 (256 systems in one arena, hipGraph KDK vs separate calls) and
two-disk-4m (4,194,304-body encounter, block time steps on 4 rungs, FP16
quadrupoles, FP64 energy drift over --steps global steps), disk-halo-1m-holes
(split steps with three black holes), disk-halo-1m-holes-retarded (the same
with retarded hole gravity at c = 10), sph-sphere-1m (SPH hydro forces) and
spin-cluster-1m (split steps with a spinning hole and frame dragging).
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
ENCOUNTER_TOP = 0.03        # two-disk encounter: the global step puts this fraction of the stars on the top rung
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector, datasheet
# SPH counting rule (add/sub/mul/div/sqrt = 1; comparisons, min/max not counted):
# a force pair 62 (sph.go:219-246 without the viscosity branch), + 19 when it
# approaches (:231-235); a density term 20 (r^2, sqrt, W, m W, the sum).
FLOPS_SPH_PAIR = 62
FLOPS_SPH_VISC = 19
FLOPS_SPH_DENSITY = 20
CONFIGS = ("disk-halo-1m", "plummer-16k", "clumps-16m", "batched-256", "two-disk-4m", "disk-halo-1m-holes",
           "disk-halo-1m-holes-retarded", "sph-sphere-1m", "spin-cluster-1m")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="disk-halo-1m", choices=CONFIGS)
    ap.add_argument("--n", type=int, default=None, help="override the body count")
    ap.add_argument("--theta", type=float, default=None)
    ap.add_argument("--eps", type=float, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-errors", action="store_true", help="skip the error measurements")
    ap.add_argument("--no-batched", action="store_true", help="disk-halo-1m: skip the batched-256 leg")
    ap.add_argument("--cpu-seconds", type=float, default=16.0,
                    help="budget of the CPU baseline samples (half on all host threads, half on one)")
    ap.add_argument("--T", type=float, default=None, help="two-disk-4m: override the global step")
    ap.add_argument("--shape", default=None,
                    help="disk-halo-1m: walk shape 'stack,staging,waves' (64-star groups, work queue) instead of "
                         "the autotuner's choice, so that profiler passes all run the same kernel")
    ap.add_argument("--calibrate", action="store_true",
                    help="also launch 3 drifts with dt = 0 over the resident bodies (k_drift reads 48 and writes "
                         "24 bytes per body): the known-size kernel that scales HBM counters")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def errors(got, want):
    from gravitas_amd.ics import error_stats
    return error_stats(got, want)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(state, theta, eps, budget):
    """The CPU oracle (C restatement of bh.Solver: serial build, ParallelFor
    walk) on the host: min(16, cores) threads (the host share of one GPU)
    and 1 thread, each with half the budget (see cpu_sample)."""
    out = cpu_sample(state, theta, eps, budget / 2, min(16, os.cpu_count() or 1))
    one = cpu_sample(state, theta, eps, budget / 2, 1)
    out["cpu_model"] = cpu_model()
    out["host_cpus"] = os.cpu_count()
    out["one_core"] = {k: one[k] for k in ("value", "unit", "cores", "sample")}
    return out


def cpu_sample(state, theta, eps, budget, workers):
    """One warm-up call, then timed calls of the full Accelerations until the
    budget is spent, on an evenly strided sample of the workload sized so
    that one call takes about a quarter of the budget."""
    import oracle
    n = len(state["px"])
    o = oracle.Solver(theta=theta, workers=workers)
    probe = min(n, 65536)
    idx = np.linspace(0, n - 1, probe).astype(np.int64)
    s = {f: np.ascontiguousarray(state[f][idx]) for f in ("px", "py", "pz", "m", "kind")}
    t0 = time.perf_counter()
    o.accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"], eps=eps)
    tp = time.perf_counter() - t0
    ns = int(min(n, max(probe, probe * (budget / 4.0) / max(tp, 1e-3))))
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


def timed_steps(solver, ds, params, steps, warmup, phase_steps=5):
    """Warm-up, then `steps` synchronous force steps (each an eager sequence
    of launches on the solver's stream, ending in one synchronisation) timed
    on the host clock; then `phase_steps` more steps with HIP events recorded
    on the solver's stream between the phases (per-phase GPU times)."""
    from gravitas_amd.hip import synchronize
    timing = solver.timing
    solver.timing = False
    for _ in range(max(3, warmup)):
        solver.accelerations(ds, params)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.accelerations(ds, params)
    synchronize()
    wall = time.perf_counter() - t0
    solver.timing = True
    phases = []
    for _ in range(max(1, phase_steps)):
        solver.accelerations(ds, params)
        phases.append(solver.phase_ms())
    solver.timing = timing
    phase_avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    return wall * 1e3 / steps, phase_avg, solver.stats()


def walk_summary(n, st, phase_avg):
    npp, npc = st["pp_interactions"], st["pc_interactions"]
    walk_s = phase_avg["walk"] * 1e-3
    flops = FLOPS_PP * npp + FLOPS_PC * npc
    achieved = flops / walk_s / 1e12
    ng = (n + 63) // 64
    return {
        "interactions_per_particle": {"pp": npp / n, "pc": npc / n},
        "tree": {"nodes": st["n_nodes"], "depth": st["depth"], "max_stack": st["max_stack"]},
        "walk_lists": {"groups": ng, "cell_entries_per_group": st["list_entries"] / ng,
                       "cell_lane_utilization": npc / max(1, 64 * st["list_entries"]),
                       "rounds_per_group": st["rounds"] / ng, "mixed_nodes_per_group": st["mixed_nodes"] / ng,
                       "max_list": st["max_list"]},
        "phases_ms": phase_avg,
        "roofline": {"bound": "valu", "kernel": "k_walk32", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": flops, "walk_ms": walk_s * 1e3,
                     "note": "FP32 vector ALU bound (no dense contraction, so no MFMA); peak = MI355X FP32 vector "
                             "157.3 TFLOP/s (packed); with packed FP32 disabled the issue ceiling is 78.6"},
    }


def measured_traffic(kernel, workload, n, theta):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a copy kernel of known
    size), or None if no such measurement exists for this configuration."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload.replace('-', '_')}_traffic.json")
    try:
        t = json.load(open(path))
    except (OSError, ValueError):
        return None
    if t.get("n") != n or t.get("theta") != theta:
        return None
    return t["traffic_bytes"]


def base_line(args, n, ms_per_step, dtype, config):
    return {
        "metric": METRIC, "value": n / (ms_per_step * 1e-3), "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config,
    }


def bench_disk_halo(args):
    from gravitas_amd import DeviceState, Params, Solver, ics
    n = args.n or 1048576
    theta = 0.5 if args.theta is None else args.theta
    eps = 0.02 if args.eps is None else args.eps
    log(f"building disk+halo model n={n}")
    state = ics.disk_halo(n, seed=2, eps=eps)
    ds = DeviceState.from_host(state)
    params = Params(softening=eps)
    if args.shape:
        stack, parts, waves = (int(v) for v in args.shape.split(","))
        solver = Solver(theta=theta, walk="fp32", timing=True, group_size=64, persistent=2, stack=stack,
                        parts=parts, waves=waves)
        table = []
        log(f"walk shape {args.shape} (no autotune)")
    else:
        solver = Solver(theta=theta, walk="fp32", timing=True)
        log("autotune the walk")
        table = solver.autotune(ds, [0, n], theta, eps, reps=3)
    log(f"timing {args.steps} steps after {args.warmup} warm-up")
    ms, phases, st = timed_steps(solver, ds, params, args.steps, args.warmup)
    acc32 = ds.acc_host()
    if args.calibrate:
        from gravitas_amd.hip import synchronize
        for _ in range(3):
            solver.drift(ds, 0.0)
        synchronize()
    out = base_line(args, n, ms, "fp32", {
        "workload": "disk-halo-1m", "n": n, "theta": theta, "eps": eps,
        "bodies": "70% exponential disk (R_d 1, z0 0.1, Q 1.5) + 30% Hernquist halo (a 5, 5x disk mass)",
        "walk": "fp32 group walk (64 stars per wavefront), Kahan-compensated per-star sums",
        "parallelism": "single GPU"})
    out["interactions_per_s"] = (st["pp_interactions"] + st["pc_interactions"]) / (ms * 1e-3)
    out.update(walk_summary(n, st, phases))
    out["roofline"]["traffic"] = measured_traffic("k_walk32", "disk-halo-1m", n, theta)
    if table:
        out["walk_autotune"] = {"chosen": tune_name(min(table, key=lambda r: r["ms"])),
                                "ms_per_walk": {tune_name(r): r["ms"] for r in table}}
    else:
        out["walk_shape"] = {"stack": solver.stack, "staging": solver.parts, "waves": solver.waves,
                             "group": 64, "queue": True}
    if not args.no_errors:
        log("FP64 walk and direct-sum sample")
        s64 = Solver(theta=theta, walk="fp64", timing=True)
        ms64, ph64, _ = timed_steps(s64, ds, params, 2, 1)
        acc64 = ds.acc_host()
        e64 = errors(acc32, acc64)
        targets = ics.sample_targets(n, 4096, 7)
        direct = solver.direct_sample(ds, params, targets).numpy()
        out["err_fp32_vs_fp64_walk"] = e64
        out["err_fp32_vs_direct"] = errors(acc32[targets], direct)
        out["err_fp64_walk_vs_direct"] = errors(acc64[targets], direct)
        out["fp64_walk"] = {"particles_per_s": n / (ms64 * 1e-3), "step_ms": ms64, "walk_ms": ph64["walk"],
                            "fp32_over_fp64_step": ms64 / ms, "fp32_over_fp64_walk": ph64["walk"] / phases["walk"]}
        out["gate_p99_le_1e-3"] = bool(e64["p99"] <= 1e-3)
        if args.theta is None:
            log("theta 0.7 line")
            s7 = Solver(theta=0.7, walk="fp32", timing=True)
            ms7, ph7, st7 = timed_steps(s7, ds, params, max(3, args.steps // 4), 1)
            acc7 = ds.acc_host()
            s7d = Solver(theta=0.7, walk="fp64")
            s7d.accelerations(ds, params)
            out["theta_0.7"] = {"particles_per_s": n / (ms7 * 1e-3), "ms_per_step": ms7, "walk_ms": ph7["walk"],
                                "err_fp32_vs_fp64_walk": errors(acc7, ds.acc_host()),
                                "err_fp32_vs_direct": errors(acc7[targets], direct),
                                "interactions_per_particle": {"pp": st7["pp_interactions"] / n,
                                                              "pc": st7["pc_interactions"] / n}}
    if not args.no_cpu:
        log("CPU baseline")
        cb = cpu_baseline(state, theta, eps, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = out["value"] / cb["value"]
    if not args.no_batched:
        # the metric's second half: batched steps/s of the 256-system workload
        del ds
        log("batched leg")
        _, _, out["batched"] = batched_measure(max(4, args.steps // 2), 2)
    return out


def bench_plummer(args):
    """Plain reference method: 16,384 equal-mass Plummer stars, theta 0.5,
    eps 0.01, one force evaluation plus one KDK step of dt 1e-3 (FP64 walk,
    bit-identical to the reference algorithm; FP32 walk alongside)."""
    from gravitas_amd import DeviceState, Params, Solver, ics
    from gravitas_amd.hip import synchronize
    n = args.n or 16384
    theta = 0.5 if args.theta is None else args.theta
    eps = 0.01 if args.eps is None else args.eps
    state = ics.plummer_equal_mass(n, seed=1)
    params = Params(softening=eps, dt=1e-3)
    res = {}
    for walk in ("fp64", "fp32"):
        solver = Solver(theta=theta, walk=walk, timing=True)
        ds = DeviceState.from_host(state)
        solver.step(ds, params, acc_valid=False)  # warm-up
        times = []
        for _ in range(args.steps):
            ds = DeviceState.from_host(state)
            synchronize()
            t0 = time.perf_counter()
            solver.step(ds, params, acc_valid=False)  # force evaluation + one KDK step (two force calls)
            synchronize()
            times.append(time.perf_counter() - t0)
        res[walk] = {"ms": float(np.median(times)) * 1e3, "acc": ds.acc_host(), "stats": solver.stats()}
    ds = DeviceState.from_host(state)
    direct = Solver().direct_sample(ds, params, np.arange(n)).numpy()
    s64 = Solver(theta=theta, walk="fp64")
    s64.accelerations(ds, params)
    a0 = ds.acc_host()
    ms = res["fp32"]["ms"]
    out = base_line(args, n, ms, "fp32", {"workload": "plummer-16k", "n": n, "theta": theta, "eps": eps,
                                          "step": "initial forces + one KDK step of dt 1e-3 (2 force evaluations)"})
    out["fp64_step_ms"] = res["fp64"]["ms"]
    out["err_fp64_tree_vs_direct"] = errors(a0, direct)
    out["interactions_per_particle"] = {"pp": res["fp32"]["stats"]["pp_interactions"] / n,
                                        "pc": res["fp32"]["stats"]["pc_interactions"] / n}
    if not args.no_cpu:
        import oracle
        o = oracle.Solver(theta=theta, workers=min(16, os.cpu_count() or 1))
        o.accelerations(state["px"], state["py"], state["pz"], state["m"], eps=eps)
        t0 = time.perf_counter()
        for _ in range(3):
            o.accelerations(state["px"], state["py"], state["pz"], state["m"], eps=eps)
        t = (time.perf_counter() - t0) / 3
        out["cpu_baseline"] = {"value": n / t, "unit": "particles/s", "cores": o.workers, "kind": "port",
                               "sample": f"full {n}-body Accelerations, mean of 3 after 1 warm-up"}
    return out


def bench_clumps(args):
    """Clustered stress case: 16,777,216 stars in 1,000 Plummer clumps, theta
    0.5, eps 1e-4; precision ladder FP64 / FP32 / FP32 + FP16 quadrupoles."""
    from gravitas_amd import DeviceState, Params, Solver, ics
    n = args.n or 16777216
    theta = 0.5 if args.theta is None else args.theta
    eps = 1e-4 if args.eps is None else args.eps
    log(f"building clumps n={n}")
    state = ics.clumps(n, 1000, seed=5)
    ds = DeviceState.from_host(state)
    params = Params(softening=eps)
    targets = ics.sample_targets(n, 4096, 7)
    direct = Solver().direct_sample(ds, params, targets).numpy()
    ladder, accs = {}, {}
    for walk in ("fp64", "fp32", "fp32-q16"):
        log(f"walk {walk}")
        solver = Solver(theta=theta, walk=walk, timing=True)
        steps = 2 if walk == "fp64" else args.steps
        ms, ph, st = timed_steps(solver, ds, params, steps, 1)
        accs[walk] = ds.acc_host()
        ladder[walk] = {"ms_per_step": ms, "particles_per_s": n / (ms * 1e-3), "phases_ms": ph,
                        "err_vs_direct": errors(accs[walk][targets], direct)}
        if walk != "fp64":
            ladder[walk]["err_vs_fp64_walk"] = errors(accs[walk], accs["fp64"])
            ladder[walk].update(walk_summary(n, st, ph))
    ms = ladder["fp32"]["ms_per_step"]
    out = base_line(args, n, ms, "fp32", {"workload": "clumps-16m", "n": n, "theta": theta, "eps": eps,
                                          "bodies": "1000 Plummer clumps (scale 0.005), log-normal masses, unit cube"})
    out["ladder"] = ladder
    out["roofline"] = ladder["fp32"]["roofline"]
    return out


def tune_name(r):
    return (f"group{r['group_size']}-{'queue' if r['persistent'] == 2 else 'grid'}"
            f"-stack{r['stack']}-staging{r['parts']}" + (f"-waves{r['waves']}" if r['waves'] != 4 else ""))


def batched_measure(steps, warmup):
    """256 independent Plummer/Hernquist systems, N log-uniform in [1024, 65536]
    (seeded), theta 0.6, eps 0.01 N^(-1/3), one arena.  Autotunes the walk
    (group size 32/64, persistent work queue or one workgroup per group) on
    eager steps, captures the best as a hipGraph KDK step and replays it;
    compared with the same step done as 256 separate per-system calls.
    Returns the body count, the graph's ms per step and the summary."""
    from gravitas_amd import DeviceState, Params, Solver, ics
    from gravitas_amd.hip import synchronize
    systems, eps = ics.batched_systems(256)
    sizes = np.array([len(s["px"]) for s in systems], dtype=np.int64)
    offsets = np.r_[0, np.cumsum(sizes)].astype(np.int64)
    state = ics.concat(*systems)
    n = int(offsets[-1])
    theta, dt = 0.6, 1e-3
    log(f"batched: {len(sizes)} systems, n={n}")
    # the library's autotuner on the batch's walk (group size, work queue,
    # stack and staging sizes), then the chosen walk in a captured KDK step
    solver = Solver(walk="fp32", timing=True)
    ds = DeviceState.from_host(state)
    table = solver.autotune(ds, offsets, theta, eps, reps=3)
    tune = {tune_name(r): r["ms"] for r in table}
    best = tune_name(min(table, key=lambda r: r["ms"]))
    log(f"autotune -> {best}")
    solver.graph_capture_step(ds, offsets, theta, eps, dt)
    solver.graph_launch(max(1, warmup))
    solver.sync()
    synchronize()
    t0 = time.perf_counter()
    solver.graph_launch(steps)
    solver.sync()
    synchronize()
    ms_graph = (time.perf_counter() - t0) * 1e3 / steps
    st = solver.stats()
    # the same KDK step as 256 separate per-system calls (one solver reused)
    sep = Solver(theta=theta, walk="fp32")
    ds2 = DeviceState.from_host(state)
    views = [ds2.view(int(offsets[b]), int(sizes[b])) for b in range(len(sizes))]
    prm = [Params(softening=eps[b], dt=dt) for b in range(len(sizes))]
    for b in range(len(sizes)):
        sep.step(views[b], prm[b], acc_valid=False)
    nsep = max(2, steps // 4)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(nsep):
        for b in range(len(sizes)):
            sep.step(views[b], prm[b], acc_valid=True)
    synchronize()
    ms_sep = (time.perf_counter() - t0) * 1e3 / nsep
    summary = {
        "systems": len(sizes), "bodies": n, "theta": theta, "eps": "0.01 N^(-1/3) per system",
        "sizes": "log-uniform [1024, 65536], seed 4",
        "step": "KDK (kick, drift, force step, kick), hipGraph replay", "walk_group": best,
        "ms_per_step": ms_graph, "steps_per_s": 1e3 / ms_graph, "steps_timed": steps,
        "separate_launches": {"ms_per_step": ms_sep, "steps_per_s": 1e3 / ms_sep, "steps_timed": nsep},
        "batched_over_separate": ms_sep / ms_graph,
        "autotune_ms_per_walk": tune,
        "interactions_per_particle": {"pp": st["pp_interactions"] / n, "pc": st["pc_interactions"] / n}}
    return n, ms_graph, summary


def bench_batched(args):
    """The batched workload on its own (see batched_measure)."""
    n, ms_graph, summary = batched_measure(args.steps, args.warmup)
    out = base_line(args, n, ms_graph, "fp32", {
        "workload": "batched-256", "systems": summary["systems"], "n": n, "theta": summary["theta"],
        "eps": summary["eps"], "sizes": summary["sizes"], "step": summary["step"],
        "walk_group": summary["walk_group"]})
    out["steps_per_s"] = summary["steps_per_s"]
    out["separate_launches"] = summary["separate_launches"]
    out["batched_over_separate"] = summary["batched_over_separate"]
    out["autotune_ms_per_walk"] = summary["autotune_ms_per_walk"]
    out["interactions_per_particle"] = summary["interactions_per_particle"]
    return out


def bench_encounter(args):
    """Two-disk encounter: two disk+halo galaxies of 2,097,152 bodies each on a
    parabolic orbit (pericentre 5 R_d, 30 deg inclination), theta 0.5, eps
    0.02, FP32 walk with FP16-stored quadrupoles.  Each global step T runs
    block time steps on 4 power-of-two rungs (8 ticks): every tick drifts all
    bodies and evaluates only those whose step ends there.  T is 4 x the 3 %
    quantile of the initial wanted steps sqrt(2 eta eps / |a|), so 3 % of the
    stars start on the top rung (the odd ticks evaluate about that share).
    Energy (kinetic
    + FP64 tree potential) and momentum are measured in FP64 on the GPU before
    and after the timed steps."""
    from gravitas_amd import DeviceState, Params, Solver, ics
    from gravitas_amd.hip import synchronize
    n = args.n or 4194304
    theta = 0.5 if args.theta is None else args.theta
    eps = 0.02 if args.eps is None else args.eps
    levels, eta = 4, 0.025
    log(f"building two-disk encounter n={n}")
    state = ics.two_disk_encounter(n, seed=3)
    ds = DeviceState.from_host(state)
    params = Params(softening=eps)
    solver = Solver(theta=theta, walk="fp32-q16")
    solver.accelerations(ds, params)  # initial forces
    acc0 = ds.acc_host()
    # full force evaluations at 4M (the rate compared with the CPU path)
    full = {}
    for walk in ("fp32-q16", "fp32"):
        sf = Solver(theta=theta, walk=walk)
        sf.accelerations(ds, params)
        synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            sf.accelerations(ds, params)
        synchronize()
        full[walk] = n / ((time.perf_counter() - t0) / 3)
    solver.accelerations(ds, params)
    a = np.linalg.norm(acc0, axis=1)
    dt_wanted = np.sqrt(2 * eta * eps / a[a > 0])
    # a star goes to the top rung (step T/8) when its wanted step is below T/4:
    # T = 4 x the 3 % quantile puts 3 % of the stars there at the start
    T = float(4 * np.quantile(dt_wanted, ENCOUNTER_TOP)) if args.T is None else args.T
    out_err = {}
    if not args.no_errors:
        log("FP64 walk and direct-sum sample of the initial forces")
        d64 = DeviceState.from_host(state)
        Solver(theta=theta, walk="fp64").accelerations(d64, params)
        acc64 = d64.acc_host()
        targets = ics.sample_targets(n, 4096, 7)
        direct = solver.direct_sample(ds, params, targets).numpy()
        out_err = {"err_fp32q16_vs_fp64_walk": errors(acc0, acc64),
                   "err_fp32q16_vs_direct": errors(acc0[targets], direct)}
        del d64
    diag0 = solver.diagnostics(ds, params)
    u0 = solver.potential_energy(ds, params)
    e0 = diag0["kinetic"] + u0
    for _ in range(args.warmup):  # warm-up steps advance the state; energy is taken after them too
        solver.block_step(ds, params, T, levels, eta, acc_valid=True)
    diag_w = solver.diagnostics(ds, params)
    e_w = diag_w["kinetic"] + solver.potential_energy(ds, params)
    solver.accelerations(ds, params)  # forces of the current positions (the potential call rebuilt the tree)
    synchronize()
    log(f"timing {args.steps} global steps of T={T} on {levels} rungs")
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos.append(solver.block_step(ds, params, T, levels, eta, acc_valid=True))
    synchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / args.steps
    diag1 = solver.diagnostics(ds, params)
    u1 = solver.potential_energy(ds, params)
    e1 = diag1["kinetic"] + u1
    ticks = 1 << (levels - 1)
    act = np.zeros(ticks)
    for inf in infos:
        for t, c in zip(inf["tick"], inf["active"]):
            act[t - 1] += c
    act /= len(infos) * n
    targets_total = sum(inf["targets"] for inf in infos)
    rungs = np.mean([inf["rung_count"] for inf in infos], axis=0) / n
    out = base_line(args, n, ms, "fp32", {
        "workload": "two-disk-4m", "n": n, "theta": theta, "eps": eps, "levels": levels, "eta": eta, "T": T,
        "bodies": "two disk+halo galaxies (2 x 2,097,152), parabolic orbit, pericentre 5 R_d, inclination 30 deg",
        "walk": "fp32 group walk, FP16-stored quadrupoles, Kahan-compensated sums",
        "step": "one global step T = 8 ticks of block steps; active stars evaluated per tick"})
    out["unit"] = "particle-steps/s"
    out["force_evaluations_per_s"] = targets_total / wall
    out["evaluations_per_global_step"] = targets_total / len(infos) / n
    out["active_fraction_per_tick"] = act.tolist()
    out["rung_fraction"] = rungs.tolist()
    out["wanted_step_quantiles"] = {q: float(np.quantile(dt_wanted, q)) for q in (0.01, 0.03, 0.1, 0.5)}
    out["energy"] = {"E_start": e0, "E_after_warmup": e_w, "E_end": e1,
                     "drift_over_timed_steps": (e1 - e_w) / abs(e_w), "drift_total": (e1 - e0) / abs(e0),
                     "global_steps": args.steps + args.warmup, "time_span": (args.steps + args.warmup) * T}
    p0, p1 = np.array(diag0["momentum"]), np.array(diag1["momentum"])
    out["momentum_change"] = float(np.linalg.norm(p1 - p0))
    out["full_force_particles_per_s"] = full
    out.update(out_err)
    if not args.no_cpu:
        log("CPU baseline")
        cb = cpu_baseline(state, theta, eps, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["full_force_gpu_over_cpu"] = full["fp32-q16"] / cb["value"]
    return out


def bench_holes(args):
    """The disk+halo model (1,048,576 bodies) with a central black hole of
    0.1 disk masses and two of 0.02 in the disk plane at R = 2 and 4 on
    circular orbits; steps of dt 2e-3 with the split scheme of
    internal/sim/split.go (tree forces between stars, direct sums for every
    pair with a hole, power-of-two substeps for the holes and the stars near
    them).  Reports steps/s, the substep counts and the FP64 momentum drift.

    disk-halo-1m-holes-retarded: the same with c = 10 and retarded gravity of
    the holes extrapolated by their velocity (retarded.go, "retarded+v"):
    every hole pull comes from the hole's position on the body's past light
    cone, read from its history of outer steps.  The warm-up steps fill the
    histories (the timed steps read up to --warmup + --steps entries; older
    retarded times extrapolate from the oldest one)."""
    from gravitas_amd import DeviceState, Params, Retard, Solver, ics
    from gravitas_amd.hip import synchronize
    n = args.n or 1048576
    theta = 0.5 if args.theta is None else args.theta
    eps = 0.02 if args.eps is None else args.eps
    dt = 2e-3
    retarded = args.config == "disk-halo-1m-holes-retarded"
    c, gravity = (10.0, "retarded+v") if retarded else (0.0, "instant")
    hist = Retard() if retarded else None
    clock = [0.0]

    def split(**kw):
        sub = solver.split_step(ds, params, retard=hist, time=clock[0], gravity=gravity, **kw)
        clock[0] += dt
        return sub
    log(f"building disk+halo with holes n={n}")
    state = ics.disk_halo(n, seed=2, eps=eps)
    state["alive"] = np.ones(n, np.uint8)
    holes = [0, 1, 2]
    state["kind"][holes] = ics.KIND_HOLE
    state["m"][holes] = [0.1, 0.02, 0.02]
    for h, R in zip(holes, (0.0, 2.0, 4.0)):
        state["px"][h], state["py"][h], state["pz"][h] = R, 0.0, 0.0
        state["vx"][h] = state["vz"][h] = 0.0
        state["vy"][h] = 0.0
    r = np.sqrt(state["px"] ** 2 + state["py"] ** 2 + state["pz"] ** 2)
    for hh, R in zip(holes[1:], (2.0, 4.0)):  # circular speed from the mass inside R (spherical estimate)
        state["vy"][hh] = float(np.sqrt(state["m"][r < R].sum() / R))
    ds = DeviceState.from_host(state)
    solver = Solver(theta=theta, walk="fp32")
    params = Params(softening=eps, dt=dt, c=c)
    d0 = solver.diagnostics(ds, params)
    subs = [split(acc_valid=False)]
    for _ in range(max(0, args.warmup - 1)):
        subs.append(split())
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        subs.append(split())
    synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    d1 = solver.diagnostics(ds, params)
    gross = float(np.abs(state["m"] * np.abs(state["vx"])).sum())
    cfg = {"workload": args.config, "n": n, "theta": theta, "eps": eps, "dt": dt,
           "holes": "0.1 at the centre, 0.02 at R = 2 and R = 4 (circular)",
           "step": "split step: tree forces + hole direct sums + substeps near holes"}
    if retarded:
        cfg.update({"c": c, "gravity": gravity, "history_entries": len(subs)})
    out = base_line(args, n, ms, "fp32", cfg)
    out["substeps"] = subs[-args.steps:]
    out["momentum_drift_over_gross"] = float(np.abs(np.array(d1["momentum"]) - np.array(d0["momentum"])).max() / gross)
    return out


def bench_sph(args):
    """SPH hydro (sph.Solver.Forces): a Plummer sphere of 1,048,576 gas bodies
    (u 0.2, velocities x 0.3, internal/sph/system_test.go:14-23), softening
    0.01, default gas parameters (gamma 5/3, 40 neighbours, alpha 1).  h starts
    at the target of the sphere's analytic density and converges in the
    warm-up; one step = one full evaluation (gas list, keys, sort, tree,
    density passes, refit, leaf sort, forces), FP64, bit-identical to the
    oracle."""
    from gravitas_amd import GasState, Hydro, ics
    from gravitas_amd.hip import synchronize
    from gravitas_amd.hydro import HydroOut
    n = args.n or 1048576
    eps = 0.01 if args.eps is None else args.eps
    log(f"building gas sphere n={n}")
    state = ics.gas_sphere(n, 9, h_from_density=True)
    g = GasState.from_host(state)
    hy = Hydro(timing=False)
    out = HydroOut(n)
    for _ in range(max(3, args.warmup)):
        hy.forces(g, eps, out=out)
    synchronize()
    log(f"timing {args.steps} evaluations")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hy.forces(g, eps, out=out)
    synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    st = hy.stats()
    hy_t = Hydro(timing=True)
    phases = []
    for _ in range(5):
        hy_t.forces(g, eps, out=out)
        phases.append(hy_t.phase_ms())
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    res = base_line(args, n, ms, "fp64", {
        "workload": "sph-sphere-1m", "n": n, "eps": eps,
        "bodies": "Plummer sphere of gas (M = a = 1), u 0.2, velocities x 0.3; gamma 5/3, 40 neighbours, alpha 1",
        "parallelism": "single GPU"})
    res["metric"] = "gas particles/s (SPH hydro forces, FP64, bit-identical to the reference order)"
    force_flops = FLOPS_SPH_PAIR * st["pairs"] + FLOPS_SPH_VISC * st["visc_pairs"]
    achieved = force_flops / (ph["force"] * 1e-3) / 1e12
    res.update({
        "neighbours_per_body": st["pairs"] / n, "density_passes_per_body": st["density_sums"] / n,
        "pairs_per_s": st["pairs"] / (ms * 1e-3),
        "tree": {"nodes": st["n_nodes"], "leaves": st["n_leaves"], "depth": st["depth"],
                 "max_density_passes": st["max_iterations"], "walk_steps_per_wave": st["walk_steps"] / ((n + 63) // 64)},
        "phases_ms": ph,
        "roofline": {"bound": "valu", "kernel": "k_hy_force", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": force_flops, "force_ms": ph["force"],
                     "density_tflops": FLOPS_SPH_DENSITY * st["density_pairs"] / (ph["density"] * 1e-3) / 1e12,
                     "note": "FP64 vector ALU bound (divisions and square roots per pair); peak = MI355X FP64 "
                             "vector 78.6 TFLOP/s"},
    })
    if not args.no_cpu:
        import oracle
        log("CPU baseline: one oracle evaluation from the converged state")
        hs = g.host()  # the input of one more GPU evaluation, then of the oracle's
        hy.forces(g, eps, out=out)
        go, gs = out.host(), g.host()
        s = {k: hs[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz", "m", "u", "h", "rho", "kind", "alive")}
        t0 = time.perf_counter()
        oo, _ = oracle.sph_forces(s, softening=eps)
        tc = time.perf_counter() - t0
        res["cpu_baseline"] = {
            "value": n / tc, "unit": "gas particles/s", "cores": 1, "kind": "port",
            "sample": f"the whole {n}-body workload, one Forces evaluation from the GPU-converged h and rho; "
                      "C restatement of sph.Solver.Forces, -O2 -ffp-contract=off, 1 thread"}
        res["gpu_over_cpu"] = res["value"] / res["cpu_baseline"]["value"]
        res["bit_identical_to_oracle"] = bool(all(np.array_equal(go[k], oo[k]) for k in
                                                  ("ax", "ay", "az", "dudt", "divv", "dt")) and
                                              np.array_equal(gs["h"], s["h"]) and np.array_equal(gs["rho"], s["rho"]))
    return res


def bench_spin(args):
    """The star cluster and spinning hole of the reference's accretion scene
    (internal/scene/scenes/11-accretion.json, without its gas cloud): a
    Plummer cluster of 1,048,576 stars (mass 1, scale 1.5), a central hole of
    mass 0.1 spinning at a* = 0.5 about z (J = a* M^2 / c, scene/spec.go:
    241-251), c = 3, softening 0.03, dt 0.002.  One step = the split step
    (split.go:164-199) with frame dragging (framedrag.go): dragOuter over the
    far stars after the first and before the last outer kick, dragFast around
    the drifts of every substep; FP32 tree walk between stars, FP64 for every
    pair with the hole and for the dragging.  Also timed: the same steps
    without the spin, one dragOuter pass over all stars (bit-identical to the
    sequential loop) and, on the CPU, the oracle's pass."""
    from gravitas_amd import DeviceState, Params, Solver, ics
    from gravitas_amd.hip import synchronize
    n = args.n or 1048576
    theta = 0.5 if args.theta is None else args.theta
    eps = 0.03 if args.eps is None else args.eps
    dt, c, hole_m, a_star = 0.002, 3.0, 0.1, 0.5
    log(f"building spinning-hole cluster n={n}")
    state = ics.plummer(n, 7, mass=1.0, a=1.5)
    state["alive"] = np.ones(n, np.uint8)
    state["kind"][0] = ics.KIND_HOLE
    state["m"][0] = hole_m
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        state[k][0] = 0.0
    for k in ("jx", "jy", "jz"):
        state[k] = np.zeros(n)
    state["jz"][0] = a_star * hole_m * hole_m / c

    def run(st, steps, warmup):
        ds = DeviceState.from_host(st)
        solver = Solver(theta=theta, walk="fp32")
        params = Params(softening=eps, dt=dt, c=c)
        subs = [solver.split_step(ds, params, acc_valid=False)]
        for _ in range(max(0, warmup - 1)):
            subs.append(solver.split_step(ds, params))
        synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            subs.append(solver.split_step(ds, params))
        synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, subs[-steps:], solver, ds

    log(f"timing {args.steps} split steps with frame dragging")
    ms, subs, solver, ds = run(state, args.steps, args.warmup)
    drag = solver.drag_stats()
    nospin = {k: v for k, v in state.items() if k not in ("jx", "jy", "jz")}
    ms0, subs0, _, _ = run(nospin, args.steps, args.warmup)
    out = base_line(args, n, ms, "fp64+fp32", {
        "workload": "spin-cluster-1m", "n": n, "theta": theta, "eps": eps, "dt": dt, "c": c,
        "bodies": "Plummer cluster (M 1, a 1.5) with a central hole of 0.1 spinning at a* = 0.5; "
                 "the accretion scene's stars and hole, without the gas",
        "step": "split step with frame dragging: tree forces (FP32 walk) + hole direct sums + substeps + "
                "dragOuter/dragFast (FP64)"})
    out["metric"] = "particles/s (split step with a spinning hole and frame dragging, c = 3)"
    out["substeps"] = subs
    out["ms_per_step_without_spin"] = ms0
    out["substeps_without_spin"] = subs0
    steps_run = args.steps + max(1, args.warmup)
    out["drag_per_step"] = {"passes": drag["passes"] / steps_run, "iterations": drag["iterations"] / steps_run,
                            "sequential_passes": drag["sequential_passes"],
                            "slow_segments": drag["slow_segments"] / steps_run}
    # one dragOuter pass over every star, from the state the steps left
    hs = solver.host_state(ds)
    for k in ("m", "kind", "alive"):
        hs[k] = state[k]
    p = Params(softening=eps, c=c)
    sv = Solver(walk="fp64")
    dsp = DeviceState.from_host(hs)
    reps = 10
    sv.drag_outer(dsp, p, 0.5 * dt)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sv.drag_outer(dsp, p, 0.5 * dt)
    synchronize()
    pass_ms = (time.perf_counter() - t0) * 1e3 / reps
    out["drag_outer_pass"] = {"bodies": n - 1, "ms": pass_ms, "passes_per_s": 1e3 / pass_ms,
                              "stats": sv.drag_stats()}
    if not args.no_cpu:
        import oracle
        keys = ("px", "py", "pz", "vx", "vy", "vz", "jx", "jy", "jz")
        one = DeviceState.from_host(hs)
        sv.drag_outer(one, p, 0.5 * dt)
        g = sv.host_state(one)
        w = {k: hs[k].copy() for k in keys}
        log("CPU baseline: one oracle dragOuter pass")
        t0 = time.perf_counter()
        oracle.drag_outer(w, hs["m"], hs["kind"], hs["alive"], 0.5 * dt, eps * eps, c)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": 1.0 / tc, "unit": "dragOuter passes/s", "cores": 1, "kind": "port",
            "sample": f"one dragOuter pass over the {n - 1} stars (the whole workload), C restatement of "
                      "framedrag.go's sequential loop, -O2 -ffp-contract=off, 1 thread"}
        out["drag_gpu_over_cpu"] = (1e3 / pass_ms) * tc
        out["drag_bit_identical_to_oracle"] = bool(all(np.array_equal(g[k], w[k]) for k in keys[3:]))
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    from gravitas_amd import gpu_available
    reason = gpu_available()
    if reason is not None:
        print(f"no GPU backend: {reason}", file=sys.stderr)
        sys.exit(1)
    fn = {"disk-halo-1m": bench_disk_halo, "plummer-16k": bench_plummer, "clumps-16m": bench_clumps,
          "batched-256": bench_batched, "two-disk-4m": bench_encounter,
          "disk-halo-1m-holes": bench_holes, "disk-halo-1m-holes-retarded": bench_holes, "sph-sphere-1m": bench_sph,
          "spin-cluster-1m": bench_spin}[args.config]
    print(json.dumps(fn(args)), flush=True)


if __name__ == "__main__":
    main()
