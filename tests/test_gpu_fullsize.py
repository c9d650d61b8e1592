"""Parity at the benchmark sizes (the smaller tests cover the edge cases):

  * gravity, the 1,048,576-body disk+halo model at eps 0.02 and theta 0.5
    and 0.7 -- the FP64 walk bit-identical to the oracle, the FP32 walk
    within the reference's GPU bound (gpu_test.go:18) and p99 <= 1e-4, both
    with the oracle's interaction counts;
  * 256 batched systems (N log-uniform in [1,024, 65,536], theta 0.6, eps
    0.01 N^(-1/3)) -- the FP64 batch bit-identical to 256 separate oracle
    runs, the FP32 batch (one workgroup per group and the work queue) within
    the bound with the FP64 counts, three graph replays of the KDK step equal
    to eager steps;
  * 16,777,216 stars in 1,000 clumps (theta 0.5, eps 1e-4) -- keys, order and
    node count equal to the oracle's build, the FP64 walk bit-identical to
    the oracle's walk on a seeded 4,096-target sample, FP32 and FP32 with
    FP16-stored quadrupoles against the FP64 walk over every star;
  * the two-disk encounter of 4,194,304 bodies -- one FP64 block step on 4
    rungs equal to the oracle's event loop (blockstep.go:130-198) driven by
    the same forces (rung populations, event ticks, active counts, state),
    FP32-Q16 active-set forces against the FP64 walk, and the FP64 energy
    drift of three FP32-Q16 global steps;
  * hydro: a 262,144-body gas sphere (h from the analytic density, then
    converged) -- every output bit-identical to the oracle.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import DeviceState, GasState, Hydro, Params, Solver, ics
from gravitas_amd.hydro import HydroOut

pytestmark = pytest.mark.gpu
WORKERS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def disk_halo_1m():
    return ics.disk_halo(1048576, seed=2, eps=0.02)


@pytest.mark.parametrize("theta", [0.5, 0.7])
def test_disk_halo_1m_walks(gpu, disk_halo_1m, theta):
    s = disk_halo_1m
    n = len(s["px"])
    say(f"[full] disk+halo {n} theta {theta}: oracle on {WORKERS} threads")
    want = oracle.Solver(theta=theta, workers=WORKERS).accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"],
                                                                     eps=0.02)
    ds = DeviceState.from_host(s)
    s64 = Solver(theta=theta, walk="fp64")
    say("[full] FP64 walk")
    s64.accelerations(ds, Params(softening=0.02))
    assert np.array_equal(ds.acc_host(), want)
    st64 = s64.stats()
    say("[full] FP32 walk")
    s32 = Solver(theta=theta, walk="fp32")
    s32.accelerations(ds, Params(softening=0.02))
    st32 = s32.stats()
    assert (st32["pp_interactions"], st32["pc_interactions"]) == (st64["pp_interactions"], st64["pc_interactions"])
    assert st32["fallbacks"] == 0
    e = np.sort(ics.rel_errors(ds.acc_host(), want))
    p50, p99, mx = e[len(e) // 2], e[int(0.99 * len(e))], e[-1]
    say(f"[full] FP32 vs FP64: p50 {p50:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p50 <= 1e-5 and p99 <= 1e-4 and mx <= 1e-3


def walk_errors(got, want):
    e = np.sort(ics.rel_errors(got, want))
    return e[len(e) // 2], e[int(0.99 * len(e))], e[-1]


def test_batched_256_systems(gpu):
    systems, eps = ics.batched_systems(256)
    sizes = np.array([len(x["px"]) for x in systems], dtype=np.int64)
    offsets = np.r_[0, np.cumsum(sizes)].astype(np.int64)
    state = ics.concat(*systems)
    theta, dt = 0.6, 1e-3
    n = int(offsets[-1])
    say(f"[full] batched {len(sizes)} systems, n={n}: oracle per system on {WORKERS} threads")
    want = np.zeros((n, 3))
    for b in range(len(sizes)):
        sl = slice(int(offsets[b]), int(offsets[b + 1]))
        want[sl] = oracle.Solver(theta=theta, workers=WORKERS).accelerations(
            state["px"][sl], state["py"][sl], state["pz"][sl], state["m"][sl], state["kind"][sl], eps=eps[b])
    say("[full] batched FP64")
    ds = DeviceState.from_host(state)
    s64 = Solver(walk="fp64")
    s64.batch_accelerations(ds, offsets, theta, eps)
    assert np.array_equal(ds.acc_host(), want)
    st64 = s64.stats()
    results = {}
    for persistent in (0, 2):
        say(f"[full] batched FP32, persistent {persistent}")
        s32 = Solver(walk="fp32", persistent=persistent)
        s32.batch_accelerations(ds, offsets, theta, eps)
        st = s32.stats()
        assert (st["pp_interactions"], st["pc_interactions"]) == (st64["pp_interactions"], st64["pc_interactions"])
        assert st["fallbacks"] == 0
        results[persistent] = ds.acc_host()
        p50, p99, mx = walk_errors(results[persistent], want)
        say(f"[full] batched FP32 vs FP64: p50 {p50:.2e} p99 {p99:.2e} max {mx:.2e}")
        assert p50 <= 1e-5 and p99 <= 1e-4 and mx <= 1e-3
    assert np.array_equal(results[0], results[2])
    say("[full] batched KDK: 3 graph replays vs eager steps")
    eager = DeviceState.from_host(state)
    se = Solver(walk="fp32")
    se.batch_accelerations(eager, offsets, theta, eps)
    graphed = DeviceState.from_host(state)
    sg = Solver(walk="fp32")
    sg.graph_capture_step(graphed, offsets, theta, eps, dt)
    for rep in range(3):
        se.batch_step(eager, offsets, theta, eps, dt, acc_valid=True)
        sg.graph_launch(1)
        sg.sync()
        ge, gg = se.host_state(eager), sg.host_state(graphed)
        for k in ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay", "az"):
            assert np.array_equal(ge[k], gg[k]), (rep, k)


def test_clumps_16m_ladder(gpu):
    n, theta, eps = 16777216, 0.5, 1e-4
    s = ics.clumps(n, 1000, seed=5)
    say(f"[full] clumps {n}: oracle build")
    o = oracle.Solver(theta=theta, workers=WORKERS)
    o.build(s["px"], s["py"], s["pz"], s["m"], s["kind"])
    ds = DeviceState.from_host(s)
    p = Params(softening=eps)
    say("[full] clumps FP64 walk")
    s64 = Solver(theta=theta, walk="fp64")
    s64.accelerations(ds, p)
    st64 = s64.stats()
    assert st64["n_nodes"] == o.n_nodes
    assert np.array_equal(s64.export("keys"), o.keys())
    order = o.order()
    assert np.array_equal(s64.export("order"), order)
    a64 = ds.acc_host()
    # the oracle's per-particle walk (bh.go:106-155) of a seeded sample on the full tree
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order), dtype=order.dtype)
    targets = ics.sample_targets(n, 4096, 7)
    want = np.array([o.walk(s["px"][i], s["py"][i], s["pz"][i], int(rank[i]), eps * eps) for i in targets])
    assert np.array_equal(a64[targets], want)
    del o
    for walk, bound in (("fp32", (1e-5, 1e-4, 1e-3)), ("fp32-q16", (1e-4, 1e-3, 1e-2))):
        say(f"[full] clumps {walk} walk")
        sv = Solver(theta=theta, walk=walk)
        sv.accelerations(ds, p)
        st = sv.stats()
        assert (st["pp_interactions"], st["pc_interactions"]) == (st64["pp_interactions"], st64["pc_interactions"])
        assert st["fallbacks"] == 0
        p50, p99, mx = walk_errors(ds.acc_host(), a64)
        say(f"[full] clumps {walk} vs FP64: p50 {p50:.2e} p99 {p99:.2e} max {mx:.2e}")
        assert p50 <= bound[0] and p99 <= bound[1] and mx <= bound[2]


def test_two_disk_4m_block_step(gpu):
    n, theta, eps, levels, eta = 4194304, 0.5, 0.02, 4, 0.025
    s = ics.two_disk_encounter(n, seed=3)
    p = Params(softening=eps)
    gpu_state = DeviceState.from_host(s)
    s64 = Solver(theta=theta, walk="fp64")

    def acc_fn(x, y, z):  # FP64 GPU forces of all bodies at the given positions
        for k, v in (("px", x), ("py", y), ("pz", z)):
            getattr(gpu_state, k).upload(np.ascontiguousarray(v))
        s64.accelerations(gpu_state, p)
        return gpu_state.acc_host()

    say(f"[full] two-disk {n}: FP64 forces, oracle walk of a 4,096-target sample")
    a0 = acc_fn(s["px"], s["py"], s["pz"])
    o = oracle.Solver(theta=theta, workers=WORKERS)
    o.build(s["px"], s["py"], s["pz"], s["m"], s["kind"])
    order = o.order()
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order), dtype=order.dtype)
    targets = ics.sample_targets(n, 4096, 7)
    want = np.array([o.walk(s["px"][i], s["py"][i], s["pz"][i], int(rank[i]), eps * eps) for i in targets])
    assert np.array_equal(a0[targets], want)
    del o
    a = np.linalg.norm(a0, axis=1)
    T = 4 * float(np.quantile(np.sqrt(2 * eta * eps / a[a > 0]), 0.03))  # ~3 % of the stars on the top rung
    say(f"[full] FP64 block step T={T:.3e}, GPU and the oracle's event loop on the same forces")
    ds = DeviceState.from_host(s)
    sv = Solver(theta=theta, walk="fp64")
    got = sv.block_step(ds, p, T, levels, eta, acc_valid=False)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    st["acc"] = a0
    ref = oracle.block_step(st, acc_fn, T, levels, eta, eps, s["kind"], np.ones(n, np.uint8))
    assert got["rung_count"] == ref["rung_count"]
    assert list(zip(got["tick"], got["active"])) == ref["events"]
    assert ref["rung_count"][0] > 0 and ref["rung_count"][levels - 1] > 0
    h = sv.host_state(ds)
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        assert np.array_equal(h[k], st[k]), k
    say("[full] FP32-Q16 active-set forces vs FP64")
    rng = np.random.default_rng(9)
    active = (rng.random(n) < 0.05).astype(np.uint8)
    d16 = DeviceState.from_host(s)
    sq = Solver(theta=theta, walk="fp32-q16")
    sq.accelerations_active(d16, p, active)
    sel = active == 1
    p50, p99, mx = walk_errors(d16.acc_host()[sel], a0[sel])
    say(f"[full] Q16 active vs FP64: p50 {p50:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p50 <= 1e-4 and p99 <= 1e-3 and mx <= 1e-2
    say("[full] FP64 energy drift of 3 FP32-Q16 global steps")
    d = DeviceState.from_host(s)
    sq.accelerations(d, p)
    e0 = sq.diagnostics(d, p)["kinetic"] + sq.potential_energy(d, p)
    sq.accelerations(d, p)
    for _ in range(3):
        sq.block_step(d, p, T, levels, eta, acc_valid=True)
    e1 = sq.diagnostics(d, p)["kinetic"] + sq.potential_energy(d, p)
    say(f"[full] energy {e0:.9e} -> {e1:.9e}, drift {abs(e1 - e0) / abs(e0):.2e}")
    assert abs(e1 - e0) <= 1e-3 * abs(e0)


def test_gas_sphere_262k_hydro(gpu):
    n = 262144
    s = ics.gas_sphere(n, 9, h_from_density=True)
    g = GasState.from_host(s)
    hy = Hydro()
    out = HydroOut(n)
    say(f"[full] gas sphere {n}: converge h on the GPU")
    for _ in range(3):
        hy.forces(g, 0.01, out=out)
    start = g.host()
    say("[full] one more evaluation, GPU and oracle")
    hy.forces(g, 0.01, out=out)
    go, gs = out.host(), g.host()
    so = {k: start[k].copy() for k in start}
    oo, evals = oracle.sph_forces(so, softening=0.01)
    for k in ("ax", "ay", "az", "dudt", "divv", "dt"):
        assert np.array_equal(go[k], oo[k]), k
    assert np.array_equal(gs["h"], so["h"]) and np.array_equal(gs["rho"], so["rho"])
    assert hy.stats()["density_sums"] == evals
