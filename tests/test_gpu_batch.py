"""GPU parity of batched steps: independent systems packed in one arena must
each get exactly what a separate force step (and the CPU oracle) gives them,
in one pass; hipGraph replays of a captured KDK step must equal eager steps.
The reference has no batching (one Simulation per engine); its per-system
contract is bh.Solver.Accelerations (internal/sim/bh/bh.go:76-102)."""
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import DeviceState, Params, Solver, ics

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_batch(sizes, seed):
    parts = []
    for i, n in enumerate(sizes):
        if n == 0:
            parts.append({k: np.zeros(0, np.uint8 if k == "kind" else np.float64)
                          for k in ("px", "py", "pz", "vx", "vy", "vz", "m", "kind")})
            continue
        s = ics.plummer(n, seed + i) if i % 2 == 0 else ics.hernquist_halo(n, seed + i, 1.0, 1.0, 20.0,
                                                                            kind=ics.KIND_STAR)
        # shift the systems apart; each one is walked alone anyway
        s["px"] = s["px"] + 10.0 * i
        parts.append(s)
    state = ics.concat(*parts)
    offsets = np.r_[0, np.cumsum(sizes)].astype(np.int64)
    return state, offsets


SIZES = [3000, 1, 0, 1024, 2500, 25, 700, 4096, 64, 1500, 3, 2200]
THETAS = [0.6, 0.5, 0.6, 0.7, 0.6, 0.9, 0.4, 0.6, 0.6, 1.0, 0.6, 0.55]


@pytest.fixture(scope="module")
def batch():
    state, offsets = make_batch(SIZES, 100)
    eps = [0.01 * max(n, 1) ** (-1.0 / 3.0) for n in SIZES]
    # some dead bodies and a hole in the mix
    alive = np.ones(len(state["px"]), np.uint8)
    alive[[5, 4000, 7000]] = 0
    state["alive"] = alive
    state["kind"][[10, 9000]] = ics.KIND_HOLE
    return state, offsets, eps


def oracle_per_instance(state, offsets, eps):
    out = np.zeros((len(state["px"]), 3))
    for b in range(len(SIZES)):
        lo, hi = offsets[b], offsets[b + 1]
        if hi == lo:
            continue
        sl = slice(lo, hi)
        o = oracle.Solver(theta=THETAS[b], workers=8)
        out[sl] = o.accelerations(state["px"][sl], state["py"][sl], state["pz"][sl], state["m"][sl],
                                  state["kind"][sl], state["alive"][sl], eps=eps[b])
    return out


def test_batch_fp64_equals_separate_oracle_runs(gpu, batch):
    state, offsets, eps = batch
    say(f"[batch] fp64 {len(SIZES)} systems, n={offsets[-1]}")
    want = oracle_per_instance(state, offsets, eps)
    ds = DeviceState.from_host(state)
    Solver(walk="fp64").batch_accelerations(ds, offsets, THETAS, eps)
    got = ds.acc_host()
    live = state["alive"] == 1
    assert np.array_equal(got[live], want[live])


def test_batch_fp32_matches_and_variants_agree(gpu, batch):
    state, offsets, eps = batch
    live = (state["alive"] == 1) & (state["kind"] != ics.KIND_HOLE)
    ds64 = DeviceState.from_host(state)
    s64 = Solver(walk="fp64")
    s64.batch_accelerations(ds64, offsets, THETAS, eps)
    ref = ds64.acc_host()
    results = {}
    for gsize in (64, 32):
        for persistent in (0, 2):  # one workgroup per group; the work queue (forced: 2)
            say(f"[batch] fp32 group={gsize} persistent={persistent}")
            ds = DeviceState.from_host(state)
            s = Solver(walk="fp32", group_size=gsize, persistent=persistent)
            s.batch_accelerations(ds, offsets, THETAS, eps)
            st = s.stats()
            assert st["pp_interactions"] == s64.stats()["pp_interactions"]
            assert st["pc_interactions"] == s64.stats()["pc_interactions"]
            acc = ds.acc_host()
            nz = live & (np.abs(ref).sum(1) > 0)  # a one-star system feels no force
            assert np.array_equal(acc[live & ~nz], ref[live & ~nz])
            e = np.sort(ics.rel_errors(acc[nz], ref[nz]))
            assert e[len(e) // 2] <= 1e-5 and e[-1] <= 1e-3
            results[(gsize, persistent)] = acc
    # the work queue does not change the arithmetic of a group
    assert np.array_equal(results[(64, 0)], results[(64, 2)])
    assert np.array_equal(results[(32, 0)], results[(32, 2)])


def test_single_system_batch_equals_plain_call(gpu):
    s = ics.plummer(5000, 7)
    say("[batch] one-system batch vs gx_accelerations_dev")
    a = DeviceState.from_host(s)
    b = DeviceState.from_host(s)
    Solver(theta=0.5, walk="fp32").accelerations(a, Params(softening=0.01))
    Solver(walk="fp32").batch_accelerations(b, [0, 5000], 0.5, 0.01)
    assert np.array_equal(a.acc_host(), b.acc_host())


def test_batch_step_and_graph_replay_equal_eager(gpu, batch):
    state, offsets, eps = batch
    dt = 1e-3
    for walk in ("fp64", "fp32"):
        say(f"[batch] graph replay vs eager KDK, walk={walk}")
        eager = DeviceState.from_host(state)
        se = Solver(walk=walk)
        se.batch_accelerations(eager, offsets, THETAS, eps)
        graphed = DeviceState.from_host(state)
        sg = Solver(walk=walk, persistent=2)
        sg.graph_capture_step(graphed, offsets, THETAS, eps, dt)  # also computes the initial forces
        for rep in range(3):
            se.batch_step(eager, offsets, THETAS, eps, dt, acc_valid=True)
            sg.graph_launch(1)
            sg.sync()
            ge, gg = se.host_state(eager), sg.host_state(graphed)
            for k in ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay", "az"):
                live = state["alive"] == 1
                assert np.array_equal(ge[k][live], gg[k][live]), (rep, k)
    # the FP64 batch step equals the oracle's KDK per system
    say("[batch] fp64 batch KDK vs oracle")
    ds = DeviceState.from_host(state)
    s = Solver(walk="fp64")
    s.batch_step(ds, offsets, THETAS, eps, dt, acc_valid=False)
    got = s.host_state(ds)
    for b in (0, 3, 7):
        lo, hi = offsets[b], offsets[b + 1]
        sl = slice(lo, hi)
        o = oracle.Solver(theta=THETAS[b], workers=8)
        st = {k: state[k][sl].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
        m, kind, alive = state["m"][sl], state["kind"][sl], state["alive"][sl]
        acc = lambda x, y, z: o.accelerations(x, y, z, m, kind, alive, eps=eps[b])
        st["acc"] = acc(st["px"], st["py"], st["pz"])
        oracle.kick_drift_kick(st, acc, dt)
        for k in ("px", "vy"):
            assert np.array_equal(got[k][sl], st[k]), (b, k)


def test_batch_of_more_than_256_systems_fp64_equals_oracle(gpu):
    # more than 256 instances: the sort orders them with two passes over the
    # instance id (one pass up to 256); small systems, some empty
    rng = np.random.default_rng(77)
    sizes = [int(n) for n in rng.integers(0, 300, 300)]
    sizes[5] = sizes[17] = 0
    state, offsets = make_batch(sizes, 500)
    thetas = [0.6] * len(sizes)
    eps = [0.01 * max(n, 1) ** (-1.0 / 3.0) for n in sizes]
    say(f"[batch] fp64 {len(sizes)} systems, n={offsets[-1]}")
    want = np.zeros((len(state["px"]), 3))
    for b, n in enumerate(sizes):
        if n == 0:
            continue
        sl = slice(offsets[b], offsets[b + 1])
        o = oracle.Solver(theta=thetas[b], workers=1)
        want[sl] = o.accelerations(state["px"][sl], state["py"][sl], state["pz"][sl], state["m"][sl],
                                   state["kind"][sl], eps=eps[b])
    ds = DeviceState.from_host(state)
    Solver(walk="fp64").batch_accelerations(ds, offsets, thetas, eps)
    assert np.array_equal(ds.acc_host(), want)
    ds32 = DeviceState.from_host(state)
    Solver(walk="fp32").batch_accelerations(ds32, offsets, thetas, eps)
    pulled = np.linalg.norm(want, axis=1) > 0  # a lone star feels nothing
    e = np.sort(ics.rel_errors(ds32.acc_host()[pulled], want[pulled]))
    assert e[len(e) // 2] <= 1e-5 and e[-1] <= 1e-3
