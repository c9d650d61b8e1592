"""The FP32 walk's tuning space and the library's autotuner (no counterpart in
the reference, whose Metal walk has one fixed shape): every candidate --
group size 64 or 32, work queue never or always, each instantiated stack and
staging size -- walks the FP64 walk's interaction set (bh.go:106-155) within
the reference's GPU bound (gpu_test.go:18, plus p99), and the autotuner times
them all, keeps the fastest and caches the choice per workload shape."""
import sys

import numpy as np
import pytest

from gravitas_amd import DeviceState, Params, Solver, ics

pytestmark = pytest.mark.gpu

# every built walk shape: (traversal stack, staging, wavefronts per SIMD)
SHAPES = ((384, 128, 4), (384, 256, 4), (384, 384, 4), (768, 256, 4), (768, 384, 4), (384, 128, 5), (384, 256, 5),
          (384, 384, 5), (384, 256, 6))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def workload():
    # a single disk+halo galaxy and, as a batch, three systems of uneven size
    s = ics.disk_halo(98304, seed=31, eps=0.02)
    systems, eps = ics.batched_systems(3, 2048, 30000, seed=8)
    b = ics.concat(*systems)
    off = np.r_[0, np.cumsum([len(x["px"]) for x in systems])].astype(np.int64)
    return s, b, off, eps


def reference(state, offsets, theta, eps):
    ds = DeviceState.from_host(state)
    s64 = Solver(walk="fp64")
    s64.batch_accelerations(ds, offsets, theta, eps)
    st = s64.stats()
    return ds.acc_host(), (st["pp_interactions"], st["pc_interactions"])


def check(acc, ref):
    e = np.sort(ics.rel_errors(acc, ref))
    p50, p99, mx = e[len(e) // 2], e[int(0.99 * len(e))], e[-1]
    assert p50 <= 1e-5 and p99 <= 1e-4 and mx <= 1e-3, (p50, p99, mx)


def test_every_candidate_walks_the_fp64_interaction_set(gpu, workload):
    s, b, off, eps = workload
    for name, state, offsets, theta, ep in (("disk+halo", s, [0, len(s["px"])], 0.5, 0.02),
                                            ("batch", b, off, 0.6, eps)):
        ref, counts = reference(state, offsets, theta, ep)
        results = {}
        for gsize in (64, 32):
            for pers in (0, 2):
                for stack, parts, waves in SHAPES:
                    say(f"[tune] {name} group {gsize} queue {pers} stack {stack} staging {parts} waves {waves}")
                    ds = DeviceState.from_host(state)
                    sv = Solver(walk="fp32", group_size=gsize, persistent=pers, stack=stack, parts=parts,
                                waves=waves)
                    sv.batch_accelerations(ds, offsets, theta, ep)
                    st = sv.stats()
                    assert (st["pp_interactions"], st["pc_interactions"]) == counts
                    assert st["fallbacks"] == 0
                    acc = ds.acc_host()
                    check(acc, ref)
                    results[(gsize, pers, stack, parts, waves)] = acc
        # the stack size, the register budget and the work queue do not change
        # a group's arithmetic (the staging size does: it sets the flushes)
        first = {}
        for stack, parts, waves in SHAPES:
            first.setdefault(parts, (stack, waves))
        for (gsize, pers, stack, parts, waves), acc in results.items():
            assert np.array_equal(acc, results[(gsize, 0, first[parts][0], parts, first[parts][1])])


def test_bad_walk_sizes_are_argument_errors(gpu):
    from gravitas_amd import _lib
    sv = Solver(walk="fp32", stack=512, parts=256)
    with pytest.raises(_lib.GxError) as e:
        sv.accelerations(DeviceState.from_host(ics.plummer(1000, 3)), Params(softening=0.01))
    assert e.value.code == _lib.GX_ERR_ARG


def test_autotuner_times_all_and_caches(gpu, workload):
    s, _, _, _ = workload
    n = len(s["px"])
    ref, counts = reference(s, [0, n], 0.5, 0.02)
    ds = DeviceState.from_host(s)
    sv = Solver(theta=0.5, walk="fp32")
    say("[tune] autotune disk+halo")
    table = sv.autotune(ds, [0, n], 0.5, 0.02, reps=2)
    assert len(table) == 2 * 2 * len(SHAPES)
    assert all((r["pp_interactions"], r["pc_interactions"]) == counts for r in table)
    assert all(0 < r["ms"] < 1e3 for r in table)
    best = min(table, key=lambda r: r["ms"])
    assert (sv.group_size, sv.persistent, sv.stack, sv.parts, sv.waves) == (
        best["group_size"], best["persistent"], best["stack"], best["parts"], best["waves"])
    check(ds.acc_host(), ref)  # the chosen walk's forces
    # the same shape again: the cached choice, the same table, no new timing
    again = sv.autotune(ds, [0, n], 0.5, 0.02, reps=2)
    assert again == table
    # the tuned solver's steps are as accurate
    sv.accelerations(ds, Params(softening=0.02))
    check(ds.acc_host(), ref)
