"""GPU frame dragging (gravitas_amd/csrc/gx_drag.hip; internal/sim/framedrag.go)
against the oracle's sequential restatement.

  * dragOuter and dragFast passes: velocities of bodies and holes and the
    holes' spins bit-identical, from 50 bodies (strong dragging) to 1M bodies
    (1,024 scan segments per spin component), with one and two spinning
    holes, dead bodies, near flags and an unsorted near list;
  * the same through the sequential fallback (no fixed-point iteration, and a
    fixed point stopped after one iteration), which must leave the state
    untouched before it runs, for a list short enough for the one-workgroup
    pass and for a longer one;
  * whole split steps with spinning holes: as the other split-step tests,
    1e-12 of the oracle (the holes' sums over bodies are tree reductions on
    the GPU), spins included, and the same substep counts.
"""
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import DeviceState, Params, Solver, ics
from gravitas_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu

SPIN = ("jx", "jy", "jz")
KEYS = ("vx", "vy", "vz") + SPIN
EPS, C = 0.01, 3.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def spin_state(n, seed, two=False, strong=False):
    s = ics.plummer(n, seed)
    s["alive"] = np.ones(n, np.uint8)
    for k in SPIN:
        s[k] = np.zeros(n)
    holes = [0, 5, 9] if n > 9 else [0, 1]
    s["kind"][holes] = ics.KIND_HOLE
    s["m"][holes] = 0.05 if not strong else 2.0
    s["jz"][holes[0]] = 0.004 if not strong else 3.0
    if two:
        s["jx"][holes[1]], s["jy"][holes[1]] = 0.0005, -0.0002
        if strong:
            s["jx"][holes[1]], s["jy"][holes[1]] = 1.5, -0.5
    if n > 40:
        s["alive"][[12, 40]] = 0
    return s


def oracle_pass(s, dt, near=None, near_list=None):
    want = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz") + SPIN}
    if near_list is None:
        oracle.drag_outer(want, s["m"], s["kind"], s["alive"], dt, EPS * EPS, C, near=near)
    else:
        oracle.drag_fast(want, s["m"], s["kind"], s["alive"], near_list, dt, EPS * EPS, C)
    return want


def gpu_pass(sv, s, dt, near=None, near_list=None):
    ds = DeviceState.from_host(s)
    p = Params(softening=EPS, c=C)
    if near_list is None:
        sv.drag_outer(ds, p, dt, near=None if near is None else DeviceArray.from_numpy(near))
    else:
        sv.drag_fast(ds, p, dt, DeviceArray.from_numpy(np.asarray(near_list, np.int32)))
    return sv.host_state(ds)


def assert_same(got, want, start):
    for k in KEYS:
        assert np.array_equal(got[k], want[k]), (k, np.flatnonzero(got[k] != want[k])[:5])
    assert not np.array_equal(want["jz"], start["jz"])  # the pass did move the spins


CASES = [(50, 1.0, False, True), (50, 0.2, True, True), (3000, 0.01, False, False), (3000, 0.01, True, False),
         (300_000, 0.01, True, False), (1_000_000, 0.01, False, False)]


@pytest.mark.parametrize("n,dt,two,strong", CASES)
def test_drag_outer_bit_identical(gpu, n, dt, two, strong):
    s = spin_state(n, 61, two=two, strong=strong)
    near = (np.random.default_rng(5).random(n) < 0.05).astype(np.uint8)
    say(f"[drag] outer n={n} dt={dt} two={two} strong={strong}")
    want = oracle_pass(s, dt, near=near)
    sv = Solver(walk="fp64")
    got = gpu_pass(sv, s, dt, near=near)
    assert_same(got, want, s)
    st = sv.drag_stats()
    say(f"[drag] stats {st}")
    assert st["passes"] == 1
    if not strong:
        assert st["sequential_passes"] == 0 and st["max_iterations"] <= 6, st


@pytest.mark.parametrize("n,dt,two,strong", [c for c in CASES if c[0] <= 300_000])
def test_drag_fast_bit_identical(gpu, n, dt, two, strong):
    s = spin_state(n, 62, two=two, strong=strong)
    rng = np.random.default_rng(6)
    cand = np.flatnonzero(s["kind"] != ics.KIND_HOLE)
    near_list = rng.permutation(cand)[: max(3, n // 20)].astype(np.int32)  # any order, dead bodies included
    say(f"[drag] fast n={n} near={len(near_list)} two={two} strong={strong}")
    want = oracle_pass(s, dt, near_list=near_list)
    got = gpu_pass(Solver(walk="fp64"), s, dt, near_list=near_list)
    assert_same(got, want, s)


@pytest.mark.parametrize("n", [2000, 20000])  # one workgroup / host-driven passes
@pytest.mark.parametrize("iterations", ["0", "1"])
def test_sequential_fallback_bit_identical(gpu, monkeypatch, iterations, n):
    monkeypatch.setenv("GX_DRAG_ITERATIONS", iterations)
    s = spin_state(n, 63, two=True)
    near = np.zeros(n, np.uint8)
    say(f"[drag] n={n} sequential fallback after {iterations} iterations")
    want = oracle_pass(s, 0.01, near=near)
    sv = Solver(walk="fp64")
    got = gpu_pass(sv, s, 0.01, near=near)
    assert_same(got, want, s)
    st = sv.drag_stats()
    assert st["sequential_passes"] == 1 and st["iterations"] == int(iterations), st


def test_no_spin_or_newtonian_leaves_the_state(gpu):
    s = spin_state(3000, 64)
    s["jz"][:] = 0
    sv = Solver(walk="fp64")
    got = gpu_pass(sv, s, 0.01)
    for k in KEYS:
        assert np.array_equal(got[k], s[k]), k
    s = spin_state(3000, 64)
    ds = DeviceState.from_host(s)
    sv.drag_outer(ds, Params(softening=EPS, c=0.0), 0.01)
    got = sv.host_state(ds)
    for k in KEYS:
        assert np.array_equal(got[k], s[k]), k


def hole_spin_state(n=5000, seed=41):
    s = ics.plummer(n, seed)
    s["alive"] = np.ones(n, np.uint8)
    s["kind"][[0, 1, 2]] = ics.KIND_HOLE
    s["m"][[0, 1, 2]] = [0.05, 0.03, 0.02]
    s["alive"][[10, 20]] = 0
    for k in SPIN:
        s[k] = np.zeros(n)
    s["jz"][0] = 0.5 * 0.05 ** 2 / 40.0  # a* = 0.5: J = a* M^2 / c (scene/spec.go:241-251)
    s["jx"][1], s["jz"][1] = 0.2 * 0.03 ** 2 / 40.0, 0.3 * 0.03 ** 2 / 40.0
    return s


def test_split_steps_with_spinning_holes_match_oracle(gpu):
    s = hole_spin_state()
    dt, c = 1e-2, 40.0
    o = oracle.Solver(theta=0.5, workers=8)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz") + SPIN}
    ds = DeviceState.from_host(s)
    sv = Solver(theta=0.5, walk="fp64")
    p = Params(softening=EPS, dt=dt, c=c)
    for rep in range(3):
        say(f"[drag] split step {rep}")
        want = oracle.split_step(o, st, s["m"], s["kind"], s["alive"], EPS, dt, c=c, acc_valid=rep > 0)
        got = sv.split_step(ds, p, acc_valid=rep > 0)
        assert got == want and got > 1, (got, want)
        h = sv.host_state(ds)
        for k in ("px", "py", "pz", "vx", "vy", "vz"):
            scale = max(1.0, np.abs(st[k]).max())
            assert np.abs(h[k] - st[k]).max() <= 1e-12 * scale, (k, np.abs(h[k] - st[k]).max())
        for k in SPIN:
            assert np.abs(h[k] - st[k]).max() <= 1e-12 * np.abs(st["jz"]).max(), k
    assert not np.array_equal(st["jz"], s["jz"])
    dstats = sv.drag_stats()
    say(f"[drag] split-step stats {dstats}")
    assert dstats["passes"] > 6 and dstats["sequential_passes"] == 0
