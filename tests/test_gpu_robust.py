"""Regression tests of the step's error handling on the GPU: overflows that a
captured step's replays meet are reported, a node arena that starts too small
is grown and the step repeated with the same result, a captured step whose
buffers a later eager step has reallocated is refused, and LastGPUSeconds
(gpu_darwin.go:155) is measured with timing off.  Every overflow here is one
the kernels detect and bound (they write nothing past a capacity)."""
import sys

import numpy as np
import pytest

from gravitas_amd import DeviceState, Params, Solver, _lib, ics

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def line_state(n, seed):
    """n stars on a segment of the x axis: the walk lists of a group stay short."""
    rng = np.random.default_rng(seed)
    s = ics.plummer(n, seed)
    s["px"] = np.sort(rng.random(n)) * 10.0
    s["py"] = np.zeros(n)
    s["pz"] = np.zeros(n)
    for k in ("vx", "vy", "vz"):
        s[k] = np.zeros(n)
    return s


def test_graph_replays_report_an_earlier_replays_overflow(gpu):
    """A replay whose lists overflow (a denser state copied into the captured
    arrays) is reported by gx_sync even when the last replay is clean."""
    n, theta, eps, dt = 32768, 0.5, 0.01, 1e-6
    a = line_state(n, 3)
    b = ics.plummer(n, 4)
    ds = DeviceState.from_host(a)
    s = Solver(walk="fp32", persistent=2)
    s.reserve(full=1, partial=1, leaves=1)  # sized by the line state's own lists at capture
    say("[robust] capture on a line state")
    s.graph_capture_step(ds, [0, n], theta, eps, dt)
    s.graph_launch(1)
    s.sync()  # clean
    say("[robust] replay on a Plummer sphere in the same arrays, then the line again")
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        getattr(ds, k).upload(np.ascontiguousarray(b[k], dtype=np.float64))
    s.graph_launch(1)
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        getattr(ds, k).upload(np.ascontiguousarray(a[k], dtype=np.float64))
    s.graph_launch(2)
    with pytest.raises(_lib.GxError) as e:
        s.sync()
    assert e.value.code == _lib.GX_ERR_CAPACITY
    # reported once: the next clean replay syncs without error
    s.graph_launch(1)
    s.sync()


def test_small_node_arena_is_grown_and_the_step_repeated(gpu):
    """A node arena far below the tree's size, after a step that left only the
    levels a shallower tree needs: the overflow is flagged, the check of the
    deepest level reads nothing past the arena, and the repeated step equals
    a plain one bit for bit."""
    s0 = ics.plummer(20000, 11)
    p = Params(softening=0.01)
    want = DeviceState.from_host(s0)
    Solver(theta=0.5, walk="fp64").accelerations(want, p)
    ds = DeviceState.from_host(s0)
    sv = Solver(theta=0.5, walk="fp64")
    say("[robust] shallow tree first, then a node arena of 512")
    sv.accelerations(DeviceState.from_host(ics.plummer(2000, 12)), p)
    sv.reserve(nodes=512)
    sv.accelerations(ds, p)
    assert np.array_equal(ds.acc_host(), want.acc_host())
    assert sv.stats()["n_nodes"] > 512


def test_capture_is_refused_after_its_buffers_were_reallocated(gpu):
    small = ics.plummer(4096, 5)
    ds = DeviceState.from_host(small)
    s = Solver(walk="fp32")
    s.graph_capture_step(ds, [0, 4096], 0.6, 0.01, 1e-4)
    s.graph_launch(1)
    s.sync()
    say("[robust] a larger eager step reallocates the arena")
    s.accelerations(DeviceState.from_host(ics.plummer(200000, 6)), Params(softening=0.01))
    with pytest.raises(_lib.GxError) as e:
        s.graph_launch(1)
    assert e.value.code == _lib.GX_ERR_ARG


def test_last_gpu_seconds_with_timing_off(gpu):
    s = Solver(theta=0.5, walk="fp32", timing=False)
    assert s.last_gpu_seconds() == 0.0
    ds = DeviceState.from_host(ics.plummer(50000, 8))
    s.accelerations(ds, Params(softening=0.01))
    t1 = s.last_gpu_seconds()
    assert 0.0 < t1 < 1.0
    s.graph_capture_step(ds, [0, 50000], 0.5, 0.01, 1e-4)
    s.graph_launch(3)
    s.sync()
    assert 0.0 < s.last_gpu_seconds() < 1.0


def test_fresh_solver_metadata_starts_clean(gpu):
    """A new solver's step metadata (flags, counters and the sticky flag word
    at its end) starts at zero whatever the memory held before: blocks filled
    with set bits are freed first, so the allocator is likely to hand them out
    again, and every solver's first step must then find no flag set."""
    from gravitas_amd.hip import DeviceArray
    junk = [DeviceArray((size,), np.uint8) for size in (512, 1024, 2048, 4096) for _ in range(256)]
    for j in junk:
        j.fill(0xFF)
    del junk, j
    say("[robust] eight solvers on freed memory full of set bits")
    for k in range(8):
        sv = Solver(walk="fp32" if k % 2 else "fp64")
        ds = DeviceState.from_host(ics.plummer(4000, 20 + k))
        sv.accelerations(ds, Params(softening=0.01))
        assert sv.stats()["flags"] == 0
