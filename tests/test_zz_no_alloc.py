"""The entry point that allocates no device memory (gx_accelerations_noalloc,
no counterpart in the reference): called through the C ABI, its forces equal
those of the allocating entry point (gx_accelerations_dev) bit for bit, on two
calls in a row (the second smaller, so it runs on what the first left in the
shared scratch), for the FP64 and FP32 walks; the size limit is an argument
error."""
import ctypes
import sys

import numpy as np
import pytest

from gravitas_amd import DeviceState, Params, Solver, _lib, accelerations_noalloc, ics

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def allocating(state, theta, eps, walk):
    ds = DeviceState.from_host(state)
    Solver(theta=theta, walk=walk).accelerations(ds, Params(softening=eps))
    return ds.acc_host()


@pytest.mark.parametrize("walk", ["fp64", "fp32"])
def test_noalloc_equals_the_allocating_entry_point(gpu, walk):
    cases = [("plummer 65536", ics.plummer(65536, 41), 0.5, 0.01),
             ("disk+halo 12000", ics.disk_halo(12000, seed=42, eps=0.02), 0.6, 0.02)]
    for name, state, theta, eps in cases:
        say(f"[noalloc] {walk} {name}")
        want = allocating(state, theta, eps, walk)
        ds = DeviceState.from_host(state)
        sentinel = np.full(len(state["px"]), 5.0)
        ds.ax.upload(sentinel), ds.ay.upload(sentinel), ds.az.upload(sentinel)
        accelerations_noalloc(ds, Params(softening=eps), theta=theta, walk=walk)
        assert np.array_equal(ds.acc_host(), want), name


def test_noalloc_with_holes_dead_bodies_and_dark_matter_off(gpu):
    s = ics.plummer(5000, 43)
    s["kind"][[3, 300]] = ics.KIND_HOLE
    s["kind"][1000:1500] = ics.KIND_DARK
    s["alive"] = np.ones(5000, np.uint8)
    s["alive"][[7, 4000]] = 0
    p = Params(softening=0.01, dark_matter_off=True)
    ds = DeviceState.from_host(s)
    Solver(theta=0.7, walk="fp64").accelerations(ds, p)
    want = ds.acc_host()
    d2 = DeviceState.from_host(s)
    say("[noalloc] holes, dead bodies, dark matter off")
    accelerations_noalloc(d2, p, theta=0.7, walk="fp64")
    assert np.array_equal(d2.acc_host(), want)


def test_noalloc_size_limit_is_an_argument_error(gpu):
    L = _lib.load()
    n = _lib.GX_NOALLOC_MAX_BODIES + 1
    rc = L.gx_accelerations_noalloc(None, None, None, None, None, None, n, 0.5, 0.01, 0, 0, None, None, None, None)
    assert rc == _lib.GX_ERR_ARG
    assert b"at most" in L.gx_error()
