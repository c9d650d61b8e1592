"""GPU parity of the plain reference method (the 16,384-star Plummer case):
one force evaluation plus one kick-drift-kick step of dt 1e-3 at theta 0.5,
softening 0.01, checked against the CPU oracle's Accelerations + KDK
(internal/sim/sim.go:262-281, 302-343) and against the FP64 direct sum."""
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import DeviceState, Params, Solver, ics

pytestmark = pytest.mark.gpu

N, THETA, EPS, DT = 16384, 0.5, 0.01, 1e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_step(s):
    o = oracle.Solver(theta=THETA, workers=8)
    acc = lambda x, y, z: o.accelerations(x, y, z, s["m"], s["kind"], eps=EPS)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    st["acc"] = acc(st["px"], st["py"], st["pz"])  # initForces
    a0 = st["acc"].copy()
    oracle.kick_drift_kick(st, acc, DT)
    return a0, st


@pytest.fixture(scope="module")
def plummer():
    return ics.plummer_equal_mass(N, seed=1)


def test_fp64_step_bit_identical(gpu, plummer):
    s = plummer
    say(f"[step] plummer n={N} fp64 force + KDK")
    a0, want = oracle_step(s)
    solver = Solver(theta=THETA, walk="fp64")
    ds = DeviceState.from_host(s)
    p = Params(softening=EPS, dt=DT)
    solver.accelerations(ds, p)
    assert np.array_equal(ds.acc_host(), a0)
    solver.step(ds, p, acc_valid=True)
    got = solver.host_state(ds)
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(np.stack([got["ax"], got["ay"], got["az"]], 1), want["acc"])


def test_step_without_valid_forces_computes_them_first(gpu, plummer):
    s = plummer
    say("[step] initForces path")
    _, want = oracle_step(s)
    solver = Solver(theta=THETA, walk="fp64")
    ds = DeviceState.from_host(s)
    solver.step(ds, Params(softening=EPS, dt=DT), acc_valid=False)
    got = solver.host_state(ds)
    assert np.array_equal(got["px"], want["px"]) and np.array_equal(got["vz"], want["vz"])


def test_fp32_step_close_and_errors_vs_direct(gpu, plummer):
    s = plummer
    say("[step] fp32 force + KDK, errors vs direct sum")
    a0, want = oracle_step(s)
    solver = Solver(theta=THETA, walk="fp32")
    ds = DeviceState.from_host(s)
    p = Params(softening=EPS, dt=DT)
    solver.step(ds, p, acc_valid=False)
    got = solver.host_state(ds)
    acc = np.stack([got["ax"], got["ay"], got["az"]], 1)
    e = np.sort(ics.rel_errors(acc, want["acc"]))
    assert e[len(e) // 2] <= 1e-6 and e[-1] <= 1e-4
    # positions move by dt*v ~ 1e-3: FP32 forces change them only at ~1e-13
    assert np.abs(got["px"] - want["px"]).max() <= 1e-10
    # tree error of the reference method itself, vs the FP64 direct sum
    direct = solver.direct_sample(DeviceState.from_host(s), p, np.arange(N)).numpy()
    ed = np.sort(ics.rel_errors(a0, direct))
    say(f"[step] theta 0.5 tree vs direct: p50 {ed[N // 2]:.2e} p99 {ed[int(0.99 * N)]:.2e} max {ed[-1]:.2e}")
    assert ed[int(0.99 * N)] <= 1e-2


def test_resume_from_saved_state_and_forces_is_exact(gpu, plummer):
    """sim.go:185-226 (LastForces / RestoreForces): a run resumed from saved
    positions, velocities and the last forces (acc_valid, no initForces)
    continues bit-identically, here across a host round trip and a fresh
    solver, for the KDK step and the FP32 walk."""
    s = plummer
    p = Params(softening=EPS, dt=DT)
    say("[step] resume from a saved state")
    a = Solver(theta=THETA, walk="fp32")
    ds = DeviceState.from_host(s)
    a.step(ds, p, acc_valid=False)
    saved = a.host_state(ds)  # the checkpoint: state plus the last force evaluation
    a.step(ds, p, acc_valid=True)
    straight = a.host_state(ds)
    state = dict(s)
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        state[k] = saved[k]
    resumed = DeviceState.from_host(state)
    for k in ("ax", "ay", "az"):
        getattr(resumed, k).upload(saved[k])
    b = Solver(theta=THETA, walk="fp32")
    b.step(resumed, p, acc_valid=True)
    again = b.host_state(resumed)
    for k in ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay", "az"):
        assert np.array_equal(again[k], straight[k]), k
