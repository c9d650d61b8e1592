"""GPU parity of steps with black holes (internal/sim/split.go:164-199): the
impulse/RESPA scheme with tree forces between stars, direct sums for every
pair with a hole and power-of-two substeps for the holes and the stars near
them, against the oracle's restatement (oracle.split_step).

Per-body sums over the holes run in the reference's hole order; the holes'
sums over the bodies are tree reductions on the GPU and sequential in the
oracle (the reference itself adds them in completion order), so the states
agree to round-off (1e-12), and the substep counts exactly."""
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import DeviceState, Params, Solver, ics

pytestmark = pytest.mark.gpu

EPS, THETA = 0.01, 0.5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def hole_state(n=5000, seed=41):
    s = ics.plummer(n, seed)
    s["alive"] = np.ones(n, np.uint8)
    s["kind"][[0, 1, 2]] = ics.KIND_HOLE
    s["m"][[0, 1, 2]] = [0.05, 0.03, 0.02]
    s["alive"][[10, 20]] = 0  # dead bodies still take part in the hole sums, as in the reference
    s["kind"][30:60] = ics.KIND_DARK
    return s


def compare(got, want, tol=1e-12):
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        scale = max(1.0, np.abs(want[k]).max())
        assert np.abs(got[k] - want[k]).max() <= tol * scale, (k, np.abs(got[k] - want[k]).max())


@pytest.mark.parametrize("c", [0.0, 40.0])
def test_split_steps_match_oracle(gpu, c):
    s = hole_state()
    dt = 1e-2
    o = oracle.Solver(theta=THETA, workers=8)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    ds = DeviceState.from_host(s)
    sv = Solver(theta=THETA, walk="fp64")
    p = Params(softening=EPS, dt=dt, c=c, dark_matter_off=True)
    for rep in range(3):
        say(f"[split] c={c} step {rep}")
        want = oracle.split_step(o, st, s["m"], s["kind"], s["alive"], EPS, dt, c=c, dm_off=True, acc_valid=rep > 0)
        got = sv.split_step(ds, p, acc_valid=rep > 0)
        assert got == want and got > 1, (got, want)
        compare(sv.host_state(ds), st)


def test_split_without_holes_is_the_plain_step(gpu):
    s = ics.plummer(4000, 42)
    p = Params(softening=EPS, dt=1e-3)
    a = DeviceState.from_host(s)
    b = DeviceState.from_host(s)
    sa = Solver(theta=THETA, walk="fp64")
    sb = Solver(theta=THETA, walk="fp64")
    say("[split] no holes: split step vs KDK step")
    assert sa.split_step(a, p, acc_valid=False) == 1
    sb.step(b, p, acc_valid=False)
    ha, hb = sa.host_state(a), sb.host_state(b)
    for k in ("px", "py", "pz", "vx", "vy", "vz", "ax"):
        assert np.array_equal(ha[k], hb[k]), k
    # and with c > 0 the oracle's relativistic drift
    c = 5.0
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    o = oracle.Solver(theta=THETA, workers=8)
    oracle.split_step(o, st, s["m"], s["kind"], np.ones(4000, np.uint8), EPS, 1e-3, c=c, acc_valid=False)
    d = DeviceState.from_host(s)
    Solver(theta=THETA, walk="fp64").split_step(d, Params(softening=EPS, dt=1e-3, c=c), acc_valid=False)
    hd = sa.host_state(d)
    for k in ("px", "vy"):
        assert np.array_equal(hd[k], st[k]), k


def test_fp32_split_step_close(gpu):
    s = hole_state(8000, 43)
    dt = 1e-2
    o = oracle.Solver(theta=THETA, workers=8)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    want = oracle.split_step(o, st, s["m"], s["kind"], s["alive"], EPS, dt, acc_valid=False)
    ds = DeviceState.from_host(s)
    say("[split] fp32 walk")
    got = Solver(theta=THETA, walk="fp32").split_step(ds, Params(softening=EPS, dt=dt), acc_valid=False)
    assert got == want
    h = Solver().host_state(ds)
    assert np.abs(h["px"] - st["px"]).max() <= 1e-8
    assert np.abs(h["vx"] - st["vx"]).max() <= 1e-5
