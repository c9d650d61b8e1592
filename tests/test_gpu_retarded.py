"""GPU parity of retarded hole gravity (internal/sim/retarded.go) in the split
step (split.go:164-199), against the oracle's restatement (oracle.split_step
with an oracle.Retard), run side by side on the same inputs:

  * the light-cone source positions (sourcePos) of a uniformly moving hole
    for many observers are bit-identical to the oracle's, inside the
    8,192-entry history and after it has wrapped around (retarded time before
    the oldest entry: linear extrapolation), and lie on the light cone;
  * a two-body run and an equal-mass hole binary (every force a single pair,
    no sums over many bodies) follow the oracle bit for bit over hundreds of
    steps, and naive retardation widens the binary's orbit (retarded_test.go:
    99-121);
  * the gravity of a hole placed during the run reaches a star 2 away after
    2/c (retarded_test.go:50-73), on the same step as in the oracle;
  * a Plummer cluster with three moving holes in both retarded modes agrees
    with the oracle to round-off (1e-12; the holes' sums over the bodies are
    tree reductions on the GPU), with the same substep counts and history
    counts;
  * instant gravity with a history set is the plain split step, bit for bit,
    and keeps no histories.
"""
import ctypes
import math
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import DeviceState, Params, Retard, Solver, _lib, ics

pytestmark = pytest.mark.gpu

KEYS = ("px", "py", "pz", "vx", "vy", "vz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def proper_velocity(v, c):
    """ProperVelocity, internal/sim/relativity.go:38-46."""
    v = np.asarray(v, float)
    ln = math.sqrt(v @ v)
    if ln > 0.99 * c:
        v = v * (0.99 * c / ln)
    return v * (1 / math.sqrt(1 - (v @ v) / (c * c)))


def bodies(spec):
    n = len(spec)
    st = {k: np.zeros(n) for k in KEYS}
    m = np.zeros(n)
    kind = np.zeros(n, np.uint8)
    for i, b in enumerate(spec):
        st["px"][i], st["py"][i], st["pz"][i] = b.get("pos", (0, 0, 0))
        st["vx"][i], st["vy"][i], st["vz"][i] = b.get("vel", (0, 0, 0))
        m[i] = b["m"]
        kind[i] = b.get("kind", ics.KIND_STAR)
    return st, m, kind, np.ones(n, np.uint8)


class Pair:
    """The oracle and the GPU stepping the same state side by side; the time
    advances by dt after every step (sim.go:285-287)."""

    def __init__(self, st, m, kind, alive, dt, eps, c, gravity, theta=0.5):
        self.st, self.m, self.kind, self.alive = st, m, kind, alive
        self.dt, self.eps, self.c, self.gravity = dt, eps, c, gravity
        self.o = oracle.Solver(theta=theta, workers=8)
        self.ro = oracle.Retard()
        self.sv = Solver(theta=theta, walk="fp64")
        self.rg = Retard()
        self.p = Params(softening=eps, dt=dt, c=c)
        self.ds = DeviceState.from_host({**st, "m": m, "kind": kind, "alive": alive})
        self.time = 0.0
        self.steps = 0

    def step(self):
        want = oracle.split_step(self.o, self.st, self.m, self.kind, self.alive, self.eps, self.dt, c=self.c,
                                 acc_valid=self.steps > 0, retard=self.ro, time=self.time, gravity=self.gravity)
        got = self.sv.split_step(self.ds, self.p, acc_valid=self.steps > 0, retard=self.rg, time=self.time,
                                 gravity=self.gravity)
        assert got == want, (self.steps, got, want)
        self.time += self.dt
        self.steps += 1

    def gpu_state(self):
        return self.sv.host_state(self.ds)

    def sources(self, k, obs):
        got, vis = self.sv.source_pos(self.rg, self.ds, self.p, k, obs, self.time, self.gravity)
        want = [self.ro.source_pos(self.st, self.kind, self.alive, k, o, self.time, self.dt, self.c, self.gravity)
                for o in np.atleast_2d(obs)]
        return got, vis, want


def assert_state_equal(pair):
    g = pair.gpu_state()
    for k in KEYS:
        assert np.array_equal(g[k], pair.st[k]), (pair.steps, k, np.abs(g[k] - pair.st[k]).max())


def moving_hole(gravity, steps):
    """retarded_test.go:10-20: a light hole moving along x at 0.3 c."""
    st, m, kind, alive = bodies([
        {"vel": proper_velocity((3, 0, 0), 10), "m": 1e-12, "kind": ics.KIND_HOLE},
        {"pos": (50, 0, 0), "m": 1e-12}])
    pair = Pair(st, m, kind, alive, 0.004, 0.01, 10.0, gravity)
    for k in range(steps):
        if k % 1000 == 0:
            say(f"[retarded] moving hole {gravity}: step {k} of {steps}")
        pair.step()
    assert_state_equal(pair)
    return pair


@pytest.mark.parametrize("steps", [500, 9000])  # 9,000 > the 8,192 entries of the ring buffer
def test_light_cone_sources_bit_identical(gpu, steps):
    pair = moving_hole("retarded", steps)
    assert len(pair.rg) == len(pair.ro) == 1
    rng = np.random.default_rng(7 + steps)
    obs = rng.uniform(-60, 60, (200, 3))
    obs[0] = (1, 4, -2) if steps == 500 else (-400, 40, -20)
    obs[1] = pair.st["px"][0] + 1e-3, 0, 0  # within one step of the hole: its current position
    got, vis, want = pair.sources(0, obs)
    assert vis.all() and all(w is not None for w in want)
    assert np.array_equal(got, np.array(want))
    # on the light cone of uniform motion from the origin at v = 3
    d = np.linalg.norm(obs[0] - got[0])
    assert abs(d - 10 * (pair.time - got[0][0] / 3)) <= 1e-6
    assert got[0][0] < pair.st["px"][0]
    if steps == 9000:
        assert pair.time - d / 10 < pair.time - (8192 - 1) * pair.dt  # before the oldest entry


def test_extrapolated_sources_point_at_the_current_position(gpu):
    pair = moving_hole("retarded+v", 500)
    got, vis, want = pair.sources(0, [(1, 4, -2), (-30, 20, 5)])
    assert vis.all() and np.array_equal(got, np.array(want))
    cur = np.array([pair.st["px"][0], pair.st["py"][0], pair.st["pz"][0]])
    assert np.abs(got - cur).max() <= 1e-6


def test_gravity_of_a_new_hole_spreads_at_the_speed_of_light(gpu):
    st, m, kind, alive = bodies([{"pos": (2, 0, 0), "m": 1e-9}, {"m": 1.0, "kind": ics.KIND_HOLE}])
    alive[1] = 0  # not there yet
    pair = Pair(st, m, kind, alive, 0.004, 0.01, 10.0, "retarded+v")
    for _ in range(10):
        pair.step()
    assert len(pair.rg) == 0
    alive[1] = 1
    pair.ds.alive.upload(alive)
    pair.ro.hole_appeared(st, 1, pair.time, pair.dt, pair.c)
    pair.rg.hole_appeared(pair.sv, pair.ds, pair.p, 1, pair.time)
    assert len(pair.rg) == 1
    placed, arrived = pair.time, None
    for _ in range(100):
        g = pair.gpu_state()
        obs = (g["px"][0], g["py"][0], g["pz"][0])
        got, vis, want = pair.sources(0, [obs])
        assert bool(vis[0]) == (want[0] is not None)
        if arrived is None and vis[0]:
            arrived = pair.time
        pair.step()
    assert arrived is not None and abs(arrived - placed - 0.2) <= 2 * pair.dt, (arrived, placed)
    assert_state_equal(pair)


def test_hole_binary_follows_oracle_and_naive_retardation_widens_it(gpu):
    """retarded_test.go:75-121 for one orbit: an equal-mass hole binary at
    v ~ 0.1 c; naive retardation widens the orbit by > 20 %."""
    mh, d, c = 0.5, 1.0, 5.0
    v = math.sqrt(mh / (2 * d))
    period = 2 * math.pi * (d / 2) / v
    drift = {}
    for gravity in ("retarded", "retarded+v"):
        st, m, kind, alive = bodies([
            {"pos": (-d / 2, 0, 0), "vel": proper_velocity((0, -v, 0), c), "m": mh, "kind": ics.KIND_HOLE},
            {"pos": (d / 2, 0, 0), "vel": proper_velocity((0, v, 0), c), "m": mh, "kind": ics.KIND_HOLE}])
        pair = Pair(st, m, kind, alive, period / 1000, 0.001, c, gravity)
        say(f"[retarded] binary {gravity}: 1000 steps")
        total = 0.0
        for _ in range(1000):
            pair.step()
            g = pair.gpu_state() if pair.steps % 100 == 0 else None
            if g is not None:
                for k in KEYS:
                    assert np.array_equal(g[k], st[k]), (gravity, pair.steps, k)
            total += math.dist((st["px"][0], st["py"][0], st["pz"][0]), (st["px"][1], st["py"][1], st["pz"][1]))
        drift[gravity] = total / 1000 / d - 1
        assert len(pair.rg) == 2
    assert drift["retarded"] >= 0.2, drift
    assert abs(drift["retarded+v"]) <= drift["retarded"] / 10, drift


def cluster(n=3000, seed=43):
    s = ics.plummer(n, seed)
    s["alive"] = np.ones(n, np.uint8)
    s["kind"][[0, 1, 2]] = ics.KIND_HOLE
    s["m"][[0, 1, 2]] = [0.05, 0.03, 0.02]
    s["vx"][[0, 1, 2]] += [0.8, -0.6, 0.3]  # moving holes, so the light cone matters
    s["alive"][[10, 20]] = 0
    return s


@pytest.mark.parametrize("gravity", ["retarded", "retarded+v"])
def test_cluster_with_moving_holes_matches_oracle(gpu, gravity):
    s = cluster()
    st = {k: s[k].copy() for k in KEYS}
    pair = Pair(st, s["m"], s["kind"], s["alive"], 1e-2, 0.01, 4.0, gravity)
    for rep in range(40):
        pair.step()
        if rep % 10 == 9:
            say(f"[retarded] cluster {gravity}: step {rep + 1}")
            g = pair.gpu_state()
            for k in KEYS:
                scale = max(1.0, np.abs(st[k]).max())
                assert np.abs(g[k] - st[k]).max() <= 1e-12 * scale, (rep, k, np.abs(g[k] - st[k]).max())
    assert len(pair.rg) == len(pair.ro) == 3
    # the sources of each hole for a few bodies agree with the oracle's to round-off
    g = pair.gpu_state()
    obs = np.stack([g["px"][100:140], g["py"][100:140], g["pz"][100:140]], axis=1)
    for k in range(3):
        got, vis, want = pair.sources(k, obs)
        assert vis.all() and all(w is not None for w in want)
        assert np.abs(got - np.array(want)).max() <= 1e-11


def test_instant_mode_with_histories_is_the_plain_split_step(gpu):
    s = cluster(1500, 44)
    for gravity, c in (("instant", 4.0), ("retarded", 0.0)):
        a = DeviceState.from_host(s)
        b = DeviceState.from_host(s)
        sa, sb = Solver(theta=0.5, walk="fp64"), Solver(theta=0.5, walk="fp64")
        p = Params(softening=0.01, dt=1e-2, c=c)
        r = Retard()
        for k in range(3):
            sa.split_step(a, p, acc_valid=k > 0)
            sb.split_step(b, p, acc_valid=k > 0, retard=r, time=k * 1e-2, gravity=gravity)
        ha, hb = sa.host_state(a), sb.host_state(b)
        for key in KEYS:
            assert np.array_equal(ha[key], hb[key]), (gravity, key)
        assert len(r) == 0


def test_unknown_gravity_mode_is_an_argument_error(gpu):
    """ParseGravityMode (retarded.go:30-38) rejects unknown modes; the C ABI
    rejects an unknown mode number with GX_ERR_ARG and leaves the state."""
    s = cluster(200, 45)
    d = DeviceState.from_host(s)
    sv = Solver(theta=0.5, walk="fp64")
    r = Retard()
    L = _lib.load()
    sub = ctypes.c_int32(0)
    ptr = lambda a: a.ptr  # noqa: E731
    rc = L.gx_split_step_retarded_dev(
        sv._h, ptr(d.px), ptr(d.py), ptr(d.pz), ptr(d.vx), ptr(d.vy), ptr(d.vz), ptr(d.m), ptr(d.kind),
        ptr(d.alive), None, None, None, None, len(d), 0.01, 0, 1e-2, 4.0, 0, ptr(d.ax), ptr(d.ay), ptr(d.az),
        r._h, 0.0, 3, ctypes.byref(sub))
    assert rc == _lib.GX_ERR_ARG and b"gravity mode" in L.gx_error()
    assert np.array_equal(sv.host_state(d)["px"], s["px"])


def test_spinning_holes_with_retarded_gravity_match_oracle(gpu):
    """Frame dragging (framedrag.go) and retarded sources in one split step:
    the spins ride along with gx_split_step_retarded_dev."""
    s = cluster(2000, 46)
    c, dt = 4.0, 1e-2
    for k in ("jx", "jy", "jz"):
        s[k] = np.zeros(len(s["px"]))
    s["jz"][0] = 0.5 * 0.05 ** 2 / c  # a* = 0.5: J = a* M^2 / c
    s["jx"][1] = 0.3 * 0.03 ** 2 / c
    keys = KEYS + ("jx", "jy", "jz")
    st = {k: s[k].copy() for k in keys}
    o, ro = oracle.Solver(theta=0.5, workers=8), oracle.Retard()
    ds = DeviceState.from_host(s)
    sv, rg = Solver(theta=0.5, walk="fp64"), Retard()
    p = Params(softening=0.01, dt=dt, c=c)
    for rep in range(15):
        want = oracle.split_step(o, st, s["m"], s["kind"], s["alive"], 0.01, dt, c=c, acc_valid=rep > 0, retard=ro,
                                 time=rep * dt, gravity="retarded+v")
        got = sv.split_step(ds, p, acc_valid=rep > 0, retard=rg, time=rep * dt, gravity="retarded+v")
        assert got == want, (rep, got, want)
    say("[retarded] spinning holes: 15 steps")
    h = sv.host_state(ds)
    for k in keys:
        scale = max(1.0, np.abs(st[k]).max()) if k in KEYS else np.abs(st["jz"]).max()
        assert np.abs(h[k] - st[k]).max() <= 1e-12 * scale, (k, np.abs(h[k] - st[k]).max())
    assert not np.array_equal(st["jz"], s["jz"])
    assert len(rg) == len(ro) == 3
