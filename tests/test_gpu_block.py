"""GPU parity of active-set force evaluations, block time steps and the FP64
diagnostics (the two-disk encounter workload at test size).

An active star's force must be exactly what a full evaluation gives it (the
tree is still built from every star); block steps follow the rung rules of
internal/sim/blockstep.go:23-90 and its event loop :130-198 applied to
gravity, checked against the oracle's restatement; diagnostics follow
internal/sim/diag.go."""
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import DeviceState, Params, Solver, ics

pytestmark = pytest.mark.gpu

EPS, THETA, ETA, LEVELS = 0.02, 0.5, 0.025, 4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def encounter():
    s = ics.two_disk_encounter(24576, seed=21)
    s["alive"] = np.ones(len(s["px"]), np.uint8)
    s["alive"][[3, 20000]] = 0
    s["kind"][[7, 15000]] = ics.KIND_HOLE
    return s


def oracle_acc(s):
    o = oracle.Solver(theta=THETA, workers=8)
    return lambda x, y, z: o.accelerations(x, y, z, s["m"], s["kind"], s["alive"], eps=EPS)


def step_T(acc, frac=0.03):
    """Global step putting about frac of the stars on the top of 4 rungs."""
    a = np.linalg.norm(acc, axis=1)
    a = a[a > 0]
    return 4 * float(np.quantile(np.sqrt(2 * ETA * EPS / a), frac))


def test_active_forces_equal_full_evaluation(gpu, encounter):
    s = encounter
    n = len(s["px"])
    rng = np.random.default_rng(5)
    active = (rng.random(n) < 0.05).astype(np.uint8)
    star = (s["alive"] == 1) & (s["kind"] != ics.KIND_HOLE)
    sel = (active == 1) & star
    say(f"[block] active set {int(sel.sum())} of {n}, fp64")
    full = DeviceState.from_host(s)
    s64 = Solver(theta=THETA, walk="fp64")
    s64.accelerations(full, Params(softening=EPS))
    want = full.acc_host()
    ds = DeviceState.from_host(s)
    sentinel = np.full(n, 7.0)
    ds.ax.upload(sentinel), ds.ay.upload(sentinel), ds.az.upload(sentinel)
    s64.accelerations_active(ds, Params(softening=EPS), active)
    got = ds.acc_host()
    assert np.array_equal(got[sel], want[sel])
    rest = ~sel & (s["kind"] != ics.KIND_HOLE)
    assert (got[rest] == 7.0).all()
    assert (got[(s["kind"] == ics.KIND_HOLE) & (s["alive"] == 1)] == 0).all()
    st64 = s64.stats()
    assert st64["n_targets"] == int(sel.sum())
    for walk, (p50, mx) in (("fp32", (1e-5, 1e-3)), ("fp32-q16", (1e-4, 1e-2))):
        for gsize, persistent in ((64, 0), (32, 2), (64, 2)):  # 2 forces the work queue (active groups)
            say(f"[block] active set, {walk} group={gsize} persistent={persistent}")
            d = DeviceState.from_host(s)
            sv = Solver(theta=THETA, walk=walk, group_size=gsize, persistent=persistent)
            sv.accelerations_active(d, Params(softening=EPS), active)
            st = sv.stats()
            assert (st["pp_interactions"], st["pc_interactions"]) == (st64["pp_interactions"],
                                                                     st64["pc_interactions"])
            assert st["fallbacks"] == 0
            e = np.sort(ics.rel_errors(d.acc_host()[sel], want[sel]))
            assert e[len(e) // 2] <= p50 and e[-1] <= mx, (walk, e[len(e) // 2], e[-1])


def test_empty_active_set_changes_nothing(gpu, encounter):
    s = encounter
    n = len(s["px"])
    ds = DeviceState.from_host(s)
    sentinel = np.full(n, 3.0)
    ds.ax.upload(sentinel)
    sv = Solver(theta=THETA, walk="fp32")
    sv.accelerations_active(ds, Params(softening=EPS), np.zeros(n, np.uint8))
    got = ds.ax.numpy()
    holes = (s["kind"] == ics.KIND_HOLE) & (s["alive"] == 1)
    assert (got[~holes] == 3.0).all() and (got[holes] == 0).all()
    assert sv.stats()["n_targets"] == 0 and sv.stats()["pp_interactions"] == 0


def test_fp64_block_steps_bit_identical_to_oracle(gpu, encounter):
    s = encounter
    acc = oracle_acc(s)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    st["acc"] = acc(st["px"], st["py"], st["pz"])
    T = step_T(st["acc"])
    ds = DeviceState.from_host(s)
    sv = Solver(theta=THETA, walk="fp64")
    p = Params(softening=EPS)
    for rep in range(2):
        say(f"[block] fp64 block step {rep}, T={T:.3e}")
        want = oracle.block_step(st, acc, T, LEVELS, ETA, EPS, s["kind"], s["alive"])
        got = sv.block_step(ds, p, T, LEVELS, ETA, acc_valid=rep > 0)
        assert got["rung_count"] == want["rung_count"], (got["rung_count"], want["rung_count"])
        assert list(zip(got["tick"], got["active"])) == want["events"]
        h = sv.host_state(ds)
        for k in ("px", "py", "pz", "vx", "vy", "vz"):
            assert np.array_equal(h[k], st[k]), (rep, k)
        assert np.array_equal(np.stack([h["ax"], h["ay"], h["az"]], 1), st["acc"])
    assert want["rung_count"][3] > 0 and want["rung_count"][0] > 0


def test_fp32_block_steps_follow_the_oracle(gpu, encounter):
    s = encounter
    acc = oracle_acc(s)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    st["acc"] = acc(st["px"], st["py"], st["pz"])
    T = step_T(st["acc"])
    want = oracle.block_step(st, acc, T, LEVELS, ETA, EPS, s["kind"], s["alive"])
    live = s["alive"] == 1
    for walk in ("fp32", "fp32-q16"):
        say(f"[block] {walk} block step vs oracle")
        ds = DeviceState.from_host(s)
        sv = Solver(theta=THETA, walk=walk)
        got = sv.block_step(ds, Params(softening=EPS), T, LEVELS, ETA, acc_valid=False)
        # FP32 forces may move a star across a rung boundary: the populations agree closely
        assert np.abs(np.array(got["rung_count"]) - np.array(want["rung_count"])).sum() <= 0.002 * len(s["px"])
        h = sv.host_state(ds)
        dx = np.abs(h["px"][live] - st["px"][live]).max()
        dv = np.abs(h["vx"][live] - st["vx"][live]).max()
        say(f"[block] {walk}: max |dx| {dx:.2e} max |dv| {dv:.2e}")
        assert dx <= 1e-6 and dv <= 1e-3


def test_diagnostics_match_exact_sums(gpu, encounter):
    """The GPU's fixed-order tree sums against math.fsum of the same terms
    (the terms are those of the oracle's diag.go restatement, which sums them
    sequentially in state order)."""
    import math
    s = encounter
    ds = DeviceState.from_host(s)
    sv = Solver(theta=THETA)
    m = s["m"]
    pos = np.stack([s["px"], s["py"], s["pz"]], 1)
    mv = np.stack([s["vx"] * m, s["vy"] * m, s["vz"] * m], 1)
    L = np.stack([pos[:, 1] * mv[:, 2] - pos[:, 2] * mv[:, 1], pos[:, 2] * mv[:, 0] - pos[:, 0] * mv[:, 2],
                  pos[:, 0] * mv[:, 1] - pos[:, 1] * mv[:, 0]], 1)
    for dm_off in (False, True):
        d = sv.diagnostics(ds, Params(softening=EPS, dark_matter_off=dm_off))
        w = oracle.diagnostics(s["px"], s["py"], s["pz"], s["vx"], s["vy"], s["vz"], m, s["kind"], dm_off)
        src = ~((s["kind"] == ics.KIND_DARK) & dm_off)
        kin = m * (s["vx"] ** 2 + s["vy"] ** 2 + s["vz"] ** 2) / 2
        terms = [kin[src]] + [mv[:, c] for c in range(3)] + [L[:, c] for c in range(3)] + [m] + \
                [pos[:, c] * m for c in range(3)]
        got = [d["kinetic"], *d["momentum"], *d["angular_momentum"], d["mass"]]
        for c, t in enumerate(terms[:8]):
            exact, gross = math.fsum(t), np.abs(t).sum()
            assert abs(got[c] - exact) <= 1e-13 * gross, (c, got[c], exact)
            assert abs(w[c] - exact) <= 1e-11 * gross, (c, w[c], exact)  # the oracle's sequential sum
        assert d["com"] == pytest.approx([math.fsum(t) / math.fsum(m) for t in terms[8:]], rel=1e-12, abs=1e-12)
