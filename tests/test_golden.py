"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py): the oracle reproduces every stored output
bit for bit from the stored inputs, so a change to the oracle that moves a
result shows up here, and the fixtures' known answers hold
(bh_test.go:137-161)."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ("px", "py", "pz", "vx", "vy", "vz")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def state(f, prefix="in_"):
    return {k[len(prefix):]: v.copy() for k, v in f.items() if k.startswith(prefix)}


def acc(o, s, eps, dm_off=False):
    n = len(s["px"])
    return o.accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"], s["alive"], eps=eps, dm_off=dm_off,
                           out=tuple(np.zeros(n) for _ in range(3)))


def test_fixture_files_hold_no_objects():
    for name in ("plummer_512.npz", "mixed_768.npz", "encounter_768.npz", "gas_768.npz", "holes_512.npz",
                 "edges.npz"):
        for k, v in load(name).items():
            assert v.dtype != object, (name, k)


def test_oracle_reproduces_plummer_fixture():
    f = load("plummer_512.npz")
    s = state(f)
    eps, theta = float(f["eps"]), float(f["theta"])
    o = oracle.Solver(theta=theta)
    assert np.array_equal(acc(o, s, eps), f["acc"])
    assert np.array_equal(o.keys(), f["keys"]) and np.array_equal(o.order(), f["order"])
    assert o.n_nodes == int(f["n_nodes"])
    assert np.array_equal(acc(oracle.Solver(theta=0.0), s, eps), f["acc_theta0"])
    assert np.array_equal(acc(oracle.Solver(theta=theta, monopole_only=True), s, eps), f["acc_mono"])
    assert o.potential_energy(s["px"], s["py"], s["pz"], s["m"], s["kind"], s["alive"], eps=eps) == f["potential"]
    st = {k: s[k].copy() for k in KEYS}
    st["acc"] = f["acc"].copy()
    oracle.kick_drift_kick(st, lambda x, y, z: acc(o, dict(s, px=x, py=y, pz=z), eps), float(f["dt"]))
    for k in KEYS:
        assert np.array_equal(st[k], f["kdk_" + k]), k
    assert np.array_equal(st["acc"], f["kdk_acc"])
    # the keys are sorted and the permutation is one
    assert (np.diff(f["keys"].astype(np.float64)) >= 0).all()
    assert np.array_equal(np.sort(f["order"]), np.arange(len(s["px"])))


def test_oracle_reproduces_mixed_fixture():
    f = load("mixed_768.npz")
    s = state(f)
    eps, theta = float(f["eps"]), float(f["theta"])
    o = oracle.Solver(theta=theta)
    a = acc(o, s, eps)
    assert np.array_equal(a, f["acc"])
    assert np.array_equal(acc(o, s, eps, dm_off=True), f["acc_dmoff"])
    holes = (s["kind"] == 1) & (s["alive"] == 1)
    assert holes.sum() == 2 and (a[holes] == 0).all()
    for c in (0, 40):
        assert o.potential_energy(s["px"], s["py"], s["pz"], s["m"], s["kind"], s["alive"], eps=eps,
                                  c=float(c)) == f[f"potential_c{c}"]
    st = {k: s[k].copy() for k in KEYS}
    sub = oracle.split_step(o, st, s["m"], s["kind"], s["alive"], eps, float(f["dt"]), acc_valid=False)
    assert sub == int(f["split_substeps"]) and sub > 1
    for k in KEYS:
        assert np.array_equal(st[k], f["split_" + k]), k


def test_oracle_reproduces_block_step_fixture():
    f = load("encounter_768.npz")
    s = state(f)
    eps = float(f["eps"])
    o = oracle.Solver(theta=float(f["theta"]))
    fn = lambda x, y, z: acc(o, dict(s, px=x, py=y, pz=z), eps)
    st = {k: s[k].copy() for k in KEYS}
    st["acc"] = fn(st["px"], st["py"], st["pz"])
    r = oracle.block_step(st, fn, float(f["T"]), int(f["levels"]), float(f["eta"]), eps, s["kind"], s["alive"])
    assert r["rung_count"] == f["rung_count"].tolist()
    assert [list(e) for e in r["events"]] == f["events"].tolist()
    assert min(r["rung_count"]) > 0  # every rung is populated
    for k in KEYS:
        assert np.array_equal(st[k], f["block_" + k]), k
    assert np.array_equal(st["acc"], f["block_acc"])


EDGES = ("two", "coincident", "leaf24", "leaf25", "single")


@pytest.mark.parametrize("case", EDGES)
def test_oracle_reproduces_edge_cases(case):
    f = load("edges.npz")
    s = {k: f[f"{case}__{k}"] for k in ("px", "py", "pz", "m", "kind", "alive")}
    got = acc(oracle.Solver(theta=0.0), s, float(f[f"{case}__eps"]))
    assert np.array_equal(got, f[f"{case}__acc"])
    assert np.isfinite(got).all()
    if case == "two":
        assert got.tolist() == [[2, 0, 0], [-1, 0, 0], [0, 0, 0]]
    if case == "single":
        assert (got == 0).all()


def test_oracle_reproduces_hydro_fixture():
    f = load("gas_768.npz")
    s = state(f)
    out, evals = oracle.sph_forces(s, softening=float(f["eps"]), step=f["step"])
    assert evals == int(f["evals"]) and evals > 700  # h was guessed: several density passes
    assert np.array_equal(s["h"], f["h_out"]) and np.array_equal(s["rho"], f["rho_out"])
    for k in ("ax", "ay", "az", "dudt", "divv", "dt", "wake"):
        assert np.array_equal(out[k], f[k]), k
    assert np.isfinite(f["wake"].view(np.float64)).any()  # the limiter woke some bodies


def test_oracle_reproduces_holes_fixture():
    """25 split steps with a spinning hole, frame dragging and retarded+v
    hole gravity (split.go, framedrag.go, retarded.go)."""
    f = load("holes_512.npz")
    s = state(f)
    c, eps, dt = float(f["c"]), float(f["eps"]), float(f["dt"])
    o, r = oracle.Solver(theta=float(f["theta"])), oracle.Retard()
    st = {k: s[k].copy() for k in KEYS + ("jx", "jy", "jz")}
    subs = [oracle.split_step(o, st, s["m"], s["kind"], s["alive"], eps, dt, c=c, acc_valid=k > 0, retard=r,
                              time=k * dt, gravity="retarded+v") for k in range(int(f["steps"]))]
    assert subs == list(f["substeps"]) and max(subs) > 1
    for k in st:
        assert np.array_equal(st[k], f["out_" + k]), k
    assert not np.array_equal(st["jz"], s["jz"])  # the spin was dragged
    for h in range(2):
        for q, want in zip(f["obs"], f[f"source{h}"]):
            got = r.source_pos(st, s["kind"], s["alive"], h, q, int(f["steps"]) * dt, dt, c, "retarded+v")
            assert got is not None and np.array_equal(got, want), (h, q)
