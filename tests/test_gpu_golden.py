"""The HIP path against the committed golden fixtures (tests/golden/*.npz):
the same stored inputs on every machine, the oracle's stored outputs as the
answer.

  * FP64 walk, keys, permutation, KDK step, block step: bit-identical;
  * potential energy and split steps (also with a spinning hole and
    retarded hole gravity): 1e-12 relative (their sums over many bodies are
    tree reductions on the GPU and sequential in the oracle);
  * FP32 group walk: the reference's GPU bound (gpu_test.go:18: median
    <= 1e-5, max <= 1e-3 relative to the FP64 forces).
"""
import sys

import numpy as np
import pytest

from gravitas_amd import DeviceState, Params, Retard, Solver, ics
from test_golden import EDGES, KEYS, load, state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_acc(s, theta, eps, walk="fp64", dm_off=False, monopole=False):
    sv = Solver(theta=theta, walk=walk, monopole_only=monopole)
    ds = DeviceState.from_host(s)
    sv.accelerations(ds, Params(softening=eps, dark_matter_off=dm_off))
    return sv, ds.acc_host()


def within_gpu_bound(got, want, mask=None):
    if mask is not None:
        got, want = got[mask], want[mask]
    e = np.sort(ics.rel_errors(got, want))
    assert e[len(e) // 2] <= 1e-5 and e[-1] <= 1e-3, (e[len(e) // 2], e[-1])


def test_plummer_fixture(gpu):
    f = load("plummer_512.npz")
    s = state(f)
    eps, theta = float(f["eps"]), float(f["theta"])
    say("[golden] plummer_512 fp64 forces, keys, order")
    sv, a = gpu_acc(s, theta, eps)
    assert np.array_equal(a, f["acc"])
    assert np.array_equal(sv.export("keys"), f["keys"])
    assert np.array_equal(sv.export("order"), f["order"])
    assert sv.stats()["n_nodes"] == int(f["n_nodes"])
    assert np.array_equal(gpu_acc(s, 0.0, eps)[1], f["acc_theta0"])
    assert np.array_equal(gpu_acc(s, theta, eps, monopole=True)[1], f["acc_mono"])
    say("[golden] plummer_512 potential, KDK step, fp32 walk")
    ds = DeviceState.from_host(s)
    u = sv.potential_energy(ds, Params(softening=eps))
    assert abs(u - float(f["potential"])) <= 1e-12 * abs(float(f["potential"]))
    sv.step(ds, Params(softening=eps, dt=float(f["dt"])), acc_valid=False)
    h = sv.host_state(ds)
    for k in KEYS:
        assert np.array_equal(h[k], f["kdk_" + k]), k
    assert np.array_equal(ds.acc_host(), f["kdk_acc"])
    for walk in ("fp32", "fp32-q16"):
        got = gpu_acc(s, theta, eps, walk=walk)[1]
        if walk == "fp32":
            within_gpu_bound(got, f["acc"])
        else:
            e = np.sort(ics.rel_errors(got, f["acc"]))
            assert e[len(e) // 2] <= 1e-4 and e[-1] <= 1e-2, (e[len(e) // 2], e[-1])


def test_mixed_fixture(gpu):
    f = load("mixed_768.npz")
    s = state(f)
    eps, theta = float(f["eps"]), float(f["theta"])
    say("[golden] mixed_768 fp64 forces, dark matter on and off")
    sv, a = gpu_acc(s, theta, eps)
    assert np.array_equal(a, f["acc"])
    assert np.array_equal(gpu_acc(s, theta, eps, dm_off=True)[1], f["acc_dmoff"])
    live_stars = (s["alive"] == 1) & (s["kind"] != ics.KIND_HOLE)
    within_gpu_bound(gpu_acc(s, theta, eps, walk="fp32")[1], f["acc"], live_stars)
    for c in (0, 40):
        say(f"[golden] mixed_768 potential c={c}")
        u = sv.potential_energy(DeviceState.from_host(s), Params(softening=eps, c=float(c)))
        want = float(f[f"potential_c{c}"])
        assert abs(u - want) <= 1e-12 * abs(want)
    say("[golden] mixed_768 split step")
    ds = DeviceState.from_host(s)
    sub = Solver(theta=theta, walk="fp64").split_step(ds, Params(softening=eps, dt=float(f["dt"])), acc_valid=False)
    assert sub == int(f["split_substeps"])
    h = sv.host_state(ds)
    for k in KEYS:
        want = f["split_" + k]
        assert np.abs(h[k] - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), k


def test_block_step_fixture(gpu):
    f = load("encounter_768.npz")
    s = state(f)
    say("[golden] encounter_768 block step")
    ds = DeviceState.from_host(s)
    sv = Solver(theta=float(f["theta"]), walk="fp64")
    r = sv.block_step(ds, Params(softening=float(f["eps"])), float(f["T"]), int(f["levels"]), float(f["eta"]),
                      acc_valid=False)
    assert r["rung_count"] == f["rung_count"].tolist()
    assert [list(e) for e in zip(r["tick"], r["active"])] == f["events"].tolist()
    h = sv.host_state(ds)
    for k in KEYS:
        assert np.array_equal(h[k], f["block_" + k]), k
    assert np.array_equal(ds.acc_host(), f["block_acc"])


@pytest.mark.parametrize("case", EDGES)
def test_edge_fixtures(gpu, case):
    f = load("edges.npz")
    s = {k: f[f"{case}__{k}"] for k in ("px", "py", "pz", "m", "kind", "alive")}
    eps = float(f[f"{case}__eps"])
    want = f[f"{case}__acc"]
    say(f"[golden] edge case {case}")
    assert np.array_equal(gpu_acc(s, 0.0, eps)[1], want)
    got = gpu_acc(s, 0.0, eps, walk="fp32")[1]
    assert np.isfinite(got).all()
    assert np.abs(got - want).max() <= 1e-5 * max(1.0, np.abs(want).max())


def test_hydro_fixture(gpu):
    from gravitas_amd import GasState, Hydro
    f = load("gas_768.npz")
    s = state(f)
    say("[golden] gas_768 hydro")
    g = GasState.from_host(s)
    hy = Hydro()
    out = hy.forces(g, softening=float(f["eps"]), step=f["step"]).host()
    h = g.host()
    assert np.array_equal(h["h"], f["h_out"]) and np.array_equal(h["rho"], f["rho_out"])
    for k in ("ax", "ay", "az", "dudt", "divv", "dt", "wake"):
        assert np.array_equal(out[k], f[k]), k
    assert hy.stats()["density_sums"] == int(f["evals"])


def test_holes_fixture(gpu):
    """25 split steps with a spinning hole, frame dragging and retarded+v hole
    gravity: substeps equal, state and spins within 1e-12, the holes' sources
    within 1e-11."""
    f = load("holes_512.npz")
    s = state(f)
    c, eps, dt = float(f["c"]), float(f["eps"]), float(f["dt"])
    ds = DeviceState.from_host(s)
    sv, r = Solver(theta=float(f["theta"]), walk="fp64"), Retard()
    p = Params(softening=eps, dt=dt, c=c)
    say("[golden] holes_512: 25 split steps with spin and retarded gravity")
    subs = [sv.split_step(ds, p, acc_valid=k > 0, retard=r, time=k * dt, gravity="retarded+v")
            for k in range(int(f["steps"]))]
    assert subs == list(f["substeps"])
    h = sv.host_state(ds)
    for k in KEYS:
        want = f["out_" + k]
        assert np.abs(h[k] - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), k
    for k in ("jx", "jy", "jz"):
        assert np.abs(h[k] - f["out_" + k]).max() <= 1e-12 * np.abs(f["out_jz"]).max(), k
    for hole in range(2):
        got, vis = sv.source_pos(r, ds, p, hole, f["obs"], int(f["steps"]) * dt, "retarded+v")
        assert vis.all() and np.abs(got - f[f"source{hole}"]).max() <= 1e-11, hole
