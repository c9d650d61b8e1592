"""GPU parity of the SPH hydro (gx_hydro_*) against the oracle
(oracle/sph_oracle.c, a restatement of internal/sph/tree.go and
sph.Solver.Forces, sph.go:83-263).

Everything is FP64 in the reference's order of operations, so the bar is
bit-identical: smoothing lengths, densities, accelerations, du/dt, div v,
time steps and wake values, the density-pass count, and the tree itself
(keys, the tree order after the leaf sort by h, every node's range, box and
hmax)."""
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import GasParams, GasState, Hydro, ics

pytestmark = pytest.mark.gpu

OUT = ("ax", "ay", "az", "dudt", "divv", "dt")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(s, softening=0.0, gas=None, active=None, step=None, hy=None):
    """One evaluation on the GPU and in the oracle from the same state;
    returns (gpu state, gpu out, oracle state, oracle out, oracle evals, hydro)."""
    gas = gas or GasParams()
    so = {k: v.copy() for k, v in s.items()}
    fill = 7.0
    oo = {k: np.full(len(s["px"]), fill) for k in OUT}
    oo, evals = oracle.sph_forces(so, softening=softening, gamma=gas.gamma, ngb=gas.neighbors, alpha=gas.alpha,
                                  active=active, step=step, out=oo)
    hy = hy or Hydro()
    g = GasState.from_host(s)
    from gravitas_amd.hydro import HydroOut
    out = HydroOut(len(s["px"]), fill=fill)
    hy.forces(g, softening=softening, gas=gas, active=active, step=step, out=out)
    return g.host(), out.host(), so, oo, evals, hy


def assert_same(gs, go, so, oo, evals, hy):
    for k in ("h", "rho"):
        assert np.array_equal(gs[k], so[k]), (k, np.abs(gs[k] - so[k]).max())
    for k in OUT:
        same = (go[k] == oo[k]) | (np.isnan(go[k]) & np.isnan(oo[k]))
        assert same.all(), (k, np.nonzero(~same)[0][:5], go[k][~same][:3], oo[k][~same][:3])
    assert np.array_equal(go["wake"], oo["wake"])
    assert hy.stats()["density_sums"] == evals


def test_cloud_from_missing_h_is_bit_identical(gpu):
    s = ics.gas_cloud(3000, 2)
    say("[hydro] cloud 3000, h guessed")
    r = run_both(s, softening=0.01)
    assert_same(*r)
    st = r[5].stats()
    assert st["n_gas"] == 3000 and st["max_iterations"] > 1 and st["pairs"] > 0


def test_tree_matches_oracle(gpu):
    s = ics.gas_cloud(5000, 3)
    s["h"] = 0.03 + 0.1 * np.random.default_rng(4).random(5000)
    say("[hydro] tree of 5000")
    gs, go, so, oo, evals, hy = run_both(s)
    assert_same(gs, go, so, oo, evals, hy)
    t = oracle.SphTree()
    gas = np.arange(5000, dtype=np.int32)
    t.build(gas, so["px"], so["py"], so["pz"], so["h"])  # final h: the hmax the force pass saw
    t.sort_leaves()
    assert np.array_equal(hy.export("keys"), t.keys())
    assert np.array_equal(hy.export("order"), t.order())
    gn, on = hy.export("nodes"), t.nodes()
    assert len(gn) == len(on) and hy.stats()["n_leaves"] == len(t.leaves())
    gi = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(gn["lo"], gn["hi"]))}
    for nd in on:
        g = gn[gi[(int(nd["lo"]), int(nd["hi"]))]]
        for f in ("minx", "miny", "minz", "maxx", "maxy", "maxz", "hmax"):
            assert g[f] == nd[f], f
        assert (g["first"] < 0) == (nd["left"] < 0)
        if nd["left"] >= 0:  # children: the same ranges; the right one is walked first
            assert (gn["lo"][g["first"]], gn["lo"][g["second"]]) == (on["lo"][nd["right"]], on["lo"][nd["left"]])


def test_gas_sphere_two_evaluations(gpu):
    s = ics.gas_sphere(20000, 9)
    hy = Hydro()
    for rep in range(2):
        say(f"[hydro] gas sphere 20000, evaluation {rep}")
        gs, go, so, oo, evals, _ = run_both(s, softening=0.05, hy=hy)
        assert_same(gs, go, so, oo, evals, hy)
        s = {**s, "h": gs["h"], "rho": gs["rho"]}


def test_mixed_kinds_active_subset_and_limiter(gpu):
    s = ics.gas_cloud(6000, 5)
    s["kind"][::11] = ics.KIND_STAR
    s["kind"][5::13] = ics.KIND_DARK
    s["alive"][[4, 40, 400]] = 0
    rng = np.random.default_rng(6)
    oracle.sph_forces(s, softening=0.02)  # a converged start: h and rho set
    active = (rng.random(6000) < 0.3).astype(np.uint8)
    step = np.where(rng.random(6000) < 0.1, 10.0, 1e-3)
    say("[hydro] mixed kinds, active subset, step limiter")
    r = run_both(s, softening=0.02, active=active, step=step)
    assert_same(*r)
    go = r[1]
    assert np.isfinite(go["wake"].view(np.float64)).any()
    ev = (active == 1) & (s["kind"] == ics.KIND_GAS) & (s["alive"] == 1)
    assert (go["ax"][~ev] == 7.0).all()


def test_coincident_bodies_and_gas_params(gpu):
    s = ics.gas_cloud(2000, 7)
    for k in ("px", "py", "pz"):
        s[k][100:140] = s[k][99]  # 41 bodies at one point: r2 == 0 pairs, equal keys
    say("[hydro] coincident bodies, gamma 1.4, 32 neighbours, alpha 0.5")
    assert_same(*run_both(s, softening=0.01, gas=GasParams(gamma=1.4, neighbors=32, alpha=0.5)))


def test_lattice_and_few_gas(gpu):
    say("[hydro] lattice 12^3")
    assert_same(*run_both(ics.gas_lattice(12, 0.1)))
    s = ics.gas_cloud(10, 8)
    s["kind"][1:] = ics.KIND_STAR
    say("[hydro] one gas body")
    gs, go, so, oo, evals, hy = run_both(s)
    assert go["dt"][0] == np.inf and go["ax"][0] == 0 and (go["ax"][1:] == 7.0).all()
    assert np.array_equal(gs["h"], so["h"])


def test_host_entry_point_equals_device(gpu):
    s = ics.gas_cloud(4000, 10)
    say("[hydro] host arrays")
    gs, go, so, oo, evals, hy = run_both(s, softening=0.01)
    hs = {k: v.copy() for k, v in s.items()}
    ho = Hydro().forces_host(hs, softening=0.01)
    for k in OUT:
        assert np.array_equal(ho[k], go[k]), k
    assert np.array_equal(hs["h"], gs["h"]) and np.array_equal(hs["rho"], gs["rho"])


def test_degenerate_inputs(gpu):
    """All gas at one point (equal keys, r2 = 0 everywhere), exactly two gas
    bodies, dead gas, and an empty active set."""
    s = ics.gas_cloud(40, 11)
    for k in ("px", "py", "pz"):
        s[k][:] = 0.5
    say("[hydro] 40 bodies at one point")
    assert_same(*run_both(s, softening=0.01))
    s = ics.gas_cloud(50, 12)
    s["kind"][2:] = ics.KIND_STAR
    say("[hydro] two gas bodies")
    assert_same(*run_both(s, softening=0.01))
    s = ics.gas_cloud(3000, 13)
    s["alive"][::5] = 0
    say("[hydro] dead gas and an empty active set")
    r = run_both(s, softening=0.01, active=np.zeros(3000, np.uint8))
    assert_same(*r)
    assert (r[1]["ax"] == 7.0).all() and r[5].stats()["pairs"] == 0
