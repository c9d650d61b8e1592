"""Parity at the benchmark sizes (the smaller tests cover the edge cases):

  * gravity: the 1,048,576-body disk+halo model at theta 0.5, eps 0.02 --
    the FP64 walk bit-identical to the oracle, the FP32 walk within the
    reference's GPU bound (gpu_test.go:18) and p99 <= 1e-4, both with the
    oracle's interaction counts;
  * hydro: a 262,144-body gas sphere (h from the analytic density, then
    converged) -- every output bit-identical to the oracle.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from gravitas_amd import DeviceState, GasState, Hydro, Params, Solver, ics
from gravitas_amd.hydro import HydroOut

pytestmark = pytest.mark.gpu
WORKERS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_disk_halo_1m_walks(gpu):
    n = 1048576
    s = ics.disk_halo(n, seed=2, eps=0.02)
    say(f"[full] disk+halo {n}: oracle on {WORKERS} threads")
    want = oracle.Solver(theta=0.5, workers=WORKERS).accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"],
                                                                   eps=0.02)
    ds = DeviceState.from_host(s)
    s64 = Solver(theta=0.5, walk="fp64")
    say("[full] FP64 walk")
    s64.accelerations(ds, Params(softening=0.02))
    assert np.array_equal(ds.acc_host(), want)
    st64 = s64.stats()
    say("[full] FP32 walk")
    s32 = Solver(theta=0.5, walk="fp32")
    s32.accelerations(ds, Params(softening=0.02))
    st32 = s32.stats()
    assert (st32["pp_interactions"], st32["pc_interactions"]) == (st64["pp_interactions"], st64["pc_interactions"])
    assert st32["fallbacks"] == 0
    e = np.sort(ics.rel_errors(ds.acc_host(), want))
    p50, p99, mx = e[len(e) // 2], e[int(0.99 * len(e))], e[-1]
    say(f"[full] FP32 vs FP64: p50 {p50:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p50 <= 1e-5 and p99 <= 1e-4 and mx <= 1e-3


def test_gas_sphere_262k_hydro(gpu):
    n = 262144
    s = ics.gas_sphere(n, 9, h_from_density=True)
    g = GasState.from_host(s)
    hy = Hydro()
    out = HydroOut(n)
    say(f"[full] gas sphere {n}: converge h on the GPU")
    for _ in range(3):
        hy.forces(g, 0.01, out=out)
    start = g.host()
    say("[full] one more evaluation, GPU and oracle")
    hy.forces(g, 0.01, out=out)
    go, gs = out.host(), g.host()
    so = {k: start[k].copy() for k in start}
    oo, evals = oracle.sph_forces(so, softening=0.01)
    for k in ("ax", "ay", "az", "dudt", "divv", "dt"):
        assert np.array_equal(go[k], oo[k]), k
    assert np.array_equal(gs["h"], so["h"]) and np.array_equal(gs["rho"], so["rho"])
    assert hy.stats()["density_sums"] == evals
