"""The CPU-or-GPU choice on a real device: perf.Calibrate's protocol
(calibrate.go:41-65) on a 20,000-star state, the oracle standing in for the
host's CPU solver and the library's host entry point (PCIe included, as the
Go binding calls it) for the GPU: the GPU is chosen, and LastGPUSeconds
(gpu_darwin.go:155) of its last step is positive with timing off."""
import sys

import pytest

import oracle
from gravitas_amd import Solver, calibrate, ics

pytestmark = pytest.mark.gpu


def test_calibration_chooses_the_gpu(gpu):
    s = ics.disk_halo(20000, seed=5, eps=0.02)
    o = oracle.Solver(theta=0.9, workers=4)
    g = Solver(theta=0.9, walk="fp32", timing=False)

    def cpu():
        return o.accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"], eps=0.02)

    def gpu_step():
        return g.accelerations_host(s["px"], s["py"], s["pz"], s["m"], s["kind"], eps=0.02)

    print("[calibrate] 20,000 stars, CPU oracle vs GPU", flush=True)
    sys.stdout.flush()
    c = calibrate.calibrate(cpu, gpu_step, bodies=len(s["px"]))
    assert c.use_gpu and c.gpu_seconds < c.cpu_seconds
    assert calibrate.use_gpu(calibrate.GPU_AUTO, c) and calibrate.use_gpu(calibrate.GPU_ON, None)
    assert 0 < g.last_gpu_seconds() <= c.gpu_seconds
