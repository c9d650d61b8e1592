"""The CPU-or-GPU choice (gravitas_amd.calibrate): perf.Calibrate's protocol
(calibrate.go:41-65: one warm-up, best of three, the faster path kept, seconds
per N log2 N) and engine.gpu()'s modes (engine.go:147-155, galaxy_set_gpu
main.go:483-488), with stand-in force calls and a stepped clock (CPU only)."""
import math

from gravitas_amd import calibrate as cal


class Clock:
    """A clock that the stand-in force calls advance by their cost."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def timed(clock, costs):
    """A force call that takes costs[k] on its k-th call (the last repeats)."""
    calls = []

    def fn():
        clock.t += costs[min(len(calls), len(costs) - 1)]
        calls.append(1)
    return fn, calls


def test_faster_gpu_is_chosen_after_warm_up_and_best_of_three():
    clk = Clock()
    cpu, cpu_calls = timed(clk, [9.0, 2.0, 1.5, 1.75])
    gpu, gpu_calls = timed(clk, [5.0, 0.375, 0.25, 0.3125])  # binary fractions: exact differences
    c = cal.calibrate(cpu, gpu, bodies=24000, gpu_available=lambda: None, clock=clk)
    assert len(cpu_calls) == 4 and len(gpu_calls) == 4  # one warm-up, three timed
    assert c.use_gpu and c.cpu_seconds == 1.5 and c.gpu_seconds == 0.25
    assert c.sec_per_nl == 0.25 / (24000 * math.log2(24000))
    assert c.bodies_per_star == 24000 / 20000


def test_slower_gpu_keeps_the_cpu():
    clk = Clock()
    cpu, _ = timed(clk, [1.0])
    gpu, _ = timed(clk, [2.0])
    c = cal.calibrate(cpu, gpu, bodies=20000, gpu_available=lambda: None, clock=clk)
    assert not c.use_gpu and c.sec_per_nl == 1.0 / cal.nlogn(20000)


def test_unavailable_or_disallowed_gpu_is_never_timed():
    for allowed, avail in ((False, None), (True, "no device")):
        clk = Clock()
        cpu, _ = timed(clk, [1.0])
        gpu, gpu_calls = timed(clk, [0.1])
        c = cal.calibrate(cpu, gpu, bodies=20000, gpu_allowed=allowed, gpu_available=lambda: avail, clock=clk)
        assert not c.use_gpu and c.gpu_seconds is None and not gpu_calls


def test_engine_gpu_modes():
    yes = cal.Calibration(True, 1e-9, 1.0, 1.0, 0.1)
    no = cal.Calibration(False, 1e-9, 1.0, 0.1, 1.0)
    up, down = (lambda: None), (lambda: "no device")
    assert cal.use_gpu(cal.GPU_AUTO, yes, up) and not cal.use_gpu(cal.GPU_AUTO, no, up)
    assert not cal.use_gpu(cal.GPU_AUTO, yes, down) and not cal.use_gpu(cal.GPU_AUTO, None, up)
    assert cal.use_gpu(cal.GPU_ON, no, up) and not cal.use_gpu(cal.GPU_ON, no, down)
    assert not cal.use_gpu(cal.GPU_OFF, yes, up)
    # out-of-range modes are clamped, as galaxy_set_gpu does
    assert not cal.use_gpu(7, yes, up) and cal.use_gpu(-3, yes, up)
