"""The engine's CPU-or-GPU choice, mirrored for a host that drives this
backend from Python.

``calibrate`` is ``perf.Calibrate`` (internal/perf/calibrate.go:41-65): one
warm-up call, then the best of three force evaluations of the same state on
the CPU and -- if allowed and available -- on the GPU, keeping the faster, and
the seconds per N log2 N of the winner.  ``use_gpu`` is ``engine.gpu()``
(internal/engine/engine.go:147-155) with the modes of ``galaxy_set_gpu``
(cmd/libgalaxy/main.go:483-488): 0 automatic (the calibration decides), 1 on
(whenever the backend is available), 2 off.

The CPU side is the host's own solver, passed in as a callable: this package
has no CPU force path (and no fallback to one).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Callable, Optional

GPU_AUTO, GPU_ON, GPU_OFF = 0, 1, 2  # galaxy_set_gpu's modes
CALIBRATION_STARS = 20_000           # calibrationStars, calibrate.go:36-38


@dataclass
class Calibration:
    """prefs.Calibration's gravity fields."""
    use_gpu: bool
    sec_per_nl: float      # seconds of the faster path per N log2 N
    bodies_per_star: float
    cpu_seconds: float
    gpu_seconds: Optional[float]


def nlogn(m: float) -> float:
    """calibrate.go:122: m log2 max(2, m)."""
    return m * math.log2(max(2.0, m))


def best_of(fn: Callable[[], object], reps: int = 3, clock=time.perf_counter) -> float:
    """One warm-up call (buffers, caches), then the best of `reps` timed calls."""
    fn()
    best = math.inf
    for _ in range(reps):
        t0 = clock()
        fn()
        best = min(best, clock() - t0)
    return best


def calibrate(cpu_accelerations: Callable[[], object], gpu_accelerations: Optional[Callable[[], object]],
              bodies: int, stars: int = CALIBRATION_STARS, gpu_allowed: bool = True,
              gpu_available: Optional[Callable[[], Optional[str]]] = None, clock=time.perf_counter) -> Calibration:
    """Calibrate (calibrate.go:41-65): time the CPU force evaluation and, if
    allowed and the backend is available, the GPU one; keep the faster.
    `bodies` is the state's length (stars plus dark matter and holes), used
    for the N log2 N scaling and bodies per requested star."""
    if gpu_available is None:
        from .solver import gpu_available as gpu_available  # noqa: PLW0127
    best = best_of(cpu_accelerations, clock=clock)
    cpu = best
    use, gpu_t = False, None
    if gpu_allowed and gpu_accelerations is not None and gpu_available() is None:
        gpu_t = best_of(gpu_accelerations, clock=clock)
        if gpu_t < best:
            best, use = gpu_t, True
    m = float(bodies)
    return Calibration(use_gpu=use, sec_per_nl=best / nlogn(m), bodies_per_star=m / stars, cpu_seconds=cpu,
                       gpu_seconds=gpu_t)


def use_gpu(mode: int, cal: Optional[Calibration], gpu_available: Optional[Callable[[], Optional[str]]] = None
            ) -> bool:
    """engine.gpu() (engine.go:147-155): off never, on when available,
    automatic when the calibration chose the GPU and it is available."""
    if gpu_available is None:
        from .solver import gpu_available as gpu_available  # noqa: PLW0127
    mode = max(GPU_AUTO, min(GPU_OFF, int(mode)))  # galaxy_set_gpu clamps the mode
    if mode == GPU_OFF:
        return False
    if mode == GPU_ON:
        return gpu_available() is None
    return bool(cal and cal.use_gpu) and gpu_available() is None
