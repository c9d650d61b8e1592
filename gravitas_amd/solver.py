"""Host-side mirror of the reference's force-provider interface.

``Solver`` mirrors ``bh.Solver`` (internal/sim/bh/bh.go:29-38) with its
``Accelerations`` (bh.go:76-102) and ``PotentialEnergy`` (potential.go:12-42);
``gpu_available`` / ``gpu_name`` / ``last_gpu_seconds`` mirror
``GPUAvailable`` / ``GPUName`` / ``LastGPUSeconds`` (gpu_darwin.go:132-155).
The computation runs in libgravitas_hip.so on the GPU; states live in HBM as
hipMalloc'd arrays (gravitas_amd.hip.DeviceArray).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .hip import DeviceArray

KIND_STAR, KIND_HOLE, KIND_DARK, KIND_GAS = 0, 1, 2, 3
DEFAULT_THETA = 0.9


@dataclass
class Params:
    """The subset of sim.Params (internal/sim/sim.go:9-23) the solver reads."""
    softening: float = 0.0
    dark_matter_off: bool = False
    dt: float = 0.0
    c: float = 0.0  # speed of light; <= 0 is Newtonian (used by the pairs with holes and their drifts)


_avail = None


def gpu_available():
    """None if the GPU backend can be used, else the reason (GPUAvailable)."""
    global _avail
    if _avail is None:
        try:
            L = _lib.load()
        except (RuntimeError, OSError) as e:
            _avail = str(e)
            return _avail
        buf = ctypes.create_string_buffer(512)
        rc = L.gx_init(buf, 512)
        _avail = None if rc == _lib.GX_OK else f"hip: {buf.value.decode()} (code {rc})"
    return _avail


def accelerations_noalloc(s: "DeviceState", p: "Params", theta: float = 0.0, walk: str = "fp32", stream=None):
    """Accelerations of s (at most GX_NOALLOC_MAX_BODIES bodies) through the
    entry point that allocates no device memory (gx_accelerations_noalloc):
    every buffer is a slice of the library's static per-GPU scratch."""
    L = _lib.load()
    w = {"fp32": _lib.GX_WALK_FP32, "fp64": _lib.GX_WALK_FP64, "fp32-q16": _lib.GX_WALK_FP32_Q16}[walk]
    _lib.check(L.gx_accelerations_noalloc(
        _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), _ptr(s.alive), len(s), float(theta),
        float(p.softening), int(p.dark_matter_off), w, _ptr(s.ax), _ptr(s.ay), _ptr(s.az),
        ctypes.c_void_p(stream) if stream else None))


def gpu_name() -> str:
    if gpu_available() is not None:
        return ""
    buf = ctypes.create_string_buffer(256)
    _lib.load().gx_name(buf, 256)
    return buf.value.decode()


class DeviceState:
    """Struct-of-arrays state resident in HBM (sim.State, internal/sim/state.go:33-52):
    float64 px, py, pz, vx, vy, vz, ax, ay, az, m; uint8 kind, alive; and,
    when the host state has "jx" (spins, state.go:48-49), float64 jx, jy, jz."""

    FIELDS = ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay", "az", "m")
    SPIN = ("jx", "jy", "jz")

    def __init__(self, **arrays):
        for k, v in arrays.items():
            setattr(self, k, v)

    @classmethod
    def from_host(cls, s: dict):
        n = len(s["px"])
        out = {}
        for f in cls.FIELDS:
            a = s.get(f)
            out[f] = DeviceArray.from_numpy(np.zeros(n) if a is None else np.asarray(a, dtype=np.float64))
        kind, alive = s.get("kind"), s.get("alive")
        out["kind"] = DeviceArray.from_numpy(np.zeros(n, np.uint8) if kind is None else np.asarray(kind, np.uint8))
        out["alive"] = DeviceArray.from_numpy(np.ones(n, np.uint8) if alive is None else np.asarray(alive, np.uint8))
        if "jx" in s:
            for f in cls.SPIN:
                out[f] = DeviceArray.from_numpy(np.asarray(s[f], dtype=np.float64))
        return cls(**out)

    @property
    def has_spin(self):
        return hasattr(self, "jx")

    def __len__(self):
        return len(self.px)

    def view(self, start, count):
        """The bodies [start, start+count) as a state of their own (no copy)."""
        extra = self.SPIN if self.has_spin else ()
        return DeviceState(**{f: getattr(self, f).view(start, count) for f in self.FIELDS + ("kind", "alive") + extra})

    def acc_host(self):
        return np.stack([self.ax.numpy(), self.ay.numpy(), self.az.numpy()], axis=1)


def _ptr(t):
    return t.ptr


class Retard:
    """The holes' histories of one simulation for retarded gravity
    (Simulation.histories, internal/sim/retarded.go:62-209) in device memory,
    kept across split steps."""

    def __init__(self):
        self._L = _lib.load()
        h = self._L.gx_retard_new()
        if not h:
            raise MemoryError("gx_retard_new")
        self._h = ctypes.c_void_p(h)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._L.gx_retard_free(self._h)
                self._h = None
        except Exception:
            pass

    def __len__(self):
        return self._L.gx_retard_count(self._h)

    def hole_appeared(self, solver: "Solver", s: DeviceState, p: Params, i: int, time: float, ids=None):
        """holeAppeared (retarded.go:157-161; AddBody of a hole, sim.go:236-243):
        hole i, added at ``time``, starts its history with its current state."""
        _lib.check(self._L.gx_retard_hole_appeared_dev(
            solver._h, self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz),
            None if ids is None else _ptr(ids), len(s), int(i), float(time), float(p.dt), float(p.c)))


class Solver:
    """bh.Solver on the GPU.

    theta: opening angle, 0 means DefaultTheta (bh.go:69-74).
    monopole_only: drop the quadrupole term (bh.go:33-35).
    walk: "fp32" (group walk, the default), "fp32-q16" (the same with
          FP16-stored quadrupoles) or "fp64" (per-particle walk, bit-identical
          to the reference's CPU walk).
    """

    def __init__(self, theta: float = 0.0, monopole_only: bool = False, walk: str = "fp32", timing: bool = False,
                 group_size: int = 64, persistent: int = 1, stack: int = 384, parts: int = 384,
                 waves: int = 5):
        reason = gpu_available()
        if reason is not None:
            raise RuntimeError(f"GPU backend unavailable: {reason}")
        self._L = _lib.load()
        h = self._L.gx_solver_new()
        if not h:
            raise _lib.GxError(_lib.GX_ERR_HIP, self._L.gx_error().decode())
        self._h = ctypes.c_void_p(h)
        self.theta = theta
        self.monopole_only = monopole_only
        self.walk = walk
        self.timing = timing
        self.group_size = group_size
        self.persistent = persistent
        self.stack = stack
        self.parts = parts
        self.waves = waves

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._L.gx_solver_free(self._h)
                self._h = None
        except Exception:
            pass

    def _configure(self):
        w = {"fp32": _lib.GX_WALK_FP32, "fp64": _lib.GX_WALK_FP64, "fp32-q16": _lib.GX_WALK_FP32_Q16}[self.walk]
        _lib.check(self._L.gx_solver_config(self._h, float(self.theta), int(self.monopole_only), w,
                                            int(self.timing)))
        _lib.check(self._L.gx_solver_tune_walk(self._h, int(self.group_size), int(self.persistent), int(self.stack),
                                               int(self.parts), int(self.waves)))

    def autotune(self, s: DeviceState, offsets, theta, eps, reps: int = 3, dark_matter_off=False) -> list:
        """The library's walk autotuner on the systems packed in s (see
        gx_solver_autotune): keeps the fastest candidate (group size, work
        queue, stack and staging sizes) for this solver and returns the
        timing table; s.ax/ay/az hold the chosen walk's forces."""
        self._configure()
        keep, off, b, th, ep = self._batch_args(offsets, theta, eps)
        table = (_lib.GxTuneEntry * 64)()
        n = ctypes.c_int32()
        _lib.check(self._L.gx_solver_autotune(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), _ptr(s.alive), off, b, th, ep,
            int(dark_matter_off), _ptr(s.ax), _ptr(s.ay), _ptr(s.az), int(reps), table, 64, ctypes.byref(n)))
        rows = [{f: getattr(table[i], f) for f, _ in _lib.GxTuneEntry._fields_} for i in range(min(n.value, 64))]
        if rows:
            best = min(rows, key=lambda r: r["ms"])
            self.group_size, self.persistent = best["group_size"], best["persistent"]
            self.stack, self.parts, self.waves = best["stack"], best["parts"], best["waves"]
        return rows

    def reserve(self, nodes=0, full=0, partial=0, leaves=0):
        """Starting capacities of the node arena and the walk lists (0 keeps
        one); they still grow on demand."""
        _lib.check(self._L.gx_solver_reserve(self._h, int(nodes), int(full), int(partial), int(leaves)))

    def set_stream(self, stream):
        """Launch on the given hipStream_t (a pointer-sized int or None)."""
        _lib.check(self._L.gx_solver_set_stream(self._h, ctypes.c_void_p(stream) if stream else None))

    # -- Accelerations (bh.go:76-102) ------------------------------------------------
    def accelerations(self, s: DeviceState, p: Params):
        """Fill s.ax/ay/az (device tensors) from star-star gravity."""
        self._configure()
        n = len(s)
        _lib.check(self._L.gx_accelerations_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), _ptr(s.alive), n,
            float(p.softening), int(p.dark_matter_off), _ptr(s.ax), _ptr(s.ay), _ptr(s.az)))

    def enqueue(self, s: DeviceState, p: Params):
        """Asynchronous Accelerations; call sync() before reading results."""
        self._configure()
        _lib.check(self._L.gx_enqueue_accelerations_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), _ptr(s.alive), len(s),
            float(p.softening), int(p.dark_matter_off), _ptr(s.ax), _ptr(s.ay), _ptr(s.az)))

    def sync(self):
        _lib.check(self._L.gx_sync(self._h))

    def accelerations_host(self, x, y, z, m, kind=None, alive=None, eps=0.0, dm_off=False, out=None):
        """Accelerations on host (numpy) arrays; the data crosses PCIe."""
        self._configure()
        n = len(x)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        x, y, z, m = f64(x), f64(y), f64(z), f64(m)
        kind = np.zeros(n, np.uint8) if kind is None else np.ascontiguousarray(kind, np.uint8)
        alive = np.ones(n, np.uint8) if alive is None else np.ascontiguousarray(alive, np.uint8)
        ax, ay, az = (np.zeros(n), np.zeros(n), np.zeros(n)) if out is None else out
        c = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _lib.check(self._L.gx_accelerations(self._h, c(x), c(y), c(z), c(m), c(kind), c(alive), n, float(eps),
                                            int(dm_off), c(ax), c(ay), c(az)))
        return np.stack([ax, ay, az], axis=1)

    # -- batches of independent systems -------------------------------------------------
    @staticmethod
    def _batch_args(offsets, theta, eps):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        b = len(off) - 1
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(theta, dtype=np.float64), (b,)))
        ep = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, dtype=np.float64), (b,)))
        keep = (off, th, ep)
        c = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        return keep, c(off), b, c(th), c(ep)

    def batch_accelerations(self, s: DeviceState, offsets, theta, eps, dark_matter_off=False):
        """Accelerations of independent systems packed in s: system b owns
        bodies [offsets[b], offsets[b+1]); theta and eps per system (or one value)."""
        self._configure()
        keep, off, b, th, ep = self._batch_args(offsets, theta, eps)
        _lib.check(self._L.gx_batch_accelerations_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), _ptr(s.alive), off, b, th, ep,
            int(dark_matter_off), _ptr(s.ax), _ptr(s.ay), _ptr(s.az)))

    def batch_step(self, s: DeviceState, offsets, theta, eps, dt, acc_valid=True, dark_matter_off=False):
        self._configure()
        keep, off, b, th, ep = self._batch_args(offsets, theta, eps)
        _lib.check(self._L.gx_batch_step_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.m),
            _ptr(s.kind), _ptr(s.alive), off, b, th, ep, int(dark_matter_off), float(dt), int(bool(acc_valid)),
            _ptr(s.ax), _ptr(s.ay), _ptr(s.az)))

    def graph_capture_step(self, s: DeviceState, offsets, theta, eps, dt, dark_matter_off=False):
        """Capture one batched KDK step as a hipGraph (after one eager force
        evaluation); replay it with graph_launch."""
        self._configure()
        keep, off, b, th, ep = self._batch_args(offsets, theta, eps)
        _lib.check(self._L.gx_graph_capture_step(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.m),
            _ptr(s.kind), _ptr(s.alive), off, b, th, ep, int(dark_matter_off), float(dt),
            _ptr(s.ax), _ptr(s.ay), _ptr(s.az)))

    def graph_launch(self, steps: int = 1):
        _lib.check(self._L.gx_graph_launch(self._h, int(steps)))

    # -- KDK step (sim.go:262-281, 302-343) -------------------------------------------
    def step(self, s: DeviceState, p: Params, acc_valid: bool = True):
        """One kick-drift-kick leapfrog step of p.dt on the device state; with
        acc_valid False the accelerations are computed first (initForces)."""
        self._configure()
        _lib.check(self._L.gx_step_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.m),
            _ptr(s.kind), _ptr(s.alive), len(s), float(p.softening), int(p.dark_matter_off), float(p.dt),
            int(bool(acc_valid)), _ptr(s.ax), _ptr(s.ay), _ptr(s.az)))

    # -- active sets and block steps ----------------------------------------------------
    def accelerations_active(self, s: DeviceState, p: Params, active):
        """Accelerations of the bodies with active[i] != 0 only (active: uint8
        DeviceArray or host array in state order); the others keep theirs."""
        self._configure()
        if not isinstance(active, DeviceArray):
            active = DeviceArray.from_numpy(np.ascontiguousarray(active, dtype=np.uint8))
        _lib.check(self._L.gx_accelerations_active_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), _ptr(s.alive), _ptr(active),
            len(s), float(p.softening), int(p.dark_matter_off), _ptr(s.ax), _ptr(s.ay), _ptr(s.az)))

    def block_step(self, s: DeviceState, p: Params, T: float, levels: int = 4, eta: float = 0.025,
                   acc_valid: bool = True) -> dict:
        """One global step of length T with power-of-two block time steps
        (rungs 0..levels-1); returns the rung populations and active counts."""
        self._configure()
        st = _lib.GxBlockStats()
        _lib.check(self._L.gx_block_step_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.m),
            _ptr(s.kind), _ptr(s.alive), len(s), float(p.softening), int(p.dark_matter_off), float(T), int(levels),
            float(eta), int(bool(acc_valid)), _ptr(s.ax), _ptr(s.ay), _ptr(s.az), ctypes.byref(st)))
        ev = min(st.events, _lib.GX_BLOCK_MAX_TICKS)
        return {"levels": st.levels, "events": st.events, "bodies": st.bodies, "targets": st.targets,
                "rung_count": list(st.rung_count)[:st.levels], "tick": list(st.tick)[:ev],
                "active": list(st.active)[:ev]}

    def split_step(self, s: DeviceState, p: Params, acc_valid: bool = True, retard: "Retard | None" = None,
                   time: float = 0.0, gravity: str = "instant", ids=None) -> int:
        """One step of p.dt with black holes (internal/sim/split.go:164-199):
        tree forces between stars, direct sums for every pair with a hole,
        power-of-two substeps for the holes and the stars near them; p.c > 0
        switches on the Paczynski-Wiita pairs and relativistic drifts, and, if
        the state carries spins, frame dragging (framedrag.go).  With
        ``retard`` (a Retard) the holes pull in the ``gravity`` mode
        (retarded.go: "instant", "retarded", "retarded+v") at simulation time
        ``time``, and the holes are recorded at time + p.dt after the step
        (sim.go:285-287); ``ids`` (a uint64 DeviceArray, or None: the state
        index) are the bodies' IDs.  Returns the number of substeps."""
        self._configure()
        sub = ctypes.c_int32(0)
        head = (self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.m),
                _ptr(s.kind), _ptr(s.alive))
        tail = (len(s), float(p.softening), int(p.dark_matter_off), float(p.dt), float(p.c), int(bool(acc_valid)),
                _ptr(s.ax), _ptr(s.ay), _ptr(s.az), ctypes.byref(sub))
        if retard is not None:
            spins = (_ptr(s.jx), _ptr(s.jy), _ptr(s.jz)) if s.has_spin else (None, None, None)
            _lib.check(self._L.gx_split_step_retarded_dev(
                *head, *spins, None if ids is None else _ptr(ids), *tail[:-1], retard._h, float(time),
                _lib.GX_GRAVITY[gravity], tail[-1]))
        elif s.has_spin:
            _lib.check(self._L.gx_split_step_spin_dev(*head, _ptr(s.jx), _ptr(s.jy), _ptr(s.jz), *tail))
        else:
            _lib.check(self._L.gx_split_step_dev(*head, *tail))
        return sub.value

    def source_pos(self, retard: "Retard", s: DeviceState, p: Params, k: int, obs, time: float, gravity: str,
                   ids=None):
        """sourcePos (retarded.go:118-155) of alive hole number k for the
        observers ``obs`` (host array, shape (m, 3)) at simulation time
        ``time``: an (m, 3) array of source positions and a bool array, false
        where the hole's gravity has not arrived."""
        obs = np.ascontiguousarray(np.atleast_2d(np.asarray(obs, np.float64)))
        m = len(obs)
        o = [DeviceArray.from_numpy(np.ascontiguousarray(obs[:, j])) for j in range(3)]
        out = [DeviceArray(m, np.float64) for _ in range(3)]
        vis = DeviceArray(m, np.uint8)
        self._configure()
        _lib.check(self._L.gx_retard_source_pos_dev(
            self._h, retard._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.kind), _ptr(s.alive),
            None if ids is None else _ptr(ids), len(s), int(k), *(_ptr(a) for a in o), m, float(time), float(p.dt),
            float(p.c), _lib.GX_GRAVITY[gravity], *(_ptr(a) for a in out), _ptr(vis)))
        return np.stack([a.numpy() for a in out], axis=1), vis.numpy().astype(bool)

    def drag_outer(self, s: DeviceState, p: Params, dt: float, near=None):
        """Simulation.dragOuter (framedrag.go:115-127) over time dt: every
        alive body that is neither a hole nor flagged in ``near`` (a uint8
        DeviceArray, or None) is dragged in state order; velocities and spins
        change in place."""
        self._configure()
        _lib.check(self._L.gx_drag_outer_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.m),
            _ptr(s.kind), _ptr(s.alive), None if near is None else _ptr(near), _ptr(s.jx), _ptr(s.jy), _ptr(s.jz),
            len(s), float(p.softening), float(dt), float(p.c)))

    def drag_fast(self, s: DeviceState, p: Params, dt: float, near_list):
        """Simulation.dragFast (framedrag.go:129-143): the bodies of
        ``near_list`` (an int32 DeviceArray of state indices) in order, then
        the holes."""
        self._configure()
        _lib.check(self._L.gx_drag_fast_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.m),
            _ptr(s.kind), _ptr(s.alive), _ptr(near_list) if len(near_list) else None, len(near_list), _ptr(s.jx),
            _ptr(s.jy), _ptr(s.jz), len(s), float(p.softening), float(dt), float(p.c)))

    def drag_stats(self) -> dict:
        st = _lib.GxDragStats()
        _lib.check(self._L.gx_drag_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def diagnostics(self, s: DeviceState, p: Params) -> dict:
        """FP64 kinetic energy of the sources, momentum, angular momentum,
        mass and centre of mass (internal/sim/diag.go) summed on the GPU."""
        out = (ctypes.c_double * _lib.GX_DIAG_VALUES)()
        _lib.check(self._L.gx_diagnostics_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.m),
            _ptr(s.kind), len(s), int(p.dark_matter_off), out))
        v = list(out)
        m = v[7]
        com = [c * (1 / m) for c in v[8:11]] if m != 0 else [0.0, 0.0, 0.0]  # c.Scale(1 / m)
        return {"kinetic": v[0], "momentum": v[1:4], "angular_momentum": v[4:7], "mass": m, "com": com}

    def kick(self, s: DeviceState, h: float):
        """v += h a over every body (enqueued on the solver's stream)."""
        _lib.check(self._L.gx_kick_dev(self._h, _ptr(s.vx), _ptr(s.vy), _ptr(s.vz), _ptr(s.ax), _ptr(s.ay),
                                       _ptr(s.az), len(s), float(h)))

    def drift(self, s: DeviceState, dt: float):
        """x += dt v over every body (enqueued on the solver's stream)."""
        _lib.check(self._L.gx_drift_dev(self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.vx), _ptr(s.vy),
                                        _ptr(s.vz), len(s), float(dt)))

    def host_state(self, s: DeviceState) -> dict:
        extra = DeviceState.SPIN if s.has_spin else ()
        return {f: getattr(s, f).numpy() for f in DeviceState.FIELDS + extra}

    # -- PotentialEnergy (potential.go:12-42) ----------------------------------------
    def potential_energy(self, s: DeviceState, p: Params) -> float:
        self._configure()
        out = ctypes.c_double()
        _lib.check(self._L.gx_potential_energy_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), _ptr(s.alive), len(s),
            float(p.softening), int(p.dark_matter_off), float(p.c), ctypes.byref(out)))
        return out.value

    def potential_exact(self, s: DeviceState, p: Params) -> float:
        """PotentialEnergyParallel (diag.go:44-64): the exact pair-model
        potential over every pair, FP64 on the GPU."""
        out = ctypes.c_double()
        _lib.check(self._L.gx_potential_exact_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), len(s), float(p.softening),
            int(p.dark_matter_off), float(p.c), ctypes.byref(out)))
        return out.value

    # -- FP64 direct-sum oracle kernel (direct.go:25-48) ------------------------------
    def direct_sample(self, s: DeviceState, p: Params, targets):
        """FP64 direct-sum accelerations of the given target indices (host array
        or int64 DeviceArray); returns an (nt, 3) float64 DeviceArray."""
        if not isinstance(targets, DeviceArray):
            t = np.asarray(targets, dtype=np.int64)
            if len(t) and (t.min() < 0 or t.max() >= len(s)):
                raise IndexError("target index out of range")
            targets = DeviceArray.from_numpy(t)
        nt = len(targets)
        out = DeviceArray((nt, 3), np.float64)
        _lib.check(self._L.gx_direct_sample_dev(
            self._h, _ptr(s.px), _ptr(s.py), _ptr(s.pz), _ptr(s.m), _ptr(s.kind), len(s), _ptr(targets), nt,
            float(p.softening), int(p.dark_matter_off), _ptr(out)))
        return out

    # -- diagnostics -------------------------------------------------------------------
    def last_gpu_seconds(self) -> float:
        """LastGPUSeconds (gpu_darwin.go:155)."""
        return float(self._L.gx_last_seconds(self._h))

    def stats(self) -> dict:
        st = _lib.GxStats()
        _lib.check(self._L.gx_get_stats(self._h, ctypes.byref(st)))
        out = {f: getattr(st, f) for f, _ in st._fields_}
        out["max_list"] = list(st.max_list)[:3]
        return out

    _EXPORT = {"keys": (0, np.uint64, ()), "order": (1, np.int32, ()), "spos": (2, np.float64, (4,)),
               "nodes": (3, None, ()), "mult": (4, np.float64, (8,))}
    NODE_DTYPE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("crit2", "f8"), ("first", "i4"),
                           ("next", "i4"), ("lo", "i4"), ("hi", "i4"), ("nchild", "i4"), ("level", "i4"),
                           ("parent", "i4"), ("pad", "i4")])

    def export(self, what: str) -> np.ndarray:
        """Copy part of the last build to host (test hook): keys, order, spos,
        nodes, mult."""
        code, dt, shape = self._EXPORT[what]
        st = self.stats()
        count = st["n_nodes"] if what in ("nodes", "mult") else st["n_stars"]
        dt = self.NODE_DTYPE if dt is None else np.dtype(dt)
        buf = np.zeros((max(count, 0),) + shape, dtype=dt)
        k = self._L.gx_export(self._h, code, buf.ctypes.data_as(ctypes.c_void_p), count)
        if k < 0:
            raise _lib.GxError(_lib.GX_ERR_ARG, self._L.gx_error().decode())
        return buf[:k]

    def phase_ms(self) -> dict:
        """GPU milliseconds per phase of the last step (timing=True); for the
        FP32 walk, "walk" is traversal + evaluation."""
        buf = (ctypes.c_float * 9)()
        k = self._L.gx_phase_ms(self._h, buf, 9)
        names = ("collect", "keys", "sort", "sorted_copies", "topology", "moments", "walk", "total", "evaluate")
        d = {names[i]: float(buf[i]) for i in range(k)}
        if d.get("evaluate", 0.0) > 0.0:
            d["traverse"] = d["walk"]
            d["walk"] = d["walk"] + d["evaluate"]
        return d
