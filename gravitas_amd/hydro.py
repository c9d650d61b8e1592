"""Host-side mirror of the reference's hydro interface.

``Hydro`` mirrors ``sph.Solver`` (internal/sph/sph.go:11-19) and its
``Forces(s, p, active, step, out)`` (sph.go:83-263), the ``sim.Hydro``
provider the engine calls beside gravity (internal/sim/sim.go:45-58).  The
work runs in libgravitas_hip.so on the GPU (gx_hydro_*); states live in HBM.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .hip import DeviceArray


@dataclass
class GasParams:
    """The fields of sim.GasParams the hydro reads (internal/sim/sim.go:26-29);
    0 takes the defaults 5/3, 40, 1."""
    gamma: float = 0.0
    neighbors: float = 0.0
    alpha: float = 0.0


class GasState:
    """Bodies in HBM with the gas fields of sim.State: float64 px..vz, m, u,
    h, rho; uint8 kind, alive."""

    FIELDS = ("px", "py", "pz", "vx", "vy", "vz", "m", "u", "h", "rho")

    def __init__(self, **arrays):
        for k, v in arrays.items():
            setattr(self, k, v)

    @classmethod
    def from_host(cls, s: dict):
        n = len(s["px"])
        out = {}
        for f in cls.FIELDS:
            a = s.get(f)
            out[f] = DeviceArray.from_numpy(np.zeros(n) if a is None else np.asarray(a, np.float64))
        kind, alive = s.get("kind"), s.get("alive")
        out["kind"] = DeviceArray.from_numpy(np.full(n, 3, np.uint8) if kind is None else np.asarray(kind, np.uint8))
        out["alive"] = DeviceArray.from_numpy(np.ones(n, np.uint8) if alive is None else np.asarray(alive, np.uint8))
        return cls(**out)

    def __len__(self):
        return len(self.px)

    def host(self) -> dict:
        return {f: getattr(self, f).numpy() for f in self.FIELDS + ("kind", "alive")}


class HydroOut:
    """sim.HydroOut (internal/sim/sim.go:60-68) in HBM: ax, ay, az, dudt,
    divv, dt (float64) and wake (float64 bits as uint64, +Inf at the start)."""

    FIELDS = ("ax", "ay", "az", "dudt", "divv", "dt")

    def __init__(self, n, fill=0.0):
        for f in self.FIELDS:
            setattr(self, f, DeviceArray.from_numpy(np.full(n, fill)))
        self.wake = DeviceArray.from_numpy(np.full(n, np.inf).view(np.uint64))

    def host(self) -> dict:
        d = {f: getattr(self, f).numpy() for f in self.FIELDS}
        d["wake"] = self.wake.numpy()
        return d


def _ptr(a):
    return None if a is None else a.ptr


class Hydro:
    """sph.Solver on the GPU: the neighbour tree and the SPH forces."""

    _EXPORT = {"keys": (0, np.uint64), "order": (1, np.int32), "nodes": (2, None)}
    NODE_DTYPE = np.dtype([("minx", "f8"), ("miny", "f8"), ("minz", "f8"), ("maxx", "f8"), ("maxy", "f8"),
                           ("maxz", "f8"), ("hmax", "f8"), ("lo", "i4"), ("hi", "i4"), ("first", "i4"),
                           ("second", "i4"), ("escape", "i4"), ("parent", "i4")])

    def __init__(self, timing: bool = False):
        self._L = _lib.load()
        h = self._L.gx_hydro_new()
        if not h:
            raise _lib.GxError(_lib.GX_ERR_NO_DEVICE, self._L.gx_error().decode(errors="replace"))
        self._h = ctypes.c_void_p(h)
        _lib.check(self._L.gx_hydro_config(self._h, int(timing)))

    def __del__(self):
        try:
            if self._h:
                self._L.gx_hydro_free(self._h)
                self._h = None
        except Exception:
            pass

    def set_stream(self, stream):
        _lib.check(self._L.gx_hydro_set_stream(self._h, ctypes.c_void_p(stream)))

    def forces(self, s: GasState, softening: float = 0.0, gas: GasParams | None = None, active=None, step=None,
               out: HydroOut | None = None) -> HydroOut:
        """Forces(s, p, active, step, out) on the device state; h and rho of
        s are updated in place.  active: uint8 per body (None: all), step:
        float64 per body (None: no limiter); host arrays are uploaded."""
        gas = gas or GasParams()
        gp = _lib.GxGasParams(gas.gamma, gas.neighbors, gas.alpha)
        if out is None:
            out = HydroOut(len(s))
        if active is not None and not isinstance(active, DeviceArray):
            active = DeviceArray.from_numpy(np.ascontiguousarray(active, np.uint8))
        if step is not None and not isinstance(step, DeviceArray):
            step = DeviceArray.from_numpy(np.ascontiguousarray(step, np.float64))
        _lib.check(self._L.gx_hydro_forces_dev(
            self._h, s.px.ptr, s.py.ptr, s.pz.ptr, s.vx.ptr, s.vy.ptr, s.vz.ptr, s.m.ptr, s.kind.ptr, s.alive.ptr,
            s.u.ptr, s.h.ptr, s.rho.ptr, len(s), ctypes.byref(gp), float(softening), _ptr(active), _ptr(step),
            out.ax.ptr, out.ay.ptr, out.az.ptr, out.dudt.ptr, out.divv.ptr, out.dt.ptr, out.wake.ptr))
        return out

    def forces_host(self, s: dict, softening: float = 0.0, gas: GasParams | None = None, active=None, step=None,
                    out: dict | None = None) -> dict:
        """The same on host arrays (gx_hydro_forces): s['h'] and s['rho'] are
        updated in place; returns out (ax, ay, az, dudt, divv, dt, wake)."""
        gas = gas or GasParams()
        gp = _lib.GxGasParams(gas.gamma, gas.neighbors, gas.alpha)
        n = len(s["px"])
        for k in ("px", "py", "pz", "vx", "vy", "vz", "m", "u", "h", "rho"):
            s[k] = np.ascontiguousarray(s[k], np.float64)
        kind = np.ascontiguousarray(s["kind"], np.uint8)
        alive = np.ascontiguousarray(s.get("alive", np.ones(n, np.uint8)), np.uint8)
        if out is None:
            out = {k: np.zeros(n) for k in HydroOut.FIELDS}
        if "wake" not in out:
            out["wake"] = np.full(n, np.inf).view(np.uint64).copy()
        act = None if active is None else np.ascontiguousarray(active, np.uint8)
        st = None if step is None else np.ascontiguousarray(step, np.float64)
        p = lambda a: None if a is None else a.ctypes.data
        _lib.check(self._L.gx_hydro_forces(
            self._h, *[p(s[k]) for k in ("px", "py", "pz", "vx", "vy", "vz", "m")], p(kind), p(alive), p(s["u"]),
            p(s["h"]), p(s["rho"]), n, ctypes.byref(gp), float(softening), p(act), p(st),
            *[p(out[k]) for k in HydroOut.FIELDS], p(out["wake"])))
        return out

    def stats(self) -> dict:
        st = _lib.GxHydroStats()
        _lib.check(self._L.gx_hydro_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _lib.GxHydroStats._fields_}

    def phase_ms(self) -> dict:
        buf = (ctypes.c_float * 6)()
        self._L.gx_hydro_phase_ms(self._h, buf, 6)
        names = ("sort", "tree", "density", "terms", "force", "total")
        return dict(zip(names, list(buf)))

    def export(self, what: str) -> np.ndarray:
        """Copy part of the last evaluation's tree to host (test hook)."""
        code, dt = self._EXPORT[what]
        dt = dt or self.NODE_DTYPE
        st = self.stats()
        count = st["n_nodes"] if what == "nodes" else st["n_gas"]
        out = np.zeros(count, dt)
        if count:
            got = self._L.gx_hydro_export(self._h, code, out.ctypes.data, count)
            if got < 0:
                _lib.check(_lib.GX_ERR_ARG)
            out = out[:got]
        return out
