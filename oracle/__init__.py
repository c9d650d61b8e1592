"""CPU oracle for the Barnes-Hut force step -- TEST INFRASTRUCTURE ONLY.

This package wraps ``bh_oracle.c``, a C restatement of the gravitas
reference's ``bh.Solver`` (internal/sim/bh/*.go) and ``direct.Solver``
(internal/sim/direct/direct.go).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import it, and only as the
checker or as the timed CPU baseline -- never inside the product path.

Every function here cites the reference lines it follows in ``bh_oracle.c``.

Pinning (what keeps the restatement honest, see tests/test_oracle.py):
  * Morton keys: naive bit loop (bh_test.go:16-31), closed-form cell().
  * Radix sort: equals numpy's stable argsort of (key) (bh_test.go:33-50).
  * Tree: brute-force recursive definition on small inputs in Python; node
    invariants (ranges partition, children in octant order, skip links).
  * Moments: per-node direct sums over the node's stars; trace(Q) = 0.
  * Walk: two-star known answer (bh_test.go:137-147); theta -> 1e-9 equals
    the direct sum to 1e-10 (bh_test.go:88-95); force = gradient of the tree
    potential (bh_test.go:332-351); quadrupole beats monopole in the far field
    (bh_test.go:309-330); coincident stars give no NaN (bh_test.go:149-161).
  * Potential: tiny theta equals the exact pair sum (bh_test.go:280-292).
  * Direct sum: numpy float64 pair sum, Kepler/momentum invariants.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I8 = ctypes.POINTER(ctypes.c_int8)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)

NODE_DTYPE = np.dtype(
    [("x", "f8"), ("y", "f8"), ("z", "f8"), ("m", "f8"),
     ("qxx", "f8"), ("qyy", "f8"), ("qzz", "f8"), ("qxy", "f8"), ("qxz", "f8"), ("qyz", "f8"),
     ("crit2", "f8"), ("first", "i4"), ("next", "i4"), ("lo", "i4"), ("hi", "i4")])

KIND_STAR, KIND_HOLE, KIND_DARK, KIND_GAS = 0, 1, 2, 3
DEFAULT_THETA = 0.9


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_create.restype = ctypes.c_void_p
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_build.restype = ctypes.c_int64
        L.or_build.argtypes = [ctypes.c_void_p, _D, _D, _D, _D, _U8, _U8, ctypes.c_int64, ctypes.c_double,
                               ctypes.c_int]
        L.or_accelerations.argtypes = [ctypes.c_void_p, _D, _D, _D, _D, _U8, _U8, ctypes.c_int64, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _D]
        L.or_accelerations_built.argtypes = [ctypes.c_void_p, ctypes.c_double, _D, _D, _D]
        L.or_accelerations_sorted.argtypes = [ctypes.c_void_p, ctypes.c_double, _D, _D, _D]
        L.or_potential_energy.restype = ctypes.c_double
        L.or_potential_energy.argtypes = [ctypes.c_void_p, _D, _D, _D, _D, _U8, _U8, ctypes.c_int64,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double]
        L.or_walk.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                              ctypes.c_double, _D]
        L.or_walk_potential.restype = ctypes.c_double
        L.or_walk_potential.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int32, ctypes.c_double]
        L.or_direct.argtypes = [_D, _D, _D, _D, _U8, ctypes.c_int64, _I64, ctypes.c_int64, ctypes.c_double,
                                ctypes.c_int, ctypes.c_int, _D, _D, _D]
        L.or_potential_exact.restype = ctypes.c_double
        L.or_potential_exact.argtypes = [_D, _D, _D, _D, _U8, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_double]
        L.or_radix_sort.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
        L.or_morton_key.restype = ctypes.c_uint64
        L.or_morton_key.argtypes = [ctypes.c_uint64] * 3
        for f in ("or_n_stars", "or_n_nodes"):
            getattr(L, f).restype = ctypes.c_int64
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.or_nodes.restype = ctypes.c_void_p
        L.or_nodes.argtypes = [ctypes.c_void_p]
        L.or_keys.restype = ctypes.c_void_p
        L.or_keys.argtypes = [ctypes.c_void_p]
        L.or_order.restype = ctypes.c_void_p
        L.or_order.argtypes = [ctypes.c_void_p]
        L.or_sorted.restype = ctypes.c_void_p
        L.or_sorted.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_monopole.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_workers.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_node_size.restype = ctypes.c_int
        L.or_kick.argtypes = [ctypes.c_int64, _D, _D, _D, _D, _D, _D, ctypes.c_double]
        L.or_drift.argtypes = [ctypes.c_int64, _D, _D, _D, _D, _D, _D, ctypes.c_double]
        L.or_rung_for.restype = ctypes.c_int
        L.or_rung_for.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.or_block_active.argtypes = [ctypes.c_int64, _I8, ctypes.c_int, ctypes.c_int, _U8]
        L.or_block_kick.argtypes = [ctypes.c_int64, _D, _D, _D, _D, _D, _D, _U8, _U8, _I8, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.or_diagnostics.argtypes = [ctypes.c_int64, _D, _D, _D, _D, _D, _D, _D, _U8, ctypes.c_int, _D]
        L.or_split_step.restype = ctypes.c_int
        _U64 = ctypes.POINTER(ctypes.c_uint64)
        L.or_split_step.argtypes = [ctypes.c_void_p, _D, _D, _D, _D, _D, _D, _D, _U8, _U8, _D, _D, _D,
                                    ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, _D, _D, _D, ctypes.c_void_p, _U64,
                                    ctypes.c_double, ctypes.c_int]
        L.or_retard_new.restype = ctypes.c_void_p
        L.or_retard_free.argtypes = [ctypes.c_void_p]
        L.or_retard_count.restype = ctypes.c_int
        L.or_retard_count.argtypes = [ctypes.c_void_p]
        L.or_hole_appeared.argtypes = [ctypes.c_void_p, _D, _D, _D, _D, _D, _D, _U64, ctypes.c_int64,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.or_source_pos.restype = ctypes.c_int
        L.or_source_pos.argtypes = [ctypes.c_void_p, _D, _D, _D, _U8, _U8, _U64, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, _D]
        L.or_drag_kick.argtypes = [_D, _D, _D, _D, _D, _D, _D, _U8, _U8, _D, _D, _D, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.or_drag_pass.argtypes = [_D, _D, _D, _D, _D, _D, _D, _U8, _U8, _D, _D, _D, ctypes.c_int64, _U8,
                                   ctypes.POINTER(ctypes.c_int32), ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double]
        assert L.or_node_size() == NODE_DTYPE.itemsize
        L.or_pair_force.restype = ctypes.c_double
        L.or_pair_force.argtypes = [ctypes.c_double] * 4
        L.or_pair_potential.restype = ctypes.c_double
        L.or_pair_potential.argtypes = [ctypes.c_double] * 5
        L.or_pair_rs.restype = ctypes.c_double
        L.or_pair_rs.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        # SPH neighbour tree and hydro (sph_oracle.c)
        for f in ("or_cbrt",):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_double]
        for f in ("or_sph_w", "or_sph_dw"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_double, ctypes.c_double]
        L.or_sph_tree_new.restype = ctypes.c_void_p
        L.or_sph_tree_free.argtypes = [ctypes.c_void_p]
        L.or_sph_tree_build.restype = ctypes.c_int32
        L.or_sph_tree_build.argtypes = [ctypes.c_void_p, _I32, ctypes.c_int32, _D, _D, _D, _D]
        L.or_sph_tree_refit_h.argtypes = [ctypes.c_void_p, _D]
        L.or_sph_tree_sort_leaves.argtypes = [ctypes.c_void_p]
        for f in ("or_sph_tree_nnodes", "or_sph_tree_nleaves"):
            getattr(L, f).restype = ctypes.c_int32
            getattr(L, f).argtypes = [ctypes.c_void_p]
        for f in ("or_sph_tree_nodes", "or_sph_tree_keys", "or_sph_tree_order", "or_sph_tree_leaves"):
            getattr(L, f).restype = ctypes.c_void_p
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.or_sph_each.restype = ctypes.c_int32
        L.or_sph_each.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, _I32, _D, ctypes.c_int32]
        L.or_sph_group_scan.restype = ctypes.c_int32
        L.or_sph_group_scan.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, _I32, _D, ctypes.c_int32]
        L.or_sph_forces.restype = ctypes.c_int64
        L.or_sph_forces.argtypes = [ctypes.c_int64, _D, _D, _D, _D, _D, _D, _D, _U8, _U8, _D, _D, _D,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _U8, _D,
                                    _D, _D, _D, _D, _D, _D, ctypes.POINTER(ctypes.c_uint64)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _u8(a):
    return a.ctypes.data_as(_U8)


def _state_arrays(x, y, z, m, kind=None, alive=None):
    x, y, z, m = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, m))
    n = len(x)
    kind = np.zeros(n, np.uint8) if kind is None else np.ascontiguousarray(kind, dtype=np.uint8)
    alive = np.ones(n, np.uint8) if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
    return x, y, z, m, kind, alive


class Solver:
    """CPU ``bh.Solver`` restatement (internal/sim/bh/bh.go:29-38).

    ``theta == 0`` means DefaultTheta = 0.9 (bh.go:69-74).  ``workers`` is the
    thread count of the walk (ParallelFor, internal/sim/parallel.go:15-40).
    """

    def __init__(self, theta: float = 0.0, monopole_only: bool = False, workers: int = 1):
        self.theta = theta
        self.monopole_only = monopole_only
        self.workers = workers
        self._L = lib()
        self._h = ctypes.c_void_p(self._L.or_create())
        self._keep = None

    def __del__(self):
        try:
            if self._h:
                self._L.or_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def theta_eff(self) -> float:
        return DEFAULT_THETA if self.theta == 0 else self.theta

    # -- build (build.go:11-51) ------------------------------------------------
    def build(self, x, y, z, m, kind=None, alive=None, dm_off=False, theta=None):
        arrs = _state_arrays(x, y, z, m, kind, alive)
        self._keep = arrs
        x, y, z, m, kind, alive = arrs
        th = self.theta_eff() if theta is None else theta
        self._L.or_set_monopole(self._h, int(self.monopole_only))
        return int(self._L.or_build(self._h, _dp(x), _dp(y), _dp(z), _dp(m), _u8(kind), _u8(alive), len(x),
                                    th, int(dm_off)))

    @property
    def n_stars(self) -> int:
        return int(self._L.or_n_stars(self._h))

    @property
    def n_nodes(self) -> int:
        return int(self._L.or_n_nodes(self._h))

    def nodes(self) -> np.ndarray:
        n = self.n_nodes
        if n == 0:
            return np.zeros(0, NODE_DTYPE)
        buf = (ctypes.c_char * (n * NODE_DTYPE.itemsize)).from_address(self._L.or_nodes(self._h))
        return np.frombuffer(buf, dtype=NODE_DTYPE).copy()

    def keys(self) -> np.ndarray:
        n = self.n_stars
        if n == 0:
            return np.zeros(0, np.uint64)
        buf = (ctypes.c_uint64 * n).from_address(self._L.or_keys(self._h))
        return np.frombuffer(buf, dtype=np.uint64).copy()

    def order(self) -> np.ndarray:
        n = self.n_stars
        if n == 0:
            return np.zeros(0, np.int32)
        buf = (ctypes.c_int32 * n).from_address(self._L.or_order(self._h))
        return np.frombuffer(buf, dtype=np.int32).copy()

    def sorted(self):
        n = self.n_stars
        out = []
        for w in range(4):
            if n == 0:
                out.append(np.zeros(0))
                continue
            buf = (ctypes.c_double * n).from_address(self._L.or_sorted(self._h, w))
            out.append(np.frombuffer(buf, dtype=np.float64).copy())
        return out

    # -- Accelerations (bh.go:76-102) ----------------------------------------
    def accelerations(self, x, y, z, m, kind=None, alive=None, eps=0.0, dm_off=False, out=None):
        x, y, z, m, kind, alive = _state_arrays(x, y, z, m, kind, alive)
        n = len(x)
        ax, ay, az = (np.zeros(n) for _ in range(3)) if out is None else out
        self._keep = (x, y, z, m, kind, alive)
        self._L.or_accelerations(self._h, _dp(x), _dp(y), _dp(z), _dp(m), _u8(kind), _u8(alive), n, self.theta,
                                 eps, int(dm_off), int(self.monopole_only), int(self.workers), _dp(ax), _dp(ay),
                                 _dp(az))
        return np.stack([ax, ay, az], axis=1)

    def accelerations_sorted(self, eps):
        """Walk every star of the last build; results in sorted order."""
        n = self.n_stars
        ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
        self._L.or_accelerations_sorted(self._h, eps, _dp(ax), _dp(ay), _dp(az))
        return np.stack([ax, ay, az], axis=1)

    def walk(self, x, y, z, self_index=-1, eps2=0.0):
        out = np.zeros(3)
        self._L.or_walk(self._h, x, y, z, self_index, eps2, _dp(out))
        return out

    def walk_potential(self, x, y, z, self_index=-1, eps2=0.0):
        return float(self._L.or_walk_potential(self._h, x, y, z, self_index, eps2))

    # -- PotentialEnergy (potential.go:12-42) ---------------------------------
    def potential_energy(self, x, y, z, m, kind=None, alive=None, eps=0.0, dm_off=False, c=0.0):
        x, y, z, m, kind, alive = _state_arrays(x, y, z, m, kind, alive)
        self._keep = (x, y, z, m, kind, alive)
        return float(self._L.or_potential_energy(self._h, _dp(x), _dp(y), _dp(z), _dp(m), _u8(kind), _u8(alive),
                                                 len(x), self.theta, eps, int(dm_off), int(self.monopole_only), c))


def direct(x, y, z, m, kind=None, eps=0.0, dm_off=False, targets=None, workers=1):
    """FP64 O(N^2) direct sum, direct.go:25-48.  Rows follow ``targets``."""
    x, y, z, m, kind, _ = _state_arrays(x, y, z, m, kind, None)
    n = len(x)
    if targets is None:
        nt, tp = n, None
    else:
        targets = np.ascontiguousarray(targets, dtype=np.int64)
        nt, tp = len(targets), targets.ctypes.data_as(_I64)
    ax, ay, az = np.zeros(nt), np.zeros(nt), np.zeros(nt)
    lib().or_direct(_dp(x), _dp(y), _dp(z), _dp(m), _u8(kind), n, tp, nt, eps, int(dm_off), int(workers),
                    _dp(ax), _dp(ay), _dp(az))
    return np.stack([ax, ay, az], axis=1)


def potential_exact(x, y, z, m, kind=None, eps=0.0, dm_off=False, c=0.0):
    """PotentialEnergyParallel (diag.go:44-64): the exact pair-model potential
    over all pairs (Paczynski-Wiita star-hole pairs for c > 0)."""
    x, y, z, m, kind, _ = _state_arrays(x, y, z, m, kind, None)
    return float(lib().or_potential_exact(_dp(x), _dp(y), _dp(z), _dp(m), _u8(kind), len(x), eps, int(dm_off),
                                          float(c)))


def pair_force(m, r2, eps2, rs):
    """PairForce, pair.go:26-34: f with acceleration f*d from mass m at |d|^2 = r2."""
    return float(lib().or_pair_force(m, r2, eps2, rs))


def pair_potential(mi, mj, r2, eps2, rs):
    """PairPotential, pair.go:37-47 (Newtonian, or Paczynski-Wiita with a
    linear continuation inside 1.5 rs)."""
    return float(lib().or_pair_potential(mi, mj, r2, eps2, rs))


def pair_rs(ki, mi, kj, mj, c):
    """PairRS, pair.go:13-22: the hole's Schwarzschild radius for a star-hole
    pair, 0 otherwise."""
    return float(lib().or_pair_rs(ki, mi, kj, mj, c))


def morton_key(x: int, y: int, z: int) -> int:
    return int(lib().or_morton_key(x, y, z))


def radix_sort(keys: np.ndarray):
    """Serial stable LSD radix sort (morton.go:25-49); returns (keys, idx)."""
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    n = len(k)
    idx = np.arange(n, dtype=np.int32)
    tk, ti = np.zeros(n, np.uint64), np.zeros(n, np.int32)
    lib().or_radix_sort(k.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                        idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                        tk.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                        ti.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n)
    return k, idx


def kick_drift_kick(state, acc_fn, dt):
    """One KDK leapfrog step (sim.go:262-281 without holes/gas): kick(dt/2),
    drift(dt), new forces, kick(dt/2).  ``state`` is a dict of float64 arrays
    px..vz; ``acc_fn(px, py, pz)`` returns an (n, 3) acceleration array and is
    assumed to hold the forces of the current positions on entry (state['acc'])."""
    L = lib()
    n = len(state["px"])
    a = state["acc"]
    ax, ay, az = (np.ascontiguousarray(a[:, i]) for i in range(3))
    L.or_kick(n, _dp(state["vx"]), _dp(state["vy"]), _dp(state["vz"]), _dp(ax), _dp(ay), _dp(az), 0.5 * dt)
    L.or_drift(n, _dp(state["px"]), _dp(state["py"]), _dp(state["pz"]), _dp(state["vx"]), _dp(state["vy"]),
               _dp(state["vz"]), dt)
    a = acc_fn(state["px"], state["py"], state["pz"])
    ax, ay, az = (np.ascontiguousarray(a[:, i]) for i in range(3))
    L.or_kick(n, _dp(state["vx"]), _dp(state["vy"]), _dp(state["vz"]), _dp(ax), _dp(ay), _dp(az), 0.5 * dt)
    state["acc"] = a
    return state


def rung_for(dt: float, T: float, max_rung: int) -> int:
    """rungFor, internal/sim/blockstep.go:37-45."""
    return lib().or_rung_for(float(dt), float(T), int(max_rung))


def block_step(state, acc_fn, T, levels, eta, eps, kind, alive):
    """One global step T with power-of-two block steps, the event loop of
    internal/sim/blockstep.go:130-198 applied to gravity (or_block_kick holds
    the per-body rules).  ``state`` holds float64 arrays px..vz and 'acc' (the
    forces of the current positions); ``acc_fn(px, py, pz)`` returns the
    (n, 3) forces of all bodies, of which only the active stars' rows are
    taken (an active star's force does not depend on which others are active).
    Returns the rung populations at the start and the events (tick, active
    stars)."""
    L = lib()
    n = len(state["px"])
    max_rung = levels - 1
    full = 1 << max_rung
    unit = T / full
    crit = 2 * eta * eps
    kind = np.ascontiguousarray(kind, np.uint8)
    alive = np.ascontiguousarray(alive, np.uint8)
    star = (alive == 1) & (kind != 1)
    hole = (alive == 1) & (kind == 1)
    rung = np.zeros(n, np.int8)
    active = np.zeros(n, np.uint8)
    i8 = lambda a: a.ctypes.data_as(_I8)

    def kick(now):
        acc = state["acc"]
        ax, ay, az = (np.ascontiguousarray(acc[:, c]) for c in range(3))
        L.or_block_kick(n, _dp(state["vx"]), _dp(state["vy"]), _dp(state["vz"]), _dp(ax), _dp(ay), _dp(az),
                        _u8(kind), _u8(alive), i8(rung), now, full, max_rung, float(T), float(crit))

    kick(0)
    rungs = np.bincount(rung, minlength=levels)[:levels].tolist()
    events = []
    now = 0
    while now < full:
        nxt = full
        for r in np.unique(rung):
            ln = full >> int(r)
            nxt = min(nxt, (now // ln + 1) * ln)
        L.or_drift(n, _dp(state["px"]), _dp(state["py"]), _dp(state["pz"]), _dp(state["vx"]), _dp(state["vy"]),
                   _dp(state["vz"]), float(nxt - now) * unit)
        now = nxt
        L.or_block_active(n, i8(rung), now, full, _u8(active))
        sel = (active == 1) & star
        a = acc_fn(state["px"], state["py"], state["pz"])
        acc = state["acc"].copy()
        acc[sel] = a[sel]
        acc[hole] = 0.0
        state["acc"] = acc
        events.append((now, int(sel.sum())))
        kick(now)
    return {"rung_count": rungs, "events": events}


def diagnostics(px, py, pz, vx, vy, vz, m, kind, dm_off=False):
    """internal/sim/diag.go sums in state order: kinetic, momentum (3),
    angular momentum (3), mass, sum m x (3)."""
    out = np.zeros(11)
    f = lambda a: np.ascontiguousarray(a, np.float64)
    arrs = [f(a) for a in (px, py, pz, vx, vy, vz, m)]
    k = np.ascontiguousarray(kind, np.uint8)
    lib().or_diagnostics(len(arrs[0]), *[_dp(a) for a in arrs], _u8(k), int(dm_off), _dp(out))
    return out


GRAVITY = {"instant": 0, "retarded": 1, "retarded+v": 2}  # sim.GravityMode, retarded.go:21-28


def _u64(a):
    return None if a is None else np.ascontiguousarray(a, np.uint64).ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class Retard:
    """The holes' histories for retarded gravity (holeHistory by body ID,
    internal/sim/retarded.go:62-209), kept across steps."""

    def __init__(self):
        self._h = ctypes.c_void_p(lib().or_retard_new())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_retard_free(self._h)
            self._h = None

    def __len__(self):
        return lib().or_retard_count(self._h)

    def hole_appeared(self, state, i, time, dt, c, ids=None):
        """holeAppeared (retarded.go:157-161): hole i was added at ``time``."""
        keep = None if ids is None else np.ascontiguousarray(ids, np.uint64)
        lib().or_hole_appeared(self._h, *(_dp(np.ascontiguousarray(state[k], np.float64))
                                          for k in ("px", "py", "pz", "vx", "vy", "vz")),
                               _u64(keep), int(i), float(time), float(dt), float(c))

    def source_pos(self, state, kind, alive, k, obs, time, dt, c, gravity, ids=None):
        """sourcePos (retarded.go:118-155) of alive hole number k for an
        observer at obs; None if its gravity has not arrived there."""
        out = np.zeros(3)
        keep = None if ids is None else np.ascontiguousarray(ids, np.uint64)
        pos = [np.ascontiguousarray(state[a], np.float64) for a in ("px", "py", "pz")]
        vis = lib().or_source_pos(self._h, *(_dp(p) for p in pos), _u8(np.ascontiguousarray(kind, np.uint8)),
                                  _u8(np.ascontiguousarray(alive, np.uint8)), _u64(keep), len(pos[0]), int(k),
                                  float(obs[0]), float(obs[1]), float(obs[2]), float(time), float(dt), float(c),
                                  GRAVITY[gravity], _dp(out))
        return out if vis else None


def split_step(solver, state, m, kind, alive, eps, dt, c=0.0, dm_off=False, acc_valid=True, retard=None, time=0.0,
               gravity="instant", ids=None):
    """One step with black holes (internal/sim/split.go:164-199: RESPA
    substeps for stars near holes, no gas) on the float64 arrays px..vz and
    'acc' of ``state``, in place; the slow forces come from ``solver`` (an
    oracle Solver).  If ``state`` holds the hole spins "jx", "jy", "jz"
    (float64, per body), frame dragging (framedrag.go) acts with c > 0 and the
    spins are updated in place.  With ``retard`` (a Retard) the holes' pull
    follows ``gravity`` (retarded.go) at simulation time ``time``, and the
    holes are recorded at time + dt after the step (sim.go:285-287); ``ids``
    are the bodies' IDs (default: the state index).  Returns the number of
    substeps."""
    L = lib()
    n = len(state["px"])
    spin = "jx" in state
    for k in ("px", "py", "pz", "vx", "vy", "vz") + (("jx", "jy", "jz") if spin else ()):
        state[k] = np.ascontiguousarray(state[k], np.float64)
    if "acc" not in state:
        state["acc"] = np.zeros((n, 3))
    ax, ay, az = (np.ascontiguousarray(state["acc"][:, c_]) for c_ in range(3))
    m = np.ascontiguousarray(m, np.float64)
    kind = np.ascontiguousarray(kind, np.uint8)
    alive = np.ascontiguousarray(alive, np.uint8)
    L.or_set_monopole(solver._h, int(solver.monopole_only))
    L.or_set_workers(solver._h, int(solver.workers))
    j = tuple(_dp(state[k]) if spin else None for k in ("jx", "jy", "jz"))
    keep = None if ids is None else np.ascontiguousarray(ids, np.uint64)
    sub = L.or_split_step(solver._h, _dp(state["px"]), _dp(state["py"]), _dp(state["pz"]), _dp(state["vx"]),
                          _dp(state["vy"]), _dp(state["vz"]), _dp(m), _u8(kind), _u8(alive), *j, n,
                          float(solver.theta), float(eps), int(dm_off), float(dt), float(c), int(bool(acc_valid)),
                          _dp(ax), _dp(ay), _dp(az), None if retard is None else retard._h, _u64(keep),
                          float(time), GRAVITY[gravity])
    state["acc"] = np.stack([ax, ay, az], axis=1)
    return sub


def drag_kick(state, m, kind, alive, i, dt, eps2, c):
    """One frame-dragging kick of body i (dragKick, framedrag.go:56-72) on
    the arrays px..vz, jx..jz of ``state``, in place."""
    L = lib()
    n = len(state["px"])
    for k in ("px", "py", "pz", "vx", "vy", "vz", "jx", "jy", "jz"):
        state[k] = np.ascontiguousarray(state[k], np.float64)
    m = np.ascontiguousarray(m, np.float64)
    kind = np.ascontiguousarray(kind, np.uint8)
    alive = np.ascontiguousarray(alive, np.uint8)
    L.or_drag_kick(*(_dp(state[k]) for k in ("px", "py", "pz", "vx", "vy", "vz")), _dp(m), _u8(kind), _u8(alive),
                   *(_dp(state[k]) for k in ("jx", "jy", "jz")), n, int(i), float(dt), float(eps2), float(c))


def _drag(state, m, kind, alive, near, near_list, fast, dt, eps2, c):
    L = lib()
    n = len(state["px"])
    for k in ("px", "py", "pz", "vx", "vy", "vz", "jx", "jy", "jz"):
        state[k] = np.ascontiguousarray(state[k], np.float64)
    m = np.ascontiguousarray(m, np.float64)
    kind = np.ascontiguousarray(kind, np.uint8)
    alive = np.ascontiguousarray(alive, np.uint8)
    near = None if near is None else np.ascontiguousarray(near, np.uint8)
    lst = np.ascontiguousarray(np.zeros(0) if near_list is None else near_list, np.int32)
    L.or_drag_pass(*(_dp(state[k]) for k in ("px", "py", "pz", "vx", "vy", "vz")), _dp(m), _u8(kind), _u8(alive),
                   *(_dp(state[k]) for k in ("jx", "jy", "jz")), n, None if near is None else _u8(near),
                   lst.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(lst), int(fast), float(dt), float(eps2),
                   float(c))


def drag_outer(state, m, kind, alive, dt, eps2, c, near=None):
    """dragOuter (framedrag.go:115-127): every alive body of ``state`` that
    is neither a hole nor flagged in ``near``, in state order, in place."""
    _drag(state, m, kind, alive, near, None, False, dt, eps2, c)


def drag_fast(state, m, kind, alive, near_list, dt, eps2, c):
    """dragFast (framedrag.go:129-143): the bodies of ``near_list`` in order,
    then the alive holes, in place."""
    _drag(state, m, kind, alive, None, near_list, True, dt, eps2, c)


# -- SPH neighbour tree and hydro (sph_oracle.c; internal/sph/*.go) ----------------------------

SPH_NODE_DTYPE = np.dtype(
    [("minx", "f8"), ("miny", "f8"), ("minz", "f8"), ("maxx", "f8"), ("maxy", "f8"), ("maxz", "f8"),
     ("hmax", "f8"), ("lo", "i4"), ("hi", "i4"), ("left", "i4"), ("right", "i4")])


def cbrt(x: float) -> float:
    """The cube root the hydro uses for smoothing-length targets (or_cbrt)."""
    return lib().or_cbrt(float(x))


def sph_w(r: float, h: float) -> float:
    """Cubic spline kernel W(r, h), internal/sph/kernel.go:9-22."""
    return lib().or_sph_w(float(r), float(h))


def sph_dw(r: float, h: float) -> float:
    """dW/dr, internal/sph/kernel.go:24-36."""
    return lib().or_sph_dw(float(r), float(h))


def _i32(a):
    return a.ctypes.data_as(_I32)


class SphTree:
    """The neighbour BVH of internal/sph/tree.go over a list of gas bodies."""

    def __init__(self):
        self._L = lib()
        self._h = ctypes.c_void_p(self._L.or_sph_tree_new())
        self._keep = None

    def __del__(self):
        try:
            if self._h:
                self._L.or_sph_tree_free(self._h)
                self._h = None
        except Exception:
            pass

    def build(self, gas, x, y, z, h):
        gas = np.ascontiguousarray(gas, np.int32)
        x, y, z, h = (np.ascontiguousarray(a, np.float64) for a in (x, y, z, h))
        self._keep = (gas, x, y, z, h)
        return int(self._L.or_sph_tree_build(self._h, _i32(gas), len(gas), _dp(x), _dp(y), _dp(z), _dp(h)))

    def refit_h(self, h):
        h = np.ascontiguousarray(h, np.float64)
        self._L.or_sph_tree_refit_h(self._h, _dp(h))

    def sort_leaves(self):
        self._L.or_sph_tree_sort_leaves(self._h)

    def _array(self, fn, count, dtype):
        if count == 0:
            return np.zeros(0, dtype)
        buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(fn(self._h))
        return np.frombuffer(buf, dtype=dtype).copy()

    def nodes(self):
        return self._array(self._L.or_sph_tree_nodes, self._L.or_sph_tree_nnodes(self._h), SPH_NODE_DTYPE)

    def leaves(self):
        return self._array(self._L.or_sph_tree_leaves, self._L.or_sph_tree_nleaves(self._h), np.int32)

    def keys(self):
        return self._array(self._L.or_sph_tree_keys, len(self._keep[0]), np.uint64)

    def order(self):
        return self._array(self._L.or_sph_tree_order, len(self._keep[0]), np.int32)

    def each(self, px, py, pz, radius, symmetric):
        """Bodies each() reports, in its visiting order, with their r^2."""
        cap = len(self._keep[0]) + 1
        out, r2 = np.zeros(cap, np.int32), np.zeros(cap)
        n = self._L.or_sph_each(self._h, px, py, pz, radius, int(symmetric), _i32(out), _dp(r2), cap)
        return out[:n], r2[:n]

    def group_scan(self, leaf, symmetric, px, py, pz, radius):
        """candidates(leaf) + scan for one body, in visiting order."""
        cap = len(self._keep[0]) + 1
        out, r2 = np.zeros(cap, np.int32), np.zeros(cap)
        n = self._L.or_sph_group_scan(self._h, int(leaf), int(symmetric), px, py, pz, radius, _i32(out), _dp(r2),
                                      cap)
        return out[:n], r2[:n]


def sph_forces(s, softening=0.0, gamma=0.0, ngb=0.0, alpha=0.0, active=None, step=None, wake=None, out=None):
    """One hydro evaluation, sph.Solver.Forces (internal/sph/sph.go:83-263),
    on the float64 arrays px..vz, m, u, h, rho and uint8 kind, alive of the
    dict ``s``; h and rho are updated in place.  Returns (out, evals): out is
    a dict ax, ay, az, dudt, divv, dt (float64) and wake (uint64 bits of
    float64; +Inf where not lowered), evals the number of density sums."""
    L = lib()
    n = len(s["px"])
    for k in ("px", "py", "pz", "vx", "vy", "vz", "m", "u", "h", "rho"):
        s[k] = np.ascontiguousarray(s[k], np.float64)
    kind = np.ascontiguousarray(s["kind"], np.uint8)
    alive = np.ascontiguousarray(s.get("alive", np.ones(n, np.uint8)), np.uint8)
    if out is None:
        out = {k: np.zeros(n) for k in ("ax", "ay", "az", "dudt", "divv", "dt")}
    if wake is None:
        wake = np.full(n, np.inf).view(np.uint64).copy()
    out["wake"] = wake
    act = None if active is None else np.ascontiguousarray(active, np.uint8)
    st = None if step is None else np.ascontiguousarray(step, np.float64)
    evals = L.or_sph_forces(n, *[_dp(s[k]) for k in ("px", "py", "pz", "vx", "vy", "vz", "m")], _u8(kind),
                            _u8(alive), _dp(s["u"]), _dp(s["h"]), _dp(s["rho"]), float(gamma), float(ngb),
                            float(alpha), float(softening), None if act is None else _u8(act),
                            None if st is None else _dp(st), *[_dp(out[k]) for k in
                                                               ("ax", "ay", "az", "dudt", "divv", "dt")],
                            wake.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out, int(evals)
