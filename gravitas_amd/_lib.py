"""ctypes binding of libgravitas_hip.so (the C ABI declared in include/gravitas_hip.h).

There is no CPU fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgravitas_hip.so")

GX_OK = 0
GX_ERR_NO_DEVICE = 1
GX_ERR_CODE_OBJECT = 2
GX_ERR_PIPELINE = 3
GX_ERR_ARG = 4
GX_ERR_HIP = 5
GX_ERR_CAPACITY = 6
GX_ERR_UNSUPPORTED = 7
GX_WALK_FP32 = 0
GX_WALK_FP64 = 1
GX_WALK_FP32_Q16 = 2

# Every symbol include/gravitas_hip.h declares.
EXPORTS = (
    "gx_init", "gx_name", "gx_error", "gx_solver_new", "gx_solver_free", "gx_solver_config",
    "gx_solver_set_stream", "gx_solver_reserve", "gx_solver_tune_walk", "gx_solver_autotune", "gx_accelerations", "gx_accelerations_dev", "gx_enqueue_accelerations_dev",
    "gx_sync", "gx_potential_energy_dev", "gx_direct_sample_dev", "gx_last_seconds", "gx_get_stats",
    "gx_phase_ms", "gx_export", "gx_step_dev", "gx_kick_dev", "gx_drift_dev", "gx_solver_tune",
    "gx_batch_accelerations_dev", "gx_batch_step_dev", "gx_graph_capture_step", "gx_graph_launch",
    "gx_accelerations_active_dev", "gx_block_step_dev", "gx_diagnostics_dev", "gx_split_step_dev",
    "gx_hydro_new", "gx_hydro_free", "gx_hydro_config", "gx_hydro_set_stream", "gx_hydro_forces_dev",
    "gx_hydro_forces", "gx_hydro_get_stats", "gx_hydro_phase_ms", "gx_hydro_export",
    "gx_split_step_spin_dev", "gx_drag_outer_dev", "gx_drag_fast_dev", "gx_drag_get_stats",
    "gx_retard_new", "gx_retard_free", "gx_retard_count", "gx_retard_hole_appeared_dev", "gx_retard_source_pos_dev",
    "gx_split_step_retarded_dev", "gx_accelerations_noalloc", "gx_potential_exact_dev",
)

GX_GRAVITY = {"instant": 0, "retarded": 1, "retarded+v": 2}  # sim.GravityMode, internal/sim/retarded.go:21-28

GX_NOALLOC_MAX_BODIES = 65536
GX_BLOCK_MAX_LEVELS = 8
GX_BLOCK_MAX_TICKS = 128
GX_DIAG_VALUES = 11


class GxStats(ctypes.Structure):
    _fields_ = [
        ("n_stars", ctypes.c_int64), ("n_holes", ctypes.c_int64), ("n_nodes", ctypes.c_int64),
        ("pp_interactions", ctypes.c_int64), ("pc_interactions", ctypes.c_int64),
        ("max_stack", ctypes.c_int32), ("flags", ctypes.c_int32), ("depth", ctypes.c_int32),
        ("walk", ctypes.c_int32), ("fallbacks", ctypes.c_int64), ("list_entries", ctypes.c_int64),
        ("rounds", ctypes.c_int64), ("mixed_nodes", ctypes.c_int64),
        ("max_list", ctypes.c_int32 * 4), ("n_targets", ctypes.c_int64),
    ]


class GxTuneEntry(ctypes.Structure):
    _fields_ = [("group_size", ctypes.c_int32), ("persistent", ctypes.c_int32), ("stack", ctypes.c_int32),
                ("parts", ctypes.c_int32), ("waves", ctypes.c_int32), ("ms", ctypes.c_float), ("pp_interactions", ctypes.c_int64),
                ("pc_interactions", ctypes.c_int64)]


class GxDragStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in ("passes", "iterations", "max_iterations", "sequential_passes",
                                              "segments", "slow_segments")]


class GxBlockStats(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_int32), ("events", ctypes.c_int32), ("bodies", ctypes.c_int64),
        ("targets", ctypes.c_int64), ("rung_count", ctypes.c_int64 * GX_BLOCK_MAX_LEVELS),
        ("tick", ctypes.c_int32 * GX_BLOCK_MAX_TICKS), ("active", ctypes.c_int64 * GX_BLOCK_MAX_TICKS),
    ]


class GxGasParams(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("neighbors", ctypes.c_double), ("alpha", ctypes.c_double)]


class GxHydroStats(ctypes.Structure):
    _fields_ = [
        ("n_gas", ctypes.c_int64), ("n_nodes", ctypes.c_int64), ("n_leaves", ctypes.c_int64),
        ("depth", ctypes.c_int32), ("max_iterations", ctypes.c_int32), ("density_sums", ctypes.c_int64),
        ("pairs", ctypes.c_int64), ("walk_steps", ctypes.c_int64), ("density_pairs", ctypes.c_int64),
        ("visc_pairs", ctypes.c_int64),
    ]


class GxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gravitas_hip error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gravitas_amd.build` "
                           "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.c_double
    I64 = ctypes.c_int64
    I = ctypes.c_int
    L.gx_init.argtypes = [ctypes.c_char_p, I]
    L.gx_init.restype = I
    L.gx_name.argtypes = [ctypes.c_char_p, I]
    L.gx_name.restype = None
    L.gx_error.restype = ctypes.c_char_p
    L.gx_solver_new.restype = P
    L.gx_solver_free.argtypes = [P]
    L.gx_solver_config.argtypes = [P, D, I, I, I]
    L.gx_solver_set_stream.argtypes = [P, P]
    step = [P, P, P, P, P, P, P, I64, D, I, P, P, P]
    for f in ("gx_accelerations", "gx_accelerations_dev", "gx_enqueue_accelerations_dev"):
        getattr(L, f).argtypes = step
        getattr(L, f).restype = I
    L.gx_sync.argtypes = [P]
    L.gx_potential_energy_dev.argtypes = [P, P, P, P, P, P, P, I64, D, I, D, P]
    L.gx_potential_exact_dev.argtypes = [P, P, P, P, P, P, I64, D, I, D, P]
    L.gx_direct_sample_dev.argtypes = [P, P, P, P, P, P, I64, P, I64, D, I, P]
    L.gx_last_seconds.argtypes = [P]
    L.gx_last_seconds.restype = D
    L.gx_get_stats.argtypes = [P, ctypes.POINTER(GxStats)]
    L.gx_phase_ms.argtypes = [P, ctypes.POINTER(ctypes.c_float), I]
    L.gx_step_dev.argtypes = [P, P, P, P, P, P, P, P, P, P, I64, D, I, D, I, P, P, P]
    L.gx_kick_dev.argtypes = [P, P, P, P, P, P, P, I64, D]
    L.gx_drift_dev.argtypes = [P, P, P, P, P, P, P, I64, D]
    L.gx_solver_tune.argtypes = [P, I, I]
    L.gx_solver_reserve.argtypes = [P, I64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.gx_solver_tune_walk.argtypes = [P, I, I, I, I, I]
    L.gx_accelerations_noalloc.argtypes = [P, P, P, P, P, P, I64, D, D, I, I, P, P, P, P]
    L.gx_solver_autotune.argtypes = [P, P, P, P, P, P, P, P, ctypes.c_int32, P, P, I, P, P, P, ctypes.c_int32,
                                     ctypes.POINTER(GxTuneEntry), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.gx_batch_accelerations_dev.argtypes = [P, P, P, P, P, P, P, P, ctypes.c_int32, P, P, I, P, P, P]
    L.gx_batch_step_dev.argtypes = [P, P, P, P, P, P, P, P, P, P, P, ctypes.c_int32, P, P, I, D, I, P, P, P]
    L.gx_graph_capture_step.argtypes = [P, P, P, P, P, P, P, P, P, P, P, ctypes.c_int32, P, P, I, D, P, P, P]
    L.gx_graph_launch.argtypes = [P, ctypes.c_int32]
    L.gx_accelerations_active_dev.argtypes = [P, P, P, P, P, P, P, P, I64, D, I, P, P, P]
    L.gx_block_step_dev.argtypes = [P, P, P, P, P, P, P, P, P, P, I64, D, I, D, ctypes.c_int32, D, I, P, P, P,
                                    ctypes.POINTER(GxBlockStats)]
    L.gx_split_step_dev.argtypes = [P, P, P, P, P, P, P, P, P, P, I64, D, I, D, D, I, P, P, P,
                                    ctypes.POINTER(ctypes.c_int32)]
    L.gx_split_step_spin_dev.argtypes = [P, P, P, P, P, P, P, P, P, P, P, P, P, I64, D, I, D, D, I, P, P, P,
                                         ctypes.POINTER(ctypes.c_int32)]
    L.gx_split_step_retarded_dev.argtypes = [P, P, P, P, P, P, P, P, P, P, P, P, P, P, I64, D, I, D, D, I, P, P, P,
                                             P, D, I, ctypes.POINTER(ctypes.c_int32)]
    L.gx_retard_new.restype = P
    L.gx_retard_new.argtypes = []
    L.gx_retard_free.argtypes = [P]
    L.gx_retard_free.restype = None
    L.gx_retard_count.argtypes = [P]
    L.gx_retard_hole_appeared_dev.argtypes = [P, P, P, P, P, P, P, P, P, I64, I64, D, D, D]
    L.gx_retard_source_pos_dev.argtypes = [P, P, P, P, P, P, P, P, I64, I64, P, P, P, I64, D, D, D, I, P, P, P, P]
    L.gx_drag_outer_dev.argtypes = [P, P, P, P, P, P, P, P, P, P, P, P, P, P, I64, D, D, D]
    L.gx_drag_fast_dev.argtypes = [P, P, P, P, P, P, P, P, P, P, P, I64, P, P, P, I64, D, D, D]
    L.gx_drag_get_stats.argtypes = [P, ctypes.POINTER(GxDragStats)]
    L.gx_diagnostics_dev.argtypes = [P, P, P, P, P, P, P, P, P, I64, I, ctypes.POINTER(ctypes.c_double)]
    L.gx_export.argtypes = [P, I, P, I64]
    L.gx_export.restype = I64
    L.gx_hydro_new.restype = P
    L.gx_hydro_free.argtypes = [P]
    L.gx_hydro_free.restype = None
    L.gx_hydro_config.argtypes = [P, I]
    L.gx_hydro_set_stream.argtypes = [P, P]
    hydro = [P] + [P] * 12 + [I64, ctypes.POINTER(GxGasParams), D, P, P] + [P] * 7
    L.gx_hydro_forces_dev.argtypes = hydro
    L.gx_hydro_forces.argtypes = hydro
    L.gx_hydro_get_stats.argtypes = [P, ctypes.POINTER(GxHydroStats)]
    L.gx_hydro_phase_ms.argtypes = [P, ctypes.POINTER(ctypes.c_float), I]
    L.gx_hydro_export.argtypes = [P, I, P, I64]
    L.gx_hydro_export.restype = I64
    _lib = L
    return L


def check(rc):
    if rc != GX_OK:
        raise GxError(rc, load().gx_error().decode(errors="replace"))
    return rc
