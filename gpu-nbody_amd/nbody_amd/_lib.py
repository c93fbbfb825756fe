"""ctypes binding of libnbody.so (the C ABI of include/nbody.h).

The library is the only compute path: if libnbody.so is missing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libnbody.so")

NB_OK = 0
NB_E_HIP = -1
NB_E_POOL = -2
NB_E_DEPTH = -3
NB_E_ARG = -4
NB_E_SYNC = -5
NB_E_STATE = -6
NB_E_STACK = -7
NB_E_MODEL = -8

NB_FP64_POSITIONS = 1
NB_FUSED_UPDATE = 2
NB_MODE_PLAIN = 0
NB_MODE_LEARNED = 1
NB_MODE_COARSE = 2

NB_MLP_IN = 24
NB_MLP_HIDDEN = 64
NB_MLP_OUT = 3
NB_MLP_PARAMS = 5955
NB_MODEL_FLOATS = NB_MLP_PARAMS + 2 * NB_MLP_IN + 2 * NB_MLP_OUT

NB_STAGE_BBOX, NB_STAGE_BUILD, NB_STAGE_SUMMARIZE, NB_STAGE_SORT, NB_STAGE_FORCE, NB_STAGE_INTEGRATE = range(1, 7)


class NbParams(C.Structure):
    _fields_ = [
        ("n_bodies", C.c_int32),
        ("theta", C.c_float),
        ("softening", C.c_float),
        ("dt", C.c_float),
        ("vote_width", C.c_int32),
        ("flags", C.c_uint32),
        ("device", C.c_int32),
        ("pool_factor", C.c_float),
        ("mode", C.c_int32),
        ("coarse_theta", C.c_float),
        ("check_tol", C.c_float),
        ("check_every", C.c_int32),
        ("check_samples", C.c_int32),
        ("coarse_alpha", C.c_float),
        ("check_adapt", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class NbCounters(C.Structure):
    _fields_ = [
        ("step", C.c_int64),
        ("cells", C.c_int64),
        ("max_depth", C.c_int64),
        ("interactions", C.c_int64),
        ("error", C.c_int64),
        ("fallbacks", C.c_int64),
        ("stage_ms", C.c_double * 8),
        ("last_check_median", C.c_double),
        ("live_cells", C.c_int64),
        ("cell_interactions", C.c_int64),
        ("check_ms", C.c_double),
        ("checks", C.c_int64),
        ("coarse_alpha", C.c_double),
        ("large_buckets", C.c_int64),
    ]


class NbTrainParams(C.Structure):
    _fields_ = [
        ("steps", C.c_int32),
        ("groups_per_step", C.c_int32),
        ("iters", C.c_int32),
        ("batch", C.c_int32),
        ("lr", C.c_float),
        ("lr_final", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("adam_eps", C.c_float),
        ("seed", C.c_uint64),
        ("reserved", C.c_int32 * 6),
    ]


class NbTrainReport(C.Structure):
    _fields_ = [
        ("samples", C.c_int32),
        ("loss_first", C.c_float),
        ("loss_last", C.c_float),
        ("gen_ms", C.c_double),
        ("train_ms", C.c_double),
        ("reserved", C.c_int64 * 4),
    ]


class NbControl(C.Structure):
    _fields_ = [
        ("steps_since_model", C.c_int64),
        ("check_seed", C.c_uint64),
        ("coarse_alpha", C.c_float),
        ("fell_back", C.c_int32),
        ("fail_streak", C.c_int32),
        ("reserved0", C.c_int32),
        ("fallbacks", C.c_int64),
        ("checks", C.c_int64),
        ("last_check_median", C.c_double),
        ("reserved", C.c_int64 * 4),
    ]


class NbError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")

_lib = None


def load() -> C.CDLL:
    """Load libnbody.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python gpu-nbody_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.nb_default_params.argtypes = [C.POINTER(NbParams), C.c_int32]
    L.nb_default_params.restype = None
    L.nb_create.argtypes = [C.POINTER(NbParams), C.POINTER(vp)]
    L.nb_destroy.argtypes = [vp]
    L.nb_destroy.restype = None
    L.nb_set_state.argtypes = [vp] + [_f32p] * 7
    L.nb_step.argtypes = [vp, C.c_int32]
    L.nb_run_stage.argtypes = [vp, C.c_int32]
    L.nb_copy_vertices.argtypes = [vp, vp, vp]
    L.nb_copy_vertices_noalloc.argtypes = [vp, vp, vp]
    L.nb_num_bodies.argtypes = [vp]
    L.nb_num_bodies.restype = C.c_int32
    L.nb_get_state.argtypes = [vp] + [vp] * 6
    L.nb_get_accel.argtypes = [vp] + [vp] * 3
    L.nb_get_positions64.argtypes = [vp] + [vp] * 3
    L.nb_get_root.argtypes = [vp, _f32p]
    L.nb_get_sorted.argtypes = [vp, _i32p]
    L.nb_get_paths.argtypes = [vp, _u64p, _i32p]
    L.nb_get_cells.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), vp, vp, vp, vp, vp]
    L.nb_get_counters.argtypes = [vp, C.POINTER(NbCounters)]
    L.nb_set_timing.argtypes = [vp, C.c_int32]
    L.nb_probe_accel.argtypes = [vp, vp, vp, vp]
    L.nb_direct_accel.argtypes = [vp, _i32p, C.c_int32, vp]
    L.nb_force_error.argtypes = [vp, C.c_int32, C.c_uint64, vp]
    L.nb_default_train_params.argtypes = [C.POINTER(NbTrainParams)]
    L.nb_default_train_params.restype = None
    L.nb_train.argtypes = [vp, C.POINTER(NbTrainParams), C.POINTER(NbTrainReport)]
    L.nb_set_model.argtypes = [vp, _f32p, C.c_int32]
    L.nb_get_model.argtypes = [vp, _f32p, C.c_int32]
    L.nb_coarse_features.argtypes = [vp, vp, vp, vp]
    L.nb_mlp_forward.argtypes = [vp, _f32p, C.c_int32, _f32p]
    L.nb_mlp_loss_grad.argtypes = [vp, _f32p, _f32p, C.c_int32, C.POINTER(C.c_float), _f32p]
    L.nb_mlp_adam.argtypes = [vp, _f32p, _f32p, C.c_int32, C.c_int32, C.c_float, _f32p]
    L.nb_diagnostics.argtypes = [vp, C.c_int32, C.c_uint64, vp]
    L.nb_get_masses.argtypes = [vp, vp]
    L.nb_restore_state.argtypes = [vp] + [vp] * 10 + [C.c_int32, C.c_int32]
    L.nb_get_control.argtypes = [vp, C.POINTER(NbControl)]
    L.nb_set_control.argtypes = [vp, C.POINTER(NbControl)]
    L.nb_last_error.argtypes = [vp]
    L.nb_last_error.restype = C.c_char_p
    L.nb_status_string.argtypes = [C.c_int]
    L.nb_status_string.restype = C.c_char_p
    for name in ("nb_create", "nb_set_state", "nb_step", "nb_run_stage", "nb_copy_vertices",
                 "nb_copy_vertices_noalloc",
                 "nb_get_state", "nb_get_accel", "nb_get_positions64", "nb_get_root",
                 "nb_get_sorted", "nb_get_paths", "nb_get_cells", "nb_get_counters",
                 "nb_set_timing", "nb_probe_accel", "nb_direct_accel", "nb_force_error",
                 "nb_train", "nb_set_model", "nb_get_model", "nb_coarse_features",
                 "nb_mlp_forward", "nb_mlp_loss_grad", "nb_mlp_adam", "nb_diagnostics",
                 "nb_get_masses", "nb_restore_state", "nb_get_control", "nb_set_control"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def ptr(a):
    """Raw pointer of a contiguous numpy array, or None."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)
