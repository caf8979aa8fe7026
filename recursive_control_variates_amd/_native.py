"""ctypes binding of librcv.so (the C ABI declared in include/rcv.h).

The library is the only compute path: if it is missing or fails to load,
``lib()`` raises, and nothing in the package falls back to another
implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librcv.so")
_lock = threading.Lock()
_lib = None

RCV_OK = 0
RCV_E_INVALID = -1
RCV_E_STATE = -2
RCV_E_HIP = -3
RCV_E_SCENE = -4

STATE_OLD = 0
STATE_NEW = 1
PARAM_TEXTURE = 0
PARAM_GRID = 1
PARAM_VALUE = 2
BSDF_DIFFUSE = 0
BSDF_NULL = 1
EMITTER_AREA = 0
EMITTER_CONSTANT = 1
INTEGRATOR_PATH = 0
INTEGRATOR_VOLPATH = 1
INTEGRATOR_VOLPATH_SIMPLE = 2
SIMPLE_NO_NEE, SIMPLE_NO_DRT, SIMPLE_NO_DRT_SUBSAMPLING, SIMPLE_NO_DRT_MIS = 1, 2, 4, 8
LOSS_L1 = 0
LOSS_L2 = 1

COUNTER_NAMES = ("paths", "collisions", "nee_steps", "traversals", "real_scatter",
                 "surface_hits", "grad_atomics", "null_events")


class RcvError(RuntimeError):
    """A librcv call returned a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"librcv error {code}: {message}")
        self.code = code


class SceneCheckError(RcvError):
    pass


class IntegratorDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("max_depth", C.c_int32), ("rr_depth", C.c_int32),
                ("hide_emitters", C.c_int32), ("samples_per_thread", C.c_int32),
                ("block_size", C.c_int32), ("flags", C.c_int32)]


class MediumDesc(C.Structure):
    _fields_ = [("sigma_t_param", C.c_int32), ("albedo_param", C.c_int32),
                ("albedo_rgb", C.c_float * 3), ("scale", C.c_float),
                ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
                ("sample_emitters", C.c_int32), ("majorant_factor", C.c_int32), ("homogeneous", C.c_int32)]


class CVState(C.Structure):
    _fields_ = [("img", C.c_void_p), ("v_n", C.c_void_p), ("w_s", C.c_void_p),
                ("mean_x", C.c_void_p), ("mean_y", C.c_void_p), ("var_x", C.c_void_p),
                ("var_y", C.c_void_p), ("cov", C.c_void_p), ("counters", C.c_void_p)]


_P = C.c_void_p
_I = C.c_int32
_I64 = C.c_int64
_F = C.c_float
_U32 = C.c_uint32

# name -> argtypes; every function returns int status except the first two
SIGNATURES = {
    "rcv_scene_create": [C.POINTER(_P)],
    "rcv_scene_destroy": [_P],
    "rcv_scene_add_param": [_P, _I, _I, _I, _I, _I, C.POINTER(_I)],
    "rcv_scene_bind_param": [_P, _I, _I, _P, _P],
    "rcv_scene_add_bsdf": [_P, _I, C.POINTER(_F), _I, C.POINTER(_I)],
    "rcv_scene_add_medium": [_P, C.POINTER(MediumDesc), C.POINTER(_I)],
    "rcv_scene_add_emitter": [_P, _I, C.POINTER(_F), C.POINTER(_I)],
    "rcv_scene_add_mesh": [_P, C.POINTER(_F), C.POINTER(_F), _I, C.POINTER(_I), _I, _I, _I, _I, _I,
                           C.POINTER(_I)],
    "rcv_scene_add_sphere": [_P, C.POINTER(_F), _F, _I, _I, _I, C.POINTER(_I)],
    "rcv_scene_add_box": [_P, C.POINTER(_F), _I, _I, _I, C.POINTER(_I)],
    "rcv_scene_add_sensor": [_P, _I, _I, _I, C.POINTER(_F), C.POINTER(_F), C.POINTER(_I)],
    "rcv_scene_commit": [_P, _P],
    "rcv_scene_check_twostate": [_P, _I, C.POINTER(_I), _I, C.POINTER(_I), _I],
    "rcv_scene_set_counting": [_P, _I],
    "rcv_scene_read_counters": [_P, C.POINTER(_I64), _I, _P],
    "rcv_scene_majorant_grid": [_P, _I, _I, _P, _I64, C.POINTER(_I), _P],
    "rcv_scene_resize_param": [_P, _I, _I, _I, _I, _P],
    "rcv_scene_set_majorant_factor": [_P, _I, _I, _P],
    "rcv_render": [_P, C.POINTER(IntegratorDesc), _I, _U32, _I, _P, _P],
    "rcv_render_twostates": [_P, C.POINTER(IntegratorDesc), _I, _U32, _I, _P, _P, _P],
    "rcv_render_backward": [_P, C.POINTER(IntegratorDesc), _I, _U32, _I, _P, _I, C.POINTER(_I),
                            C.POINTER(_P), _P],
    "rcv_render_forward": [_P, C.POINTER(IntegratorDesc), _I, _U32, _I, _I, C.POINTER(_I), C.POINTER(_P), _P,
                           _P],
    "rcv_cv_step": [C.POINTER(CVState), _P, _P, _I64, _I, _F, _F, _I, _P],
    "rcv_welford_update": [_P, _P, _P, _P, _I64, _F, _F, _P],
    "rcv_loss": [_P, _P, _I64, _I, _P, _P, _P],
    "rcv_loss_noalloc": [_P, _P, _I64, _I, _P, _P, _P],
    "rcv_adam_step": [_P, _P, _P, _P, _P, _I64, _I, _F, _F, _F, _F, _F, _F, _P],
}


def lib_path() -> str:
    return _LIB_PATH


def lib():
    """Loads librcv.so once; raises if it is absent (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(_LIB_PATH):
            raise ImportError(
                f"librcv.so not found at {_LIB_PATH}; build it with "
                "`python -m recursive_control_variates_amd.build_native` (hipcc, gfx950)")
        handle = C.CDLL(_LIB_PATH)
        handle.rcv_version.restype = C.c_int
        handle.rcv_version.argtypes = []
        handle.rcv_last_error.restype = C.c_char_p
        handle.rcv_last_error.argtypes = []
        for name, args in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = handle
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc == RCV_OK:
        return
    msg = lib().rcv_last_error().decode(errors="replace")
    if rc == RCV_E_SCENE:
        raise SceneCheckError(rc, msg)
    raise RcvError(rc, f"{what}: {msg}" if what else msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def ptr(t) -> int:
    """Device (or host) address of a contiguous tensor."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("librcv needs contiguous tensors")
    return t.data_ptr()


def stream_handle(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def f32_array(values):
    vals = [float(v) for v in values]
    return (C.c_float * len(vals))(*vals)


def i32_array(values):
    vals = [int(v) for v in values]
    return (C.c_int32 * len(vals))(*vals)
