"""Loader of liblps.so (the in-tree HIP library).  There is no fallback: if the
library is missing or does not load, every solver call raises."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from .abi import CGridInfo, CParams, CStepStats, CBox, CSphere, CPlane, CTrainCfg, CTuning

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblps.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "lps.h")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class LibraryMissing(RuntimeError):
    pass


_lib = None


def header_functions() -> list:
    """Names of the functions declared in include/lps.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(lps_\w+)\s*\(", text, flags=re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run `python -m liquidsimulation_amd.build`")
    # PyTorch brings its own ROCm runtime libraries.  When both are used in one
    # process, liblps must bind to the HIP runtime torch loads: loaded first,
    # liblps would pull in the system's runtime, torch's libraries would then
    # run against it and find no device.  So torch, when installed, loads first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.lps_create.argtypes = [C.POINTER(CParams), C.POINTER(vp)]
    L.lps_destroy.argtypes = [vp]
    L.lps_device_info.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_int32)]
    L.lps_set_params.argtypes = [vp, C.POINTER(CParams)]
    L.lps_get_derived.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), vp]
    L.lps_get_delta.argtypes = [vp, C.c_float, C.POINTER(C.c_float)]
    L.lps_index_thresholds.argtypes = [C.POINTER(CParams), _f32p]
    L.lps_set_particles.argtypes = [vp, _f32p, vp, vp, vp, C.c_uint32]
    L.lps_add_particles.argtypes = [vp, _f32p, vp, vp, vp, C.c_uint32]
    L.lps_copy_positions_device.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint64]
    L.lps_get_particles.argtypes = [vp, vp, vp, vp, vp, vp, C.c_uint32]
    L.lps_get_count.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.lps_set_colliders.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32]
    L.lps_set_scenes.argtypes = [vp, C.c_uint32, _u32p, vp, _f32p, vp, vp, vp]
    L.lps_set_scene_colliders.argtypes = [vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32]
    L.lps_get_scene_count.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.lps_get_scene_stats.argtypes = [vp, vp, C.c_uint32]
    L.lps_step.argtypes = [vp, C.c_float, C.c_uint32, vp]
    L.lps_set_tuning.argtypes = [vp, C.POINTER(CTuning)]
    L.lps_get_tuning.argtypes = [vp, C.POINTER(CTuning)]
    L.lps_autotune.argtypes = [vp, C.c_float, C.c_uint32, C.POINTER(CTuning)]
    L.lps_get_grid.argtypes = [vp, C.POINTER(CGridInfo), vp, vp, vp, vp, C.c_uint32]
    L.lps_get_fields.argtypes = [vp, vp, vp, vp, vp]
    L.lps_mlp_set_weights.argtypes = [vp, _f32p, _f32p, C.c_float]
    L.lps_mlp_get_weights.argtypes = [vp, _f32p, _f32p, C.POINTER(C.c_float)]
    L.lps_mlp_capture.argtypes = [vp, C.c_int32]
    L.lps_dump_pairs.argtypes = [vp, vp, vp, vp, C.c_uint32]
    L.lps_mlp_collect.argtypes = [vp, C.c_uint64, C.c_uint32]
    L.lps_mlp_collected.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.lps_mlp_train.argtypes = [vp, C.POINTER(CTrainCfg), vp]
    L.lps_mlp_forward.argtypes = [vp, _f32p, C.c_uint32, _f32p]
    L.lps_mlp_grad.argtypes = [vp, _f32p, _f32p, C.c_uint32, C.c_float, _f32p, C.POINTER(C.c_float)]
    L.lps_get_energy.argtypes = [vp, vp, vp, C.c_uint32]
    L.lps_mlp_grad_noalloc.argtypes = [vp, _f32p, _f32p, C.c_uint32, C.c_float, _f32p, C.POINTER(C.c_float)]
    L.lps_profile_enable.argtypes = [vp, C.c_int32]
    L.lps_profile_query.argtypes = [vp, C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
    L.lps_profile_reset.argtypes = [vp]
    L.lps_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.lps_status_string.argtypes = [C.c_int]
    L.lps_status_string.restype = C.c_char_p
    _lib = L
    return L


class LPSError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        name = lib().lps_status_string(status).decode()
        super().__init__(f"{where}: {name} ({status})")


def check(status: int, where: str) -> None:
    if status != 0:
        raise LPSError(status, where)


def ptr(a):
    return None if a is None else a.ctypes.data
