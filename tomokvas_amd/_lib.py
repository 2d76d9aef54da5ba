"""ctypes binding of libtomokvas_fdk.so (the C ABI in include/tomokvas/fdk.h).

The library is built in place by ``python -m tomokvas_amd.build``.  Loading
fails loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtomokvas_fdk.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "tomokvas", "fdk.h")

FDK_OK = 0
FDK_ERR_INVALID_ARGUMENT = 1
FDK_ERR_NOT_INITIALIZED = 2
FDK_ERR_HIP = 3
FDK_ERR_OUT_OF_MEMORY = 4
FDK_ERR_CALLBACK = 5
FDK_ERR_UNSUPPORTED = 6

F32, F16, BF16, F64 = 0, 1, 2, 3
FILTER_RAMLAK, FILTER_SHEPPLOGAN, FILTER_USER = 0, 1, 2

GET_BLOCK_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_int, C.c_int)

_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_SIGS = {
    "fdk_create": (_i, [_i, C.POINTER(_vp)]),
    "fdk_destroy": (_i, [_vp]),
    "fdk_get_free_mem_info": (C.c_size_t, [_i]),
    "fdk_last_error_string": (C.c_char_p, []),
    "fdk_set_gpu_id": (_i, [_vp, _i]),
    "fdk_set_block_size": (_i, [_vp, _i]),
    "fdk_set_allocation_limit": (_i, [_vp, C.c_longlong]),
    "fdk_set_collecting": (_i, [_vp, _i]),
    "fdk_set_show_info": (_i, [_vp, _i]),
    "fdk_set_axis_direction": (_i, [_vp, _i, _i, _i]),
    "fdk_set_sinogram_properties": (_i, [_vp, _vp, _i, _f, _i, _i, _f, _f, _f, _f]),
    "fdk_set_reconstruction_properties": (_i, [_vp, _i, _f, _i, _f, _i, _f, _f]),
    "fdk_set_scale_xyz": (_i, [_vp, _f]),
    "fdk_set_filter": (_i, [_vp, _i, _vp, _i]),
    "fdk_set_storage": (_i, [_vp, _i]),
    "fdk_initialize": (_i, [_vp]),
    "fdk_calculate_reconstruction": (_i, [_vp, GET_BLOCK_FN, _vp]),
    "fdk_device_volume": (_vp, [_vp]),
    "fdk_device_sinogram": (_vp, [_vp]),
    "fdk_get_stage_times": (_i, [_vp, _vp]),
    "fdk_get_geometry": (_i, [_vp, _vp, _vp]),
    "fdk_prepare": (_i, [_vp]),
    "fdk_weight": (_i, [_vp, _i, _vp, _i, _vp]),
    "fdk_filter": (_i, [_vp, _i, _vp, _i, _i, _i, _vp, _vp]),
    "fdk_stack_bytes": (C.c_size_t, [_vp, _i, _i]),
    "fdk_pack": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _vp, _vp]),
    "fdk_backproject": (_i, [_vp, _i, _vp, _vp, _i, _i, _vp, _i, _vp]),
    "fdk_forward_project": (_i, [_vp, _i, _vp, _i, _i, _vp, _vp]),
    "fdk_tap_gradient": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp]),
    "fdk_accumulate": (_i, [_vp, _vp, C.c_size_t, _vp]),
    "fdk_compare": (_i, [_vp, _vp, C.c_size_t, _vp, _vp]),
    "fdk_phantom_project": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "fdk_poisson_log": (_i, [_vp, C.c_size_t, C.c_double, C.c_uint64, _vp]),
    "fdk_probe_l2_read": (_i, [C.c_size_t, C.POINTER(C.c_double)]),
}

_lib = None


class FDKError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def header_functions() -> list:
    """Names of every function the C ABI header declares."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(fdk_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(n for n in names if n != "fdk_get_block_fn"))


def load():
    """Load the library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m tomokvas_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    s = load().fdk_last_error_string()
    return s.decode() if s else ""


def check(rc: int, what: str = "") -> None:
    if rc != FDK_OK:
        raise FDKError(rc, f"{what}: {last_error()}" if what else last_error())
