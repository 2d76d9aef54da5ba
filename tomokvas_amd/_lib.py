"""ctypes binding of libtomokvas_fdk.so (the C ABI in include/tomokvas/fdk.h and
include/tomokvas/fdk_io.h).

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
HEADERS = [HEADER] + [os.path.join(os.path.dirname(HERE), "include", "tomokvas", h)
                      for h in ("fdk_io.h", "fdk_driver.h", "fdk_rsq.h", "fdk_tiff.h")]

FDK_OK = 0
FDK_ERR_INVALID_ARGUMENT = 1
FDK_ERR_NOT_INITIALIZED = 2
FDK_ERR_HIP = 3
FDK_ERR_OUT_OF_MEMORY = 4
FDK_ERR_CALLBACK = 5
FDK_ERR_UNSUPPORTED = 6

F32, F16, BF16, F64 = 0, 1, 2, 3
F16_PACKED = 4   # FP16 storage interpolated in packed FP16 (opt-in, FP16 weights)
FILTER_RAMLAK, FILTER_SHEPPLOGAN, FILTER_USER = 0, 1, 2
VOLUME_RESIDENT, VOLUME_STREAMED = 0, 1

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
    "fdk_reconstruct_host": (_i, [_vp, _vp, GET_BLOCK_FN, _vp]),
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
    "fdk_tap_gradient_noalloc": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp]),
    "fdk_accumulate": (_i, [_vp, _vp, C.c_size_t, _vp]),
    "fdk_compare": (_i, [_vp, _vp, C.c_size_t, _vp, _vp]),
    "fdk_phantom_project": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "fdk_poisson_log": (_i, [_vp, C.c_size_t, C.c_double, C.c_uint64, _vp]),
    "fdk_probe_l2_read": (_i, [C.c_size_t, C.POINTER(C.c_double)]),
    "fdk_set_volume_mode": (_i, [_vp, _i]),
    "fdk_get_volume_chunk": (_i, [_vp, C.POINTER(_i)]),
    "fdk_get_angle_chunks": (_i, [_vp, C.POINTER(_i)]),
    # file formats (fdk_io.h)
    "fdk_cra_open": (_i, [C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "fdk_cra_close": (None, [_vp]),
    "fdk_cra_get_meta": (_i, [_vp, _vp]),
    "fdk_cra_set_scan_block": (_i, [_vp, _i]),
    "fdk_cra_split": (_i, [_vp, C.c_longlong, C.POINTER(_i)]),
    "fdk_cra_set_angle_block": (_i, [_vp, _i]),
    "fdk_cra_block_projections": (_i, [_vp, C.POINTER(_i)]),
    "fdk_cra_rotation": (_i, [_vp, C.POINTER(C.c_double)]),
    "fdk_cra_load": (_i, [_vp, _vp]),
    "fdk_bin_open": (_i, [C.c_char_p, C.c_char_p, _i, _i, _i, _i, _i, C.POINTER(_vp)]),
    "fdk_bin_set_scan_block": (_i, [_vp, _i]),
    "fdk_bin_get_block": (C.POINTER(C.c_float), [_vp, _i, _i]),
    "fdk_bin_flush": (_i, [_vp]),
    "fdk_bin_close": (_i, [_vp]),
    # the reconstruction program (fdk_driver.h)
    "fdk_setting_defaults": (_i, [_vp]),
    "fdk_setting_load": (_i, [C.c_char_p, _vp]),
    "fdk_setting_save": (_i, [C.c_char_p, _vp]),
    "fdk_reconstruct_file": (_i, [C.c_char_p, _vp, _vp, _vp]),
    # Scanco RSQ (fdk_rsq.h)
    "fdk_rsq_default_params": (_i, [_vp]),
    "fdk_rsq_open": (_i, [C.c_char_p, C.POINTER(_vp)]),
    "fdk_rsq_close": (None, [_vp]),
    "fdk_rsq_get_header": (_i, [_vp, _vp]),
    "fdk_rsq_sizes": (_i, [_vp, _vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "fdk_rsq_load": (_i, [_vp, _vp, _i, _i, _vp]),
    "fdk_rsq_process": (_i, [_vp, C.c_double, C.c_double, _vp, _i, _i, _vp, _vp]),
    # TIFF export (fdk_tiff.h)
    "fdk_tiff_write_u16": (_i, [C.c_char_p, _vp, _i, _i]),
    "fdk_volume_minmax": (_i, [_vp, C.c_size_t, C.POINTER(C.c_double), _vp]),
    "fdk_volume_to_u16": (_i, [_vp, C.c_size_t, C.c_double, C.c_double, _vp, _vp]),
    "fdk_tiff_export_device": (_i, [_vp, _i, _i, _i, C.c_char_p, _i, C.POINTER(C.c_double)]),
    "fdk_tiff_export_host": (_i, [_vp, _i, _i, _i, C.c_double, C.c_double, C.c_char_p, _i]),
}


class CraMeta(C.Structure):
    """fdk_cra_meta (include/tomokvas/fdk_io.h)."""
    _fields_ = [("raw_file", C.c_char * 256), ("source_name", C.c_char * 256),
                ("sample_spacing", C.c_double), ("distance", C.c_double),
                ("detector_x", C.c_int), ("detector_z", C.c_int),
                ("num_of_projections", C.c_int), ("num_of_blocks", C.c_int),
                ("x_dimension", C.c_int), ("z_dimension", C.c_int),
                ("rotate_angle", C.c_double), ("clockwise", C.c_int), ("num_centres", C.c_int),
                ("centre_of_rotation", C.c_double * 64), ("scale", C.c_double),
                ("distance_px", C.c_float), ("delta_fi", C.c_float)]



class Setting(C.Structure):
    """fdk_setting (include/tomokvas/fdk_driver.h)."""
    _fields_ = [("gpu_id", C.c_int), ("mem_limit_mib", C.c_int), ("alloc_limit_mib", C.c_int),
                ("slices_per_block", C.c_int), ("x_direction", C.c_int),
                ("y_direction", C.c_int), ("z_direction", C.c_int),
                ("vertical_center", C.c_double), ("wait_for_enter", C.c_int)]


class RunOptions(C.Structure):
    """fdk_run_options (include/tomokvas/fdk_driver.h)."""
    _fields_ = [("storage", C.c_int), ("filter", C.c_int), ("angle_block_bytes", C.c_longlong),
                ("verbose", C.c_int), ("volume_mode", C.c_int)]


class RunStats(C.Structure):
    """fdk_run_stats (include/tomokvas/fdk_driver.h)."""
    _fields_ = [("scan_blocks", C.c_int), ("angle_blocks", C.c_int),
                ("seconds_total", C.c_double), ("seconds_load_wait", C.c_double),
                ("seconds_engine", C.c_double), ("seconds_flush", C.c_double),
                ("bytes_read", C.c_double), ("bytes_written", C.c_double)]


class RsqHeader(C.Structure):
    """fdk_rsq_header (include/tomokvas/fdk_rsq.h)."""
    _fields_ = [("check", C.c_char * 17)] + [(n, C.c_int) for n in (
        "data_type", "nr_of_bytes", "nr_of_blocks", "patient_index", "scanner_id", "dimx_p",
        "dimy_p", "dimz_p", "dimx_um", "dimy_um", "dimz_um", "mu_scaling", "nr_of_samples",
        "nr_of_projections", "scandist_um", "scanner_type", "reference_line_um", "energy",
        "intensity", "data_offset")] + [
        ("bh_x", C.c_int * 3), ("bh_y", C.c_double * 3), ("bh_a", C.c_double),
        ("bh_b", C.c_double), ("file_size", C.c_longlong)]


class RsqParams(C.Structure):
    """fdk_rsq_params (include/tomokvas/fdk_rsq.h)."""
    _fields_ = [("sin_x", C.c_int), ("sin_y", C.c_int), ("header_offset", C.c_longlong),
                ("axis_move", C.c_int), ("distance", C.c_double), ("delta_fi", C.c_double),
                ("flip", C.c_int), ("invert", C.c_int), ("complete", C.c_int),
                ("minimal", C.c_int)]


_lib = None


class FDKError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def header_functions() -> list:
    """Names of every function the C ABI headers declare."""
    txt = "".join(open(h).read() for h in HEADERS)
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
