"""Minimal HIP runtime calls over ctypes (device buffers for host-side helpers
that move data to and from the GPU without torch)."""
from __future__ import annotations

import ctypes as C

_h = None


def hip():
    global _h
    if _h is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        _h = h
    return _h


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    """hipMalloc'd bytes, freed on close or collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _check(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    def upload(self, arr) -> None:
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        _check(hip().hipMemcpy(self.ptr, C.c_void_p(arr.ctypes.data), arr.nbytes, 1),
               "hipMemcpy H2D")

    def download(self, arr) -> None:
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        _check(hip().hipMemcpy(C.c_void_p(arr.ctypes.data), self.ptr, arr.nbytes, 2),
               "hipMemcpy D2H")

    def close(self) -> None:
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
