"""Scanco RSQ raw projections: ``class rsq`` (FBPG/rsq_load.h:63-135) over the
C ABI of include/tomokvas/fdk_rsq.h, with the preprocessing on the GPU.

``Rsq`` keeps the reference's method names and defaults.  ``LoadImages``
reads raw 16-bit sinograms; ``ProccessAndCopy`` uploads them, runs the
flat/dark + beam-hardening correction and the cut or half->full-fan rebinning
on the GPU and returns FP64 sinograms [n][sin_y_cpl][sin_x_cpl] as the
reference does; ``process_to_device`` writes FP32 (or FP64) straight into a
device buffer such as the engine's projection stack (``fdk_device_sinogram``),
whose layout [v][beta][u] is the sinogram order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._hip import DeviceBuffer


class Rsq:
    def __init__(self, filename: str):
        self.filename = filename
        self._L = _lib.load()
        self.p = _lib.RsqParams()
        _lib.check(self._L.fdk_rsq_default_params(C.byref(self.p)), "fdk_rsq_default_params")
        self._h = C.c_void_p()
        self.header = None
        self.detector_height = 248        # rsq_load.cpp:20
        self.zsize = 0
        self.idata = None
        self.first = 0

    # -- setters (rsq_load.h:78-92) ---------------------------------------------------
    def SetOffset(self, offset: int) -> None:
        self.p.header_offset = int(offset)

    def SetAxisMove(self, axis: int) -> None:
        self.p.axis_move = int(axis)

    def SetParallel(self, para: bool = True) -> None:
        if para:
            raise NotImplementedError("the fan->parallel rebinning (Half2Para) is not provided")

    def SetMinimal(self, minimal: bool = True) -> None:
        self.p.minimal = int(bool(minimal))

    def SetCompleting(self, complete: bool = True) -> None:
        self.p.complete = int(bool(complete))

    def SetFlip(self, flip: bool = True) -> None:
        self.p.flip = int(bool(flip))

    def SetInvert(self, invert: bool = True) -> None:
        self.p.invert = int(bool(invert))

    def SetDistance(self, distance: float) -> None:
        self.p.distance = float(distance)

    def SetDeltaFi(self, delta_fi: float) -> None:
        self.p.delta_fi = float(delta_fi)

    def SetDetectorHeight(self, height: int) -> None:
        self.detector_height = int(height)

    def GetDetectorHeight(self) -> int:
        return self.detector_height

    # -- file -------------------------------------------------------------------------
    def _open(self) -> None:
        if not self._h.value:
            _lib.check(self._L.fdk_rsq_open(self.filename.encode(), C.byref(self._h)),
                       "fdk_rsq_open")
            h = _lib.RsqHeader()
            _lib.check(self._L.fdk_rsq_get_header(self._h, C.byref(h)), "fdk_rsq_get_header")
            self.header = h

    def _sizes(self):
        self._open()
        n, sx, sy = C.c_int(), C.c_int(), C.c_int()
        _lib.check(self._L.fdk_rsq_sizes(self._h, C.byref(self.p), C.byref(n), C.byref(sx),
                                         C.byref(sy)), "fdk_rsq_sizes")
        return n.value, sx.value, sy.value

    def GetNum(self) -> int:
        return self._sizes()[0]

    def GetSize(self):
        """(sin_x_cpl, sin_y_cpl): the processed sinogram size."""
        _, sx, sy = self._sizes()
        return sx, sy

    def GetNumLoaded(self) -> int:
        return 0 if self.idata is None else self.idata.shape[0]

    def GetNoOfWholeBlocks(self) -> int:
        return self.zsize // self.detector_height if self.detector_height else 0

    def LoadImages(self, first: int | None = None, last: int | None = None) -> int:
        """LoadImages(first, last), or every sinogram without arguments
        (rsq_load.cpp:161-268); 0 on success, -1 on a bad range or short file."""
        if first is not None and last is None:   # LoadImages(unsigned): one whole detector
            h = self.detector_height                 # before the header sets it, as :163-168
            first, last = first * h, h - 1 + first * h
        self._open()
        self.detector_height = self.header.nr_of_projections
        self.zsize = self.header.dimz_p
        per = 2 * self.p.sin_x * (self.p.sin_y + 2)
        n = max(self.header.file_size - self.p.header_offset, 0) // per
        if first is None:
            first, last = 0, n - 1
        if last >= n or last < first or first < 0:
            return -1
        buf = np.empty((last + 1 - first, self.p.sin_y + 2, self.p.sin_x), np.uint16)
        if self._L.fdk_rsq_load(self._h, C.byref(self.p), first, last, buf.ctypes.data) != 0:
            self.idata = None
            return -1
        self.idata, self.first = buf, first
        return 0

    def LoadImagesUpper(self, k: int) -> int:
        h = self.detector_height
        return self.LoadImages((h - 1) // 2 + k * h, h - 1 + k * h)

    def LoadImagesLower(self, k: int) -> int:
        h = self.detector_height
        return self.LoadImages(k * h, (h - 1) // 2 + k * h)

    # -- preprocessing on the GPU ---------------------------------------------------------
    def process_to_device(self, d_out: int, dtype: int = _lib.F32, stream=None) -> None:
        """Process the loaded sinograms into device memory at d_out
        ([n][sin_y_cpl][sin_x_cpl] of dtype, FDK_F32 or FDK_F64)."""
        if self.idata is None:
            raise RuntimeError("LoadImages first")
        raw = DeviceBuffer(self.idata.nbytes)
        try:
            raw.upload(self.idata)
            _lib.check(self._L.fdk_rsq_process(C.byref(self.p), self.header.bh_a,
                                               self.header.bh_b, raw.ptr, self.idata.shape[0],
                                               dtype, C.c_void_p(d_out), stream),
                       "fdk_rsq_process")
            from ._hip import hip
            hip().hipDeviceSynchronize()
        finally:
            raw.close()

    def ProccessAndCopy(self, out: np.ndarray | None = None) -> np.ndarray:
        """The loaded sinograms as FP64 [n][sin_y_cpl][sin_x_cpl]
        (rsq_load.cpp:270-297)."""
        sx, sy = self.GetSize()
        n = self.GetNumLoaded()
        if out is None:
            out = np.empty((n, sy, sx), np.float64)
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.size != n * sy * sx:
            raise ValueError("out must be a C-contiguous float64 array [n][sin_y_cpl][sin_x_cpl]")
        dev = DeviceBuffer(out.nbytes)
        try:
            self.process_to_device(dev.ptr.value, _lib.F64)
            dev.download(out)
        finally:
            dev.close()
        return out

    def close(self) -> None:
        if self._h.value:
            self._L.fdk_rsq_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
