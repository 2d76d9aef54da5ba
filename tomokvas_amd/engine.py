"""Host-side mirror of the reference engine interface over the C ABI.

``FilteredBackProjection`` keeps the member names, argument order, argument
meaning and status-code convention of ``class FilteredBackProjection``
(FBPG/kernel.cuh:82-151) as FBPG/para_TC.cpp:130-153 calls it: setters, then
``Initialize()``, then ``CalculateReconstruction()``; 0 means success.  The
``bin_siz`` output sink of ``SetReconstructionProperties`` becomes any object
with ``GetBlock(slab, collecting) -> float32 buffer of Nx*Ny*B``
(FBPG/mucat_file.h:155); ``VolumeSink`` is an in-memory one.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import F16, F32, BF16, F64, FILTER_RAMLAK, FILTER_SHEPPLOGAN, FILTER_USER  # noqa: F401


class VolumeSink:
    """In-memory stand-in for ``bin_siz`` (FBPG/mucat_file.cpp:576-642).

    Holds the whole volume [Nz][Ny][Nx]; ``GetBlock(n, collecting)`` returns a
    host buffer for slab n.  With ``collecting`` the buffer is pre-filled with
    the slab's current contents (the partial sum of earlier angle blocks), as
    the reference's I/O thread prefetches it from the .bin file.
    """

    def __init__(self, Nx: int, Ny: int, Nz: int, B: int):
        self.Nx, self.Ny, self.Nz, self.B = Nx, Ny, Nz, B
        self.volume = np.zeros((Nz, Ny, Nx), np.float32)
        self._buf = np.zeros(B * Ny * Nx, np.float32)
        self._pending = None
        self.calls = []

    def _flush(self):
        if self._pending is not None:
            n = self._pending
            z0 = n * self.B
            nz = min(self.B, self.Nz - z0)
            self.volume[z0:z0 + nz] = self._buf[: nz * self.Ny * self.Nx].reshape(nz, self.Ny, self.Nx)
            self._pending = None

    def GetBlock(self, n: int, collecting: bool):
        self._flush()
        self.calls.append((n, bool(collecting)))
        z0 = n * self.B
        nz = min(self.B, self.Nz - z0)
        self._buf[:] = 0
        if collecting:
            self._buf[: nz * self.Ny * self.Nx] = self.volume[z0:z0 + nz].ravel()
        self._pending = n
        return self._buf

    def finish(self):
        self._flush()
        return self.volume


class FilteredBackProjection:
    """Python mirror of the reference engine (FBPG/kernel.cuh:82-151)."""

    def __init__(self, gpu_id: int = 0):
        self._L = _lib.load()
        h = C.c_void_p()
        _lib.check(self._L.fdk_create(gpu_id, C.byref(h)), "fdk_create")
        self._h = h
        self._sino_ref = None
        self._sink = None
        self._cb = None

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.fdk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- setters (kernel.cuh:89-101) ---------------------------------------------------
    def SetGPUid(self, gpu_id: int = 0) -> int:
        return self._L.fdk_set_gpu_id(self._h, gpu_id)

    def SetBlockSize(self, sz: int = 16) -> int:
        return self._L.fdk_set_block_size(self._h, sz)

    def SetAllocationLimit(self, nbytes: int = 1 << 30) -> int:
        return self._L.fdk_set_allocation_limit(self._h, nbytes)

    def SetCollecting(self, collecting: bool = False) -> int:
        return self._L.fdk_set_collecting(self._h, int(bool(collecting)))

    def SetShowInfo(self, show: bool = False) -> int:
        return self._L.fdk_set_show_info(self._h, int(bool(show)))

    def SetAxisDirection(self, X: bool = False, Y: bool = False, Z: bool = False) -> int:
        return self._L.fdk_set_axis_direction(self._h, int(X), int(Y), int(Z))

    def SetSinogramProperties(self, sinogram, sinogram_X: int, centre_X: float, sinogram_Y: int,
                              sinogram_Z: int, centre_Z: float, delta_fi: float, distance: float,
                              sampling: float) -> int:
        """``sinogram``: float32 [Nv][Nang][Nu] host array, or None when the
        projections will be written straight into device memory."""
        ptr = None
        if sinogram is not None:
            arr = np.ascontiguousarray(sinogram, dtype=np.float32)
            if arr.size != sinogram_X * sinogram_Y * sinogram_Z:
                return _lib.FDK_ERR_INVALID_ARGUMENT
            self._sino_ref = arr
            ptr = arr.ctypes.data
        return self._L.fdk_set_sinogram_properties(
            self._h, ptr, sinogram_X, centre_X, sinogram_Y, sinogram_Z, centre_Z, delta_fi,
            distance, sampling)

    def SetReconstructionProperties(self, output_file, reconstruction_X: int, centre_X: float,
                                    reconstruction_Y: int, centre_Y: float,
                                    reconstruction_Z: int, centre_Z: float, rot: float) -> int:
        self._sink = output_file
        return self._L.fdk_set_reconstruction_properties(
            self._h, reconstruction_X, centre_X, reconstruction_Y, centre_Y, reconstruction_Z,
            centre_Z, rot)

    def SetScaleXYZ(self, scale: float) -> int:
        return self._L.fdk_set_scale_xyz(self._h, scale)

    # -- extensions -----------------------------------------------------------------
    def SetFilter(self, kind: int = FILTER_RAMLAK, taps=None) -> int:
        if taps is None:
            return self._L.fdk_set_filter(self._h, kind, None, 0)
        t = np.ascontiguousarray(taps, np.float64)
        self._taps_ref = t
        return self._L.fdk_set_filter(self._h, kind, t.ctypes.data, t.size)

    def SetStorage(self, dtype: int = F32) -> int:
        return self._L.fdk_set_storage(self._h, dtype)

    def SetVolumeMode(self, mode: int = _lib.VOLUME_RESIDENT) -> int:
        """VOLUME_RESIDENT (whole volume in HBM) or VOLUME_STREAMED (two chunks
        in HBM, slabs leave through the sink while the next chunk computes)."""
        return self._L.fdk_set_volume_mode(self._h, mode)

    def GetVolumeChunk(self) -> int:
        n = C.c_int()
        _lib.check(self._L.fdk_get_volume_chunk(self._h, C.byref(n)), "fdk_get_volume_chunk")
        return n.value

    # -- run ----------------------------------------------------------------------------
    def Initialize(self) -> int:
        rc = self._L.fdk_initialize(self._h)
        self._sino_ref = None     # borrowed only until Initialize returns
        return rc

    def _with_sink(self, run) -> int:
        """run(get_block_fn, user) with the sink of SetReconstructionProperties."""
        sink = self._sink
        if sink is None:
            # no sink: the volume stays in HBM (device_volume_ptr)
            return run(_lib.GET_BLOCK_FN(), None)
        native = getattr(sink, "native_get_block", None)
        if native is not None:
            # a C writer (fileio.BinSiz): the engine calls it directly
            fn, user = native
            return run(fn, user)
        keep = {}

        def get_block(user, n, collecting):
            buf = sink.GetBlock(n, bool(collecting))
            arr = np.ascontiguousarray(buf, dtype=np.float32)
            if arr is not buf:        # the engine must write into the sink's own memory
                return None
            keep["buf"] = arr
            return arr.ctypes.data

        cb = _lib.GET_BLOCK_FN(get_block)
        self._cb = cb
        rc = run(cb, None)
        if hasattr(sink, "finish"):
            sink.finish()
        return rc

    def CalculateReconstruction(self) -> int:
        return self._with_sink(
            lambda fn, user: self._L.fdk_calculate_reconstruction(self._h, fn, user))

    def ReconstructHost(self, sinogram) -> int:
        """Initialize + CalculateReconstruction with the host stack
        [Nv][Nang][Nu] uploaded in angle chunks behind the computation (set the
        sinogram properties with None first)."""
        arr = np.ascontiguousarray(sinogram, dtype=np.float32)
        return self._with_sink(
            lambda fn, user: self._L.fdk_reconstruct_host(self._h, arr.ctypes.data, fn, user))

    @staticmethod
    def GetFreeMemInfo(gpu_id: int) -> int:
        return _lib.load().fdk_get_free_mem_info(gpu_id)

    # -- introspection ----------------------------------------------------------------------
    def stage_times_ms(self):
        out = (C.c_float * 4)()
        _lib.check(self._L.fdk_get_stage_times(self._h, out), "stage times")
        return list(out)

    def geometry(self) -> dict:
        iv = (C.c_int * 8)()
        fv = (C.c_float * 12)()
        _lib.check(self._L.fdk_get_geometry(self._h, iv, fv), "geometry")
        keys_i = ["Nu", "Nang", "Nv", "Nx", "Ny", "Nz", "B", "p"]
        keys_f = ["cX", "cZ", "dbeta", "D", "tau", "cx", "cy", "cz", "rot", "scale", "flips",
                  "storage"]
        d = dict(zip(keys_i, list(iv)))
        d.update(zip(keys_f, list(fv)))
        return d

    def device_volume_ptr(self) -> int:
        return self._L.fdk_device_volume(self._h) or 0

    def device_sinogram_ptr(self) -> int:
        return self._L.fdk_device_sinogram(self._h) or 0

    def prepare(self) -> None:
        _lib.check(self._L.fdk_prepare(self._h), "fdk_prepare")
