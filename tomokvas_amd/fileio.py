"""The reference's projection container and volume file, over the C ABI of
include/tomokvas/fdk_io.h (host code).

``CraCon`` mirrors ``cra_con`` (FBPG/mucat_file.h:37-95): the .CRA text header,
the split of the angles into blocks that fit a byte budget, and the loading of
one block sinogram-major [v][beta][u].  ``BinSiz`` mirrors ``bin_siz``
(FBPG/mucat_file.h:129-193): the .bin volume and .siz header, slab buffers
handed to the engine with the previous slab written and the next prefetched on
an I/O thread.  A ``BinSiz`` passed to ``FilteredBackProjection.
SetReconstructionProperties`` is called by the engine natively (no Python in
the slab loop).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class CraCon:
    """Reader of <path>/<filename>.CRA and the raw projection file it names."""

    def __init__(self, path: str, filename: str):
        self.path, self.filename = path, filename
        self._L = _lib.load()
        self._h = C.c_void_p()
        self._meta = None
        self._scan = 0
        self._nblocks = 1
        self._sin_block_size = 5 * 1024 ** 3   # the reference's default (mucat_file.h:63)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.fdk_cra_close(self._h)
            self._h = C.c_void_p()

    def _need(self):
        if not self._h.value:
            raise RuntimeError("call GetMetaData first")

    # GetMetaData (mucat_file.cpp:60-160): 0, or the error code
    def GetMetaData(self) -> int:
        if self._h.value:
            self._L.fdk_cra_close(self._h)
            self._h = C.c_void_p()
        rc = self._L.fdk_cra_open(self.path.encode(), self.filename.encode(), C.byref(self._h))
        if rc == _lib.FDK_OK:
            m = _lib.CraMeta()
            _lib.check(self._L.fdk_cra_get_meta(self._h, C.byref(m)), "fdk_cra_get_meta")
            self._meta = m
        return rc

    @property
    def meta(self) -> "_lib.CraMeta":
        self._need()
        return self._meta

    def PrintMetaData(self) -> None:
        m = self.meta
        print("\nMetaData: ")
        print(f"   file name:             {m.raw_file.decode()}")
        print(f"   source name:           {m.source_name.decode()}")
        print(f"   sample spacing:        {m.sample_spacing} u")
        print(f"   distance:              {m.distance} c")
        print(f"   detector x pixels:     {m.detector_x}")
        print(f"   detector y pixels:     {m.detector_z}")
        print(f"   number of projections: {m.num_of_projections}")
        print(f"   number of blocks:      {m.num_of_blocks}")
        print(f"   x dimension:           {m.x_dimension}")
        print(f"   z dimension:           {m.z_dimension}")
        print(f"   rotate angle:          {m.rotate_angle}")
        print(f"   clockwise:             {int(m.clockwise)}")
        for i in range(m.num_of_blocks):
            print(f"   block[{i}] centre:       {m.centre_of_rotation[i]}")
        print(f"   reconstruction scale:  {m.scale}")

    # GetSize / GetRecSize (mucat_file.cpp:265-282)
    def GetSize(self):
        m = self.meta
        return m.detector_x, m.num_of_projections, m.detector_z

    def GetRecSize(self):
        m = self.meta
        rx = m.detector_x if m.x_dimension == 0 else m.x_dimension
        return rx, rx, m.z_dimension

    def GetCentreOfRotation(self) -> float:
        """Centre of the active scan block (mucat_file.cpp:284-286)."""
        return self.meta.centre_of_rotation[self._scan]

    def GetDistance(self) -> float:
        return self.meta.distance_px

    def GetDFI(self) -> float:
        return self.meta.delta_fi

    def GetSampling(self) -> float:
        return self.meta.sample_spacing

    def GetScale(self) -> float:
        return self.meta.scale

    def GetNumberOfBlocks(self) -> int:
        return self.meta.num_of_blocks

    def SetActiveBlock(self, block: int) -> int:
        self._need()
        self._scan = self._L.fdk_cra_set_scan_block(self._h, int(block))
        return self._scan

    def SetSinBlockMaxSize(self, size: int = 5 * 1024 ** 3) -> int:
        self._sin_block_size = int(size)
        return self._sin_block_size

    def CalculateSinBlocks(self) -> int:
        self._need()
        n = C.c_int()
        _lib.check(self._L.fdk_cra_split(self._h, self._sin_block_size, C.byref(n)), "fdk_cra_split")
        self._nblocks = n.value
        return n.value

    def GetSinNoBlocks(self) -> int:
        return self._nblocks

    def SetSinActiveBlock(self, block: int) -> int:
        self._need()
        return self._L.fdk_cra_set_angle_block(self._h, int(block))

    def GetSinBlockSize(self):
        self._need()
        n = C.c_int()
        _lib.check(self._L.fdk_cra_block_projections(self._h, C.byref(n)), "fdk_cra_block_projections")
        m = self.meta
        return m.detector_x, n.value, m.detector_z

    def GetRotRad(self) -> float:
        self._need()
        r = C.c_double()
        _lib.check(self._L.fdk_cra_rotation(self._h, C.byref(r)), "fdk_cra_rotation")
        return r.value

    def LoadData(self, out: np.ndarray | None = None) -> np.ndarray:
        """The active angle block as float32 [detector_z][n][detector_x]."""
        x, n, z = self.GetSinBlockSize()
        if out is None:
            out = np.empty((z, n, x), np.float32)
        if out.dtype != np.float32 or not out.flags.c_contiguous or out.size != z * n * x:
            raise ValueError("out must be a C-contiguous float32 array of the block's size")
        _lib.check(self._L.fdk_cra_load(self._h, out.ctypes.data), "fdk_cra_load")
        return out


class BinSiz:
    """Writer of <path>/<filename>.bin and .siz."""

    def __init__(self, path: str, filename: str):
        self.path, self.filename = path, filename
        self._L = _lib.load()
        self._h = C.c_void_p()
        self.im_x = self.im_y = self.im_z = 0
        self.scan_blocks = 1
        self.slices_per_block = 16

    def SetScanBlocksCount(self, sb: int) -> None:
        self.scan_blocks = int(sb)

    def SetImageSize(self, x: int, y: int, z: int) -> None:
        self.im_x, self.im_y, self.im_z = int(x), int(y), int(z)

    def SetSlicesPerBlock(self, z: int = 16) -> None:
        """Sizes are complete: creates the .bin file at its full size."""
        self.slices_per_block = int(z)
        if self._h.value:
            raise RuntimeError("the volume file is already open")
        _lib.check(self._L.fdk_bin_open(self.path.encode(), self.filename.encode(), self.im_x,
                                        self.im_y, self.im_z, self.scan_blocks,
                                        self.slices_per_block, C.byref(self._h)), "fdk_bin_open")

    def SetScanBlockOffset(self, offset: int) -> None:
        _lib.check(self._L.fdk_bin_set_scan_block(self._h, int(offset)), "fdk_bin_set_scan_block")

    def GetBufferSize(self) -> float:
        """MiB held in slab buffers (mucat_file.cpp:517-520)."""
        return 3 * self.slices_per_block * self.im_x * self.im_y * 4 / 1024.0 / 1024.0

    def GetBlock(self, n: int, collecting: bool = False) -> np.ndarray:
        p = self._L.fdk_bin_get_block(self._h, int(n), int(bool(collecting)))
        if not p:
            raise RuntimeError(f"GetBlock({n}) failed: {_lib.last_error()}")
        return np.ctypeslib.as_array(p, shape=(self.slices_per_block * self.im_x * self.im_y,))

    @property
    def native_get_block(self):
        """(fdk_get_block_fn, user) for fdk_calculate_reconstruction."""
        fn = C.cast(self._L.fdk_bin_get_block, _lib.GET_BLOCK_FN)
        return fn, self._h

    def FlushBuffer(self) -> None:
        _lib.check(self._L.fdk_bin_flush(self._h), "fdk_bin_flush")

    def close(self) -> None:
        """Flush, write the .siz header, close (the reference's destructor)."""
        if self._h.value:
            h, self._h = self._h, C.c_void_p()
            _lib.check(self._L.fdk_bin_close(h), "fdk_bin_close")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
