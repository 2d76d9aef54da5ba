"""TIFF export: ``class image_file`` (FBPG/image_file.h:19-46) over the C ABI
of include/tomokvas/fdk_tiff.h, and the export of a volume held in HBM.

``ImageFile`` keeps the reference's calls: ``AppendImages`` appends float
slices to a temporary file and tracks the running minimum and maximum,
``ConvertToTiff`` writes every slice as <output_dir>/%08d.tif scaled to 16 bits
with that range, and the temporary file is removed when the object goes
(image_file.cpp:27-30).  ``export_device_volume`` does the whole job for a
volume in device memory, with the range and the 16-bit scaling on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

BW16, BW8, RGB8 = 1, 2, 3     # image_file.h:13-15


class ImageFile:
    def __init__(self, tmp_filename: str, output_dir: str):
        self.filename, self.output_dir = tmp_filename, output_dir
        self.im_size_x = self.im_size_y = self.block_size = self.im_count = 0
        self.max = self.min = None
        self.tiff_type = BW16
        self._L = _lib.load()

    def SetImageSize(self, x: int, y: int) -> None:
        self.im_size_x, self.im_size_y = int(x), int(y)
        self.block_size = self.im_size_x * self.im_size_y

    def SetTiffType(self, tiff_type: int) -> None:
        if tiff_type != BW16:
            raise NotImplementedError("only 16-bit grey TIFF (BW16) is written: the reference's "
                                      "BW8/RGB8 paths never convert their pixels")
        self.tiff_type = tiff_type

    def AppendImages(self, data: np.ndarray, cnt: int) -> int:
        """Append cnt slices (image_file.cpp:32-64); NaN does not move the range."""
        d = np.ascontiguousarray(data, np.float32).ravel()[: self.block_size * cnt]
        fin = d[~np.isnan(d)]
        if fin.size:
            lo, hi = float(fin.min()), float(fin.max())
            self.min = lo if self.min is None else min(self.min, lo)
            self.max = hi if self.max is None else max(self.max, hi)
        with open(self.filename, "ab") as f:
            f.write(d.tobytes())
        self.im_count += cnt
        return 0

    def ConvertToTiff(self) -> int:
        """image_file.cpp:78-128: every appended slice as %08d.tif."""
        if self.im_count == 0:
            return 0
        vol = np.fromfile(self.filename, np.float32, count=self.block_size * self.im_count)
        lo = 0.0 if self.min is None else self.min
        hi = 0.0 if self.max is None else self.max
        _lib.check(self._L.fdk_tiff_export_host(vol.ctypes.data, self.im_size_x, self.im_size_y,
                                                self.im_count, lo, hi,
                                                self.output_dir.encode(), 0),
                   "fdk_tiff_export_host")
        return 0

    def __del__(self):
        try:
            os.remove(self.filename)
        except OSError:
            pass


def export_device_volume(d_vol: int, x: int, y: int, z: int, output_dir: str,
                         first_index: int = 0):
    """Write a device volume [z][y][x] as %08d.tif slices; returns (min, max)."""
    mm = (C.c_double * 2)()
    _lib.check(_lib.load().fdk_tiff_export_device(C.c_void_p(d_vol), x, y, z,
                                                   output_dir.encode(), first_index, mm),
               "fdk_tiff_export_device")
    return mm[0], mm[1]
