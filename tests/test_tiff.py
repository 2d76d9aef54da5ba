"""TIFF export on the host (include/tomokvas/fdk_tiff.h, tomokvas_amd/
imagefile.py), CPU only: the file structure read back with a reader written
from the TIFF 6.0 specification, the 16-bit scaling against the oracle, and
image_file's append/convert flow.  The GPU path is in tests/test_gpu_tiff.py."""
import os

import numpy as np
import pytest

from oracle import image as O
from tiff_util import read_tiff
from tomokvas_amd import _lib
from tomokvas_amd.imagefile import BW8, ImageFile


def test_scaling_closed_forms():
    x = np.array([-2.0, 6.0, 2.0, np.nan, 9.0, -5.0], np.float32)
    u = O.to_u16(x, -2.0, 6.0)
    assert list(u) == [0, 65535, 32767, 0, 65535, 0]
    assert list(O.to_u16(x, 1.0, 1.0)) == [0] * 6


@pytest.mark.parametrize("w,h", [(5, 1), (7, 3), (64, 48)])
def test_writer_structure_and_pixels(tmp_path, w, h):
    img = np.random.default_rng(w * h).integers(0, 65536, (h, w)).astype(np.uint16)
    path = str(tmp_path / "a.tif")
    _lib.check(_lib.load().fdk_tiff_write_u16(path.encode(), img.ctypes.data, w, h))
    tags, got = read_tiff(path)
    assert np.array_equal(got, img)
    # the fields TiffWrite sets for BW16 (image_file.cpp:142-183)
    assert tags[256] == [w] and tags[257] == [h] and tags[262] == [1]
    assert tags[278] == [1] and tags[284] == [1] and tags[279] == [2 * w] * h
    assert sorted(tags) == [256, 257, 258, 259, 262, 273, 277, 278, 279, 284]
    assert all(o % 2 == 0 for o in tags[273])       # word-aligned strips


def test_image_file_append_and_convert(tmp_path):
    out = tmp_path / "tif"
    out.mkdir()
    tmp = str(tmp_path / "slices.tmp")
    f = ImageFile(tmp, str(out))
    f.SetImageSize(6, 4)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((3, 4, 6)).astype(np.float32)
    b = (rng.standard_normal((2, 4, 6)) * 3).astype(np.float32)
    b[1, 2, 3] = np.nan
    f.AppendImages(a, 3)
    f.AppendImages(b, 2)
    assert f.im_count == 5
    f.ConvertToTiff()
    vol = np.concatenate([a, b])
    lo, hi = np.nanmin(vol), np.nanmax(vol)
    assert (f.min, f.max) == (lo, hi)
    for k in range(5):
        _, img = read_tiff(str(out / f"{k:08d}.tif"))
        assert np.array_equal(img, O.to_u16(vol[k], lo, hi))
    del f
    assert not os.path.exists(tmp)


def test_only_sixteen_bit_grey(tmp_path):
    f = ImageFile(str(tmp_path / "t"), str(tmp_path))
    with pytest.raises(NotImplementedError):
        f.SetTiffType(BW8)
    with pytest.raises(_lib.FDKError):
        _lib.check(_lib.load().fdk_tiff_write_u16(str(tmp_path / "x.tif").encode(), None, 0, 0))


def test_host_export_many_slices_and_failure(tmp_path):
    """fdk_tiff_export_host writes its slices on several threads: every file is
    there with the oracle's pixels, and a directory that does not exist fails
    with a message naming a file in it."""
    L = _lib.load()
    rng = np.random.default_rng(7)
    vol = rng.standard_normal((37, 5, 9)).astype(np.float32)
    lo, hi = float(vol.min()), float(vol.max())
    out = tmp_path / "many"
    out.mkdir()
    _lib.check(L.fdk_tiff_export_host(vol.ctypes.data, 9, 5, 37, lo, hi, str(out).encode(), 11))
    assert sorted(os.listdir(out)) == [f"{k:08d}.tif" for k in range(11, 48)]
    for k in (0, 17, 36):
        _, img = read_tiff(str(out / f"{k + 11:08d}.tif"))
        assert np.array_equal(img, O.to_u16(vol[k], lo, hi))
    missing = tmp_path / "absent"
    with pytest.raises(_lib.FDKError, match="absent"):
        _lib.check(L.fdk_tiff_export_host(vol.ctypes.data, 9, 5, 37, lo, hi,
                                          str(missing).encode(), 0))
