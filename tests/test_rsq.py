"""The RSQ reader's host side (include/tomokvas/fdk_rsq.h), CPU only: header
and beam-hardening line against the oracle's parse, sizes, the raw load, and
errors.  The GPU preprocessing is tested in tests/test_gpu_rsq.py."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import rsq as R
from rsq_util import synthetic_raw, write_rsq
from tomokvas_amd import _lib
from tomokvas_amd.rsq import Rsq


def test_header_and_sizes(tmp_path):
    raw = synthetic_raw(5, 40, 24, seed=1)
    path = str(tmp_path / "s.rsq")
    write_rsq(path, raw, dimz=10, nr_of_projections=2, tail=100)
    r = Rsq(path)
    r.p.sin_x, r.p.sin_y = 24, 40
    h = R.parse_header(open(path, "rb").read(4096))
    assert r.GetNum() == 5                        # the 100 spare bytes are no sinogram
    hd = r.header
    assert hd.check == h["check"] and (hd.dimx_p, hd.dimz_p) == (24, 10)
    assert hd.nr_of_projections == 2 and hd.data_offset == h["data_offset"]
    assert list(hd.bh_x) == h["bh_x"] and list(hd.bh_y) == h["bh_y"]
    assert (hd.bh_a, hd.bh_b) == (h["bh_a"], h["bh_b"])
    assert hd.file_size == R.DATA_OFFSET + raw.nbytes + 100
    for complete, minimal, axis in [(False, False, 3), (True, False, 3), (True, True, 2)]:
        r.SetCompleting(complete)
        r.SetMinimal(minimal)
        r.SetAxisMove(axis)
        r.SetDistance(50.0)
        r.SetDeltaFi(2 * math.pi / 40)
        assert r.GetSize() == R.sizes(24, 40, axis, complete, minimal, 50.0, 2 * math.pi / 40)


def test_defaults_are_the_reference_constructor():
    p = _lib.RsqParams()
    _lib.check(_lib.load().fdk_rsq_default_params(C.byref(p)))
    assert (p.sin_x, p.sin_y, p.header_offset, p.axis_move) == (1024, 2000, 3584, 21)
    assert p.distance == 125810. / 24. and p.delta_fi == 2 * math.pi / 2000
    assert not (p.flip or p.invert or p.complete or p.minimal)


def test_load_ranges(tmp_path):
    raw = synthetic_raw(6, 30, 16, seed=2)
    path = str(tmp_path / "s.rsq")
    write_rsq(path, raw, dimz=6, nr_of_projections=3)
    r = Rsq(path)
    r.p.sin_x, r.p.sin_y = 16, 30
    assert r.LoadImages(2, 4) == 0 and r.GetNumLoaded() == 3
    assert np.array_equal(r.idata, raw[2:5])
    assert r.LoadImages() == 0 and np.array_equal(r.idata, raw)
    # whole detectors of nr_of_projections = 3 rows once the header is read
    assert r.LoadImages(1) == 0 and np.array_equal(r.idata, raw[3:6])
    assert r.GetNoOfWholeBlocks() == 2
    assert r.LoadImagesLower(0) == 0 and np.array_equal(r.idata, raw[0:2])
    assert r.LoadImagesUpper(0) == 0 and np.array_equal(r.idata, raw[1:3])
    assert r.LoadImages(4, 6) == -1 and r.LoadImages(3, 2) == -1


def test_errors(tmp_path):
    with pytest.raises(_lib.FDKError):
        Rsq(str(tmp_path / "none.rsq")).GetNum()
    path = str(tmp_path / "nobh.rsq")
    write_rsq(path, synthetic_raw(1, 4, 4, seed=3))
    data = bytearray(open(path, "rb").read())
    data[R.BH_OFFSET:R.BH_OFFSET + 8] = b"Nothing "
    open(path, "wb").write(bytes(data))
    with pytest.raises(_lib.FDKError):
        Rsq(path).GetNum()
    with pytest.raises(NotImplementedError):
        Rsq(path).SetParallel(True)
