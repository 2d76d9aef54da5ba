"""TIFF export from a volume in HBM (fdk_tiff_export_device): range and 16-bit
scaling on the GPU must write the same files as the host path, byte for byte,
and match the oracle's scaling."""
import ctypes as C
import filecmp
import os

import numpy as np
import pytest

from oracle import image as O
from tiff_util import read_tiff
from tomokvas_amd import _lib
from tomokvas_amd._hip import DeviceBuffer, hip
from tomokvas_amd.imagefile import export_device_volume

from gpu_util import engine, say, scan

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(3, 5, 7), (70, 33, 40)])
def test_device_export_equals_host_export(tmp_path, shape):
    z, y, x = shape
    rng = np.random.default_rng(z)
    vol = (rng.standard_normal(shape) * 4).astype(np.float32)
    vol[0, 0, 0], vol[-1, -1, -1] = np.nan, -0.0
    dev = DeviceBuffer(vol.nbytes)
    dev.upload(vol)
    say(f"tiff export {shape}")
    a, b = tmp_path / "dev", tmp_path / "host"
    a.mkdir()
    b.mkdir()
    lo, hi = export_device_volume(dev.ptr.value, x, y, z, str(a), first_index=3)
    dev.close()
    assert (lo, hi) == (float(np.nanmin(vol)), float(np.nanmax(vol)))
    _lib.check(_lib.load().fdk_tiff_export_host(vol.ctypes.data, x, y, z, lo, hi,
                                                str(b).encode(), 3))
    names = sorted(os.listdir(a))
    assert names == [f"{k:08d}.tif" for k in range(3, 3 + z)] == sorted(os.listdir(b))
    assert all(filecmp.cmp(a / n, b / n, shallow=False) for n in names)
    _, img = read_tiff(str(a / names[-1]))
    assert np.array_equal(img, O.to_u16(vol[-1], lo, hi))


def test_reconstruction_to_tiff(tmp_path):
    g = scan(48, 40, 30, 48, 44, 20, B=8)
    s = np.random.default_rng(2).standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    fbp = engine(g, s, None)
    _lib.check(fbp.Initialize())
    _lib.check(fbp.CalculateReconstruction())
    ptr = fbp.device_volume_ptr()
    vol = np.empty((g.Nz, g.Ny, g.Nx), np.float32)
    assert hip().hipMemcpy(C.c_void_p(vol.ctypes.data), C.c_void_p(ptr), vol.nbytes, 2) == 0
    lo, hi = export_device_volume(ptr, g.Nx, g.Ny, g.Nz, str(tmp_path))
    fbp.close()
    assert (lo, hi) == (float(vol.min()), float(vol.max()))
    for k in (0, g.Nz // 2, g.Nz - 1):
        _, img = read_tiff(str(tmp_path / f"{k:08d}.tif"))
        assert np.array_equal(img, O.to_u16(vol[k], lo, hi))


@pytest.mark.parametrize("n", [1, 6, 4 * 1000 + 3, 100003])
def test_range_and_scaling_at_any_alignment(n):
    """fdk_volume_minmax and fdk_volume_to_u16 read four floats at a time when
    the pointers allow it and one at a time otherwise; every start offset and a
    length with a remainder give the oracle's range and 16-bit values."""
    L = _lib.load()
    rng = np.random.default_rng(n)
    vol = (rng.standard_normal(n + 8) * 3).astype(np.float32)
    vol[::7] = np.nan
    vol[1] = -0.0
    dev = DeviceBuffer(vol.nbytes)
    out = DeviceBuffer(2 * (n + 8))
    dev.upload(vol)
    try:
        for off in (0, 1, 2, 3, 4):
            x = vol[off:off + n]
            say(f"tiff range/scaling n={n} float offset {off}")
            mm = (C.c_double * 2)()
            rc = L.fdk_volume_minmax(C.c_void_p(dev.ptr.value + 4 * off), n, mm, None)
            if np.isnan(x).all():
                assert rc == _lib.FDK_ERR_INVALID_ARGUMENT   # no value to scale by
                continue
            assert rc == 0
            lo, hi = float(np.nanmin(x)), float(np.nanmax(x))
            assert (mm[0], mm[1]) == (lo, hi), off
            for oo in (0, 1):
                _lib.check(L.fdk_volume_to_u16(C.c_void_p(dev.ptr.value + 4 * off), n, lo, hi,
                                               C.c_void_p(out.ptr.value + 2 * oo), None))
                hip().hipDeviceSynchronize()
                got = np.empty(n + 8, np.uint16)
                out.download(got)
                assert np.array_equal(got[oo:oo + n], O.to_u16(x, lo, hi)), (off, oo)
    finally:
        dev.close()
        out.close()
