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
