"""C-ABI library and host-side logic -- CPU only (no kernel launches)."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

from tomokvas_amd import _lib, geometry, phantoms
from tomokvas_amd.engine import VolumeSink


def test_library_exports_every_header_function():
    lib = _lib.load()
    names = _lib.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding declares every one of them
    assert set(names) == set(_lib._SIGS), set(names) ^ set(_lib._SIGS)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


def test_library_has_gfx950_code_object():
    # the embedded offload bundle names its target
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_header_compiles_as_c_and_cpp(tmp_path):
    inc = os.path.join(os.path.dirname(_lib.HERE), "include")
    src = tmp_path / "t.c"
    src.write_text('#include "tomokvas/fdk.h"\nint main(void){fdk_ctx* c=0;(void)c;return FDK_OK;}\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-c", str(src), "-o",
                    str(tmp_path / "t.o")], check=True)
    hpp = tmp_path / "t.cpp"
    hpp.write_text('#include "tomokvas/FilteredBackProjection.hpp"\nint main(){return 0;}\n')
    subprocess.run(["g++", "-std=c++14", "-Wall", "-Werror", "-I", inc, "-c", str(hpp), "-o",
                    str(tmp_path / "t2.o")], check=True)


def test_null_context_is_an_error_not_a_crash():
    lib = _lib.load()
    assert lib.fdk_set_block_size(None, 16) == _lib.FDK_ERR_INVALID_ARGUMENT
    assert "block size" in _lib.last_error()
    assert lib.fdk_initialize(None) == _lib.FDK_ERR_INVALID_ARGUMENT
    assert lib.fdk_destroy(None) == _lib.FDK_OK
    assert lib.fdk_poisson_log(None, 10, 1e5, 1, None) == _lib.FDK_ERR_INVALID_ARGUMENT


def test_flat_detector_conversion():
    # tau = pitch * SOD / SDD, D = SOD / tau  (virtual detector at the isocentre)
    g = geometry.flat_detector(128, 128, 180, 0.4, 500, 1000, 128, 128, 128, 1.0)
    assert g.tau == pytest.approx(0.02)          # cm
    assert g.D == pytest.approx(2500.0)
    assert g.cX == 63.5 and g.cZ == 63.5 and g.cx == 63.5
    assert g.p == 256
    assert g.dbeta == float(np.float32(2 * math.pi / 180))
    w = geometry.workloads()
    assert w["cfg3"].geom.voxel_updates == 1024 ** 3 * 1800
    assert w["cfg3"].geom.p == 4096 and w["cfg5"].geom.p == 4096 and w["cfg4"].geom.p == 512
    assert w["cfg2"].geom.fov_px == 1024 and w["cfg3"].geom.D == pytest.approx(10000.0)


def test_volume_sink_slab_semantics():
    s = VolumeSink(4, 3, 7, 3)
    for n in range(3):
        buf = s.GetBlock(n, False)
        assert buf.size == 4 * 3 * 3
        buf[:] = n + 1
    vol = s.finish()
    assert np.all(vol[0:3] == 1) and np.all(vol[3:6] == 2) and np.all(vol[6:7] == 3)
    # collecting: the buffer arrives holding the partial sum
    buf = s.GetBlock(1, True)
    assert np.all(buf[: 3 * 12] == 2)
    buf[:] += 10
    s.finish()
    assert np.all(s.volume[3:6] == 12)


def test_phantom_tables():
    sl = phantoms.shepp_logan(64.0)
    assert sl.shape == (10, 19)
    assert sl[0, 15] == 1.0 and sl[1, 15] == -0.98
    e = phantoms.random_ellipsoids(20251202, 64, 2048.0)
    assert e.shape == (64, 19)
    assert np.array_equal(e, phantoms.random_ellipsoids(20251202, 64, 2048.0))


def test_l2_probe_rejects_bad_footprints():
    # argument checks run before any HIP call
    lib = _lib.load()
    gbs = C.c_double(0.0)
    for fp in (0, 1000, 3 << 20, 128 << 20):
        assert lib.fdk_probe_l2_read(fp, C.byref(gbs)) == _lib.FDK_ERR_INVALID_ARGUMENT
    assert lib.fdk_probe_l2_read(2 << 20, None) == _lib.FDK_ERR_INVALID_ARGUMENT


def test_para_tc_host_program_builds_against_the_cpp_class(tmp_path):
    # instantiates SetReconstructionProperties<Sink> and the GetBlock trampoline
    # of the drop-in class and links against the library (the GPU test runs it)
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp"))
    import build_host
    exe = build_host.build(out=str(tmp_path / "para_tc_host"))
    assert os.path.getsize(exe) > 0
    env = dict(os.environ, LD_LIBRARY_PATH=_lib.HERE)   # the rpath is relative to tests/cpp
    r = subprocess.run([exe], capture_output=True, text=True, env=env)
    assert r.returncode == 2 and "usage" in r.stderr
