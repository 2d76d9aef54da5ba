"""The threaded host code under ThreadSanitizer, on the CPU: the .bin
writer's I/O thread (collecting slabs over angle blocks), the reconstruction
program's angle-block loader thread next to it (fdk_reconstruct_file over a
stub engine that follows the engine's slab contract), the .CRA reader's
parallel positioned reads, and the TIFF export's writer threads.  Built by g++
with -fsanitize=thread from the library's host sources (tests/cpp/build_tsan.py);
a data race, or any other ThreadSanitizer report, fails the test.  The race the
reference's writer has (bin_siz's sync_flag, set by its I/O thread itself,
FBPG/mucat_file.cpp:576-642) is the kind this catches."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpp"))

DX, DZ, N, SCANS = 96, 64, 60, 2


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    import build_tsan
    path = build_tsan.build(str(tmp_path_factory.mktemp("tsan") / "tsan_host"))
    # instrumented: the ThreadSanitizer runtime is linked in
    syms = subprocess.run(["nm", path], capture_output=True, text=True, check=True).stdout
    assert "__tsan_func_entry" in syms or "__tsan_init" in syms
    return path


def run(exe, *args):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66 second_deadlock_stack=1")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300, env=env)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-6000:]
    assert r.returncode == 0, r.stderr[-3000:]
    return r.stdout


def test_bin_writer_io_thread(exe, tmp_path):
    run(exe, "bin", str(tmp_path))
    x, y, z, B = 7, 5, 13, 3
    vol = np.fromfile(tmp_path / "tsan.bin", np.float32).reshape(3, z, y, x)
    want = np.repeat(3.0 + 2.0 * np.arange(-(-z // B)), B)[:z]      # slab s: (1 + s) + (2 + s)
    assert np.array_equal(vol, np.broadcast_to(want[None, :, None, None], vol.shape))


def test_program_loader_and_writer_threads(exe, tmp_path):
    lines = ["FILE NAME:              scan.raw", "SOURCE NAME:            tsan",
             "SAMPLE SPACING:         50", "DISTANCE:               1",
             f"DETECTOR X PIXELS:      {DX}", f"DETECTOR Z PIXELS:      {DZ}",
             f"NUMBER OF PROJECTIONS:  {N}", "X DIMENSION:            24",
             "Z DIMENSION:            20", "ROTATE ANGLE:           0",
             f"NUMBER OF BLOCKS:       {SCANS}", "CENTRE:                 47.5",
             "CENTRE:                 48"]
    (tmp_path / "scan.CRA").write_text("\n".join(lines) + "\n")
    rng = np.random.default_rng(4)
    raw = rng.uniform(0, 1, (SCANS * N, DZ, DX)).astype(np.float32)
    raw.tofile(tmp_path / "scan.raw")
    per = 25                                   # projections per angle block: 25, 25, 10
    out = run(exe, "driver", str(tmp_path / "scan.CRA"), str(per * DX * DZ * 4))
    assert f"scan blocks {SCANS} angle blocks 3" in out
    rx, ry, rz = [int(v) for v in (tmp_path / "scan.siz").read_text().split()[:3]]
    rz //= SCANS
    vol = np.fromfile(tmp_path / "scan.bin", np.float32).reshape(SCANS, rz, ry, rx)
    for s in range(SCANS):
        # the stub engine adds each angle block's projection sum to every voxel
        acc = np.float32(0)
        for a0 in range(0, N, per):
            blk = raw[s * N + a0: s * N + min(a0 + per, N)].astype(np.float64).sum()
            acc = np.float32(acc + np.float32(blk))
        assert np.all(vol[s] == acc), s


def test_tiff_writer_threads(exe, tmp_path):
    run(exe, "tiff", str(tmp_path))
    files = sorted(os.listdir(tmp_path))
    assert files == [f"{k:08d}.tif" for k in range(40)]
    from tiff_util import read_tiff
    x, y = 33, 17
    i = np.arange(x * y * 40, dtype=np.uint64)
    vol = ((i * 2654435761) % 1000).astype(np.float32) / np.float32(999.0)
    for k in (0, 17, 39):
        _, img = read_tiff(tmp_path / files[k])
        s = vol[k * x * y:(k + 1) * x * y].astype(np.float64)
        want = np.clip(65535 * s, 0, 65535).astype(np.uint16).reshape(y, x)
        assert np.array_equal(img, want), k
