"""The reconstruction program end to end on the GPU: .CRA file -> angle blocks
-> engine (collecting) -> .bin/.siz, against one engine pass over each scan
block's whole stack.

Splitting the angles changes only the FP32 summation order and the float
rounding of each block's start angle (|d rot| <= 2.4e-7 rad here, a shift of
< 2e-5 px at the volume's edge); the projections are smooth, so the files agree
with the one-pass volume to 1e-5 of its largest value.  With FP16 storage the
filtered samples are rounded to FP16 after filtering, and the filter pairs
detector rows differently when a block holds fewer angles (two real rows per
complex transform), so a sample that lies next to an FP16 rounding boundary can
round to the other neighbour (2^-11 of it): 1e-4 there.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from tomokvas_amd import _lib
from tomokvas_amd.build import EXE
from tomokvas_amd.driver import LoadSetting, SaveSetting, default_setting, reconstruct_file
from tomokvas_amd.engine import FilteredBackProjection, VolumeSink
from tomokvas_amd.fileio import CraCon

from gpu_util import say

pytestmark = pytest.mark.gpu

DX, DZ, N, SCANS = 96, 64, 60, 2
CENTRES = (47.5, 48.25)


def smooth_projections(clockwise):
    """[scan*N + j][v][u]: a Gaussian whose centre moves with the angle."""
    j = np.arange(SCANS * N, dtype=np.float64)[:, None, None]
    v = np.arange(DZ)[None, :, None]
    u = np.arange(DX)[None, None, :]
    beta = 2 * math.pi * (j % N) / N * (-1 if clockwise else 1)
    u0 = 47.5 + 14 * np.sin(beta + 0.3) + (j // N)
    v0 = 31.5 + 6 * np.cos(beta)
    return (np.exp(-((u - u0) ** 2 / 180.0 + (v - v0) ** 2 / 120.0)) * (1.0 + 0.2 * (j // N))
            ).astype(np.float32)


def write_scan(folder, clockwise):
    lines = ["FILE NAME:              scan.raw", "SOURCE NAME:            test",
             "SAMPLE SPACING:         50", "DISTANCE:               1",
             f"DETECTOR X PIXELS:      {DX}", f"DETECTOR Z PIXELS:      {DZ}",
             f"NUMBER OF PROJECTIONS:  {N}", "X DIMENSION:            64",
             "Z DIMENSION:            40", "ROTATE ANGLE:           17.5",
             f"NUMBER OF BLOCKS:       {SCANS}"]
    lines += [f"CENTRE:                 {c}" for c in CENTRES]
    if clockwise:
        lines.append("CLOCKWISE:")
    with open(os.path.join(folder, "scan.CRA"), "w") as f:
        f.write("\n".join(lines) + "\n")
    smooth_projections(clockwise).tofile(os.path.join(folder, "scan.raw"))
    return os.path.join(folder, "scan.CRA")


def one_pass(folder, setting, storage):
    """Each scan block reconstructed from its whole stack in one engine run."""
    cra = CraCon(folder, "scan")
    assert cra.GetMetaData() == 0
    cra.SetSinBlockMaxSize(1 << 40)
    assert cra.CalculateSinBlocks() == 1
    rx, ry, rz = cra.GetRecSize()
    out = []
    for s in range(SCANS):
        cra.SetActiveBlock(s)
        cra.SetSinActiveBlock(0)
        data = cra.LoadData()
        sink = VolumeSink(rx, ry, rz, setting.slices_per_block)
        with FilteredBackProjection(setting.gpu_id) as e:
            _lib.check(e.SetBlockSize(setting.slices_per_block))
            _lib.check(e.SetAxisDirection(setting.x_direction, setting.y_direction,
                                          setting.z_direction))
            _lib.check(e.SetSinogramProperties(
                data, DX, cra.GetCentreOfRotation(), N, DZ,
                (DZ - 1) / 2.0 + setting.vertical_center, cra.GetDFI(), cra.GetDistance(),
                cra.GetSampling() * 1e-4))
            _lib.check(e.SetReconstructionProperties(sink, rx, (rx - 1) / 2.0, ry, (ry - 1) / 2.0,
                                                     rz, 0.0, cra.GetRotRad()))
            _lib.check(e.SetScaleXYZ(1.0))
            _lib.check(e.SetStorage({"f32": _lib.F32, "f16": _lib.F16}[storage]))
            _lib.check(e.Initialize())
            _lib.check(e.CalculateReconstruction())
        out.append(sink.finish())
    return np.stack(out)


def read_volume(folder, shape):
    siz = open(os.path.join(folder, "scan.siz")).read().split()
    assert [int(t) for t in siz] == [shape[3], shape[2], shape[0] * shape[1], 0]
    return np.fromfile(os.path.join(folder, "scan.bin"), np.float32).reshape(shape)


@pytest.mark.parametrize("clockwise,storage,volume", [(False, "f32", "auto"),
                                                       (True, "f32", "streamed"),
                                                       (True, "f16", "auto")])
def test_file_driver_matches_one_pass(tmp_path, clockwise, storage, volume):
    folder = str(tmp_path)
    cra_path = write_scan(folder, clockwise)
    s = default_setting()
    # 40 slices of 6: a streamed volume keeps 2 * 32 + 6 = 70 >= 40 slices, i.e.
    # all of them; the wrapping ring is covered by tests/test_gpu_stream.py
    s.slices_per_block, s.x_direction, s.vertical_center = 6, 1, 1.5
    say(f"driver clockwise={clockwise} storage={storage} volume={volume}: 2 scan blocks x 3 "
        "angle blocks")
    st = reconstruct_file(cra_path, s, storage=storage, angle_block_bytes=25 * DX * DZ * 4,
                          volume=volume)
    assert (st.scan_blocks, st.angle_blocks) == (SCANS, 3)          # 25 + 25 + 10
    assert st.bytes_read == SCANS * N * DX * DZ * 4
    got = read_volume(folder, (SCANS, 40, 64, 64))
    want = one_pass(folder, s, storage)
    scale = np.abs(want).max()
    err = np.abs(got - want).max() / scale
    say(f"  max |file - one pass| / max = {err:.2e}; {st.seconds_total:.3f} s")
    assert scale > 0 and err < (1e-5 if storage == "f32" else 1e-4)
    # the two scan blocks differ (other centre, other data) and sit in order
    assert np.abs(got[0] - got[1]).max() > 0.05 * scale


def test_command_line_program(tmp_path):
    folder = str(tmp_path)
    cra_path = write_scan(folder, False)
    s = default_setting()
    s.slices_per_block, s.wait_for_enter = 8, 1
    SaveSetting(s, os.path.join(folder, "setting.xml"))
    say("tomokvas-fbp with 1 MiB angle blocks (42 + 18 projections)")
    r = subprocess.run([EXE, "--angle-block-mib", "1", cra_path], capture_output=True,
                       text=True, cwd=folder, stdin=subprocess.DEVNULL, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "split to 2 block(s)" in r.stdout and "total time" in r.stdout
    got = read_volume(folder, (SCANS, 40, 64, 64))
    found, s2 = LoadSetting(os.path.join(folder, "setting.xml"))
    assert found and s2.slices_per_block == 8
    want = one_pass(folder, s2, "f32")
    err = np.abs(got - want).max() / np.abs(want).max()
    say(f"  max |file - one pass| / max = {err:.2e}")
    assert err < 1e-5
