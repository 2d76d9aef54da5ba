"""The reconstruction program (include/tomokvas/fdk_driver.h, tomokvas-fbp):
settings file and command line, CPU only.  The engine runs are in
tests/test_gpu_driver.py.

The settings file below is the reference's own FBPG/setting.xml (a data file);
the defaults are those of FBPG/para_TC.cpp:287-296.
"""
import os
import subprocess

import pytest

from tomokvas_amd import _lib
from tomokvas_amd.build import EXE
from tomokvas_amd.driver import LoadSetting, SaveSetting, default_setting, reconstruct_file

REFERENCE_SETTING = """<GPU_FBP>
    <GPU ID="0" MemLimit_MiB="1848" AllocLimit_MiB="1024" />
    <RECONSTRUCTION SimultaneouslyReconstructedSlices="30" X_direction="0" Y_direction="0" Z_direction="0" DetectorVerticalCenterOffset="0" />
    <GENERAL WaitForEnter="1" />
</GPU_FBP>
"""

FIELDS = ("gpu_id", "mem_limit_mib", "alloc_limit_mib", "slices_per_block", "x_direction",
          "y_direction", "z_direction", "vertical_center", "wait_for_enter")


def as_tuple(s):
    return tuple(getattr(s, f) for f in FIELDS)


def test_defaults():
    assert as_tuple(default_setting()) == (0, 5 * 1024, 1024, 16, 0, 0, 0, 0.0, 1)


def test_reference_setting_file(tmp_path):
    p = tmp_path / "setting.xml"
    p.write_text(REFERENCE_SETTING)
    found, s = LoadSetting(str(p))
    assert found
    assert as_tuple(s) == (0, 1848, 1024, 30, 0, 0, 0, 0.0, 1)


def test_partial_and_missing(tmp_path):
    p = tmp_path / "setting.xml"
    # absent <GPU>: its defaults; absent attributes keep theirs
    p.write_text('<GPU_FBP>\n<RECONSTRUCTION X_direction="1" Z_direction=\'2\' '
                 'DetectorVerticalCenterOffset="-3.25"/>\n<GENERAL WaitForEnter="0"></GENERAL>'
                 "\n</GPU_FBP>")
    found, s = LoadSetting(str(p))
    assert found
    assert as_tuple(s) == (0, 5 * 1024, 1024, 16, 1, 0, 1, -3.25, 0)
    p.write_text("<SOMETHING_ELSE/>")
    found, s = LoadSetting(str(p))
    assert not found and as_tuple(s) == as_tuple(default_setting())
    found, s = LoadSetting(str(tmp_path / "absent.xml"))
    assert not found and as_tuple(s) == as_tuple(default_setting())


def test_save_round_trip(tmp_path):
    s = default_setting()
    s.gpu_id, s.slices_per_block, s.y_direction, s.vertical_center = 3, 24, 1, 0.1
    p = str(tmp_path / "s.xml")
    SaveSetting(s, p)
    found, t = LoadSetting(p)
    assert found and as_tuple(t) == as_tuple(s)
    # the layout of the reference's file: element per line, attributes in its order
    lines = open(p).read().splitlines()
    assert lines[0] == "<GPU_FBP>" and lines[-1] == "</GPU_FBP>"
    assert lines[1].split()[:2] == ["<GPU", 'ID="3"']
    assert lines[2].split()[1] == 'SimultaneouslyReconstructedSlices="24"'


def test_command_line_usage_and_errors(tmp_path):
    assert os.path.exists(EXE), "build first: python -m tomokvas_amd.build"
    r = subprocess.run([EXE], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 255 and "usage:" in r.stderr
    r = subprocess.run([EXE, "--storage", "f8", "x.CRA"], capture_output=True, text=True,
                       cwd=tmp_path)
    assert r.returncode == 2
    # a missing .CRA fails before any GPU call; the absent setting.xml is
    # written with the defaults, as the reference does
    r = subprocess.run([EXE, "--quiet", str(tmp_path / "none.CRA")], capture_output=True,
                       text=True, cwd=tmp_path)
    assert r.returncode == 1 and "none.CRA" in r.stderr
    found, s = LoadSetting(str(tmp_path / "setting.xml"))
    assert found and as_tuple(s) == as_tuple(default_setting())
    # every option is accepted (the run then fails on the missing file, not on usage)
    for opts in (["--streamed"], ["--resident"], ["--storage", "bf16", "--shepp-logan"],
                 ["--angle-block-mib", "64", "--setting", str(tmp_path / "s2.xml")]):
        r = subprocess.run([EXE, "--quiet", *opts, str(tmp_path / "none.CRA")],
                           capture_output=True, text=True, cwd=tmp_path)
        assert r.returncode == 1 and "none.CRA" in r.stderr, opts
    assert os.path.exists(tmp_path / "s2.xml")     # --setting names the file written
    r = subprocess.run([EXE, "--bogus", "x.CRA"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 255 and "usage:" in r.stderr


def test_driver_rejects_bad_setting(tmp_path):
    s = default_setting()
    s.slices_per_block = 0
    with pytest.raises(_lib.FDKError):
        reconstruct_file(str(tmp_path / "x.CRA"), s)
