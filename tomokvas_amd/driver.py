"""The reference's reconstruction program (FBPG/para_TC.cpp) over the C ABI of
include/tomokvas/fdk_driver.h (host code).

``LoadSetting`` / ``SaveSetting`` read and write the settings file
(FBPG/para_TC.cpp:192-302, FBPG/setting.xml); ``reconstruct_file`` runs main()
(FBPG/para_TC.cpp:5-175) on one .CRA file.  The same program ships as the
executable ``tomokvas_amd/tomokvas-fbp``:

    tomokvas-fbp [--storage f32|f16|bf16] [--angle-block-mib N] scan/name.CRA
"""
from __future__ import annotations

import ctypes as C

from . import _lib

STORAGE = {"f32": _lib.F32, "f16": _lib.F16, "bf16": _lib.BF16}


def default_setting() -> "_lib.Setting":
    s = _lib.Setting()
    _lib.check(_lib.load().fdk_setting_defaults(C.byref(s)), "fdk_setting_defaults")
    return s


def LoadSetting(path: str = "setting.xml"):
    """(found, setting): found is False when the file has no <GPU_FBP> root, and
    the setting then holds the defaults (the reference returns false)."""
    s = _lib.Setting()
    rc = _lib.load().fdk_setting_load(path.encode(), C.byref(s))
    return rc == _lib.FDK_OK, s


def SaveSetting(setting, path: str = "setting.xml") -> None:
    _lib.check(_lib.load().fdk_setting_save(path.encode(), C.byref(setting)), "fdk_setting_save")


def reconstruct_file(cra_path: str, setting=None, storage: str = "f32",
                     shepp_logan: bool = False, angle_block_bytes: int = 0,
                     verbose: bool = False, volume: str = "auto") -> "_lib.RunStats":
    """Reconstruct <dir>/<name>.CRA into <dir>/<name>.bin and .siz; returns the
    run's statistics (blocks, seconds, bytes).  volume: "auto" (streamed when
    the volume exceeds 45 % of the free HBM), "resident" or "streamed"."""
    if setting is None:
        setting = default_setting()
    opt = _lib.RunOptions(STORAGE[storage],
                          _lib.FILTER_SHEPPLOGAN if shepp_logan else _lib.FILTER_RAMLAK,
                          int(angle_block_bytes), int(bool(verbose)),
                          {"auto": -1, "resident": _lib.VOLUME_RESIDENT,
                           "streamed": _lib.VOLUME_STREAMED}[volume])
    st = _lib.RunStats()
    _lib.check(_lib.load().fdk_reconstruct_file(cra_path.encode(), C.byref(setting),
                                                 C.byref(opt), C.byref(st)),
               "fdk_reconstruct_file")
    return st
