"""The L2-resident read probe that prices the backprojector's gather roofline."""
import ctypes as C

import pytest

from tomokvas_amd import _lib

from gpu_util import say

pytestmark = pytest.mark.gpu


def test_l2_probe_reports_on_chip_bandwidth():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    lib = _lib.load()
    res = {}
    for fp in (2 << 20, 64 << 20):
        gbs = C.c_double(0.0)
        _lib.check(lib.fdk_probe_l2_read(fp, C.byref(gbs)), "probe")
        res[fp] = gbs.value
        say(f"  read bandwidth, {fp >> 20} MiB footprint: {gbs.value:.0f} GB/s")
    # resident in L2: well above the 8 TB/s of HBM, below any plausible on-chip peak
    assert 8000 < res[2 << 20] < 200000
    # 64 MiB exceeds every XCD's 4 MiB L2: the small footprint must be faster
    assert res[2 << 20] > res[64 << 20]
