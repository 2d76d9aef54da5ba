"""The RSQ preprocessing on the GPU (fdk_rsq_process) against the oracle
(oracle/rsq.py) on the same raw file.

The kernel evaluates the oracle's FP64 expressions in the same order without
FMA contraction; only the device's log and atan may differ from the host's in
the last bit, so FP64 outputs agree to 1e-12 of the largest value and FP32
outputs are the FP64 values rounded (within 1 ulp)."""
import math

import numpy as np
import pytest

from oracle import rsq as R
from rsq_util import synthetic_raw, write_rsq
from tomokvas_amd import _lib
from tomokvas_amd._hip import DeviceBuffer
from tomokvas_amd.rsq import Rsq

from gpu_util import say

pytestmark = pytest.mark.gpu

SX, SY, N, D = 96, 360, 4, 300.0


@pytest.fixture(scope="module")
def rsq_file(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("rsq") / "scan.rsq")
    raw = synthetic_raw(N, SY, SX, seed=11)
    # A = 1.1, B = 0.1: line integrals between about -3.2 and 0.05
    write_rsq(path, raw, bh=(0, 1, 2, 0.1, 1.2, 2.3), dimz=N, nr_of_projections=2)
    return path, raw


MODES = [  # complete, minimal, flip, invert, axis_move
    (False, False, False, False, 5),
    (False, False, True, True, 0),
    (True, False, False, False, 7),
    (True, False, True, True, 7),
    (True, True, False, True, 3),
    (True, True, True, False, 3),
    # axis on the last column: one output column, no redundant half
    (True, False, True, False, SX - 1),
    (False, False, False, True, SX - 1),
]


@pytest.mark.parametrize("complete,minimal,flip,invert,axis", MODES)
def test_gpu_preprocessing_matches_oracle(rsq_file, complete, minimal, flip, invert, axis):
    path, raw = rsq_file
    r = Rsq(path)
    r.p.sin_x, r.p.sin_y = SX, SY
    r.SetAxisMove(axis)
    r.SetDistance(D)
    r.SetDeltaFi(2 * math.pi / SY)
    r.SetCompleting(complete)
    r.SetMinimal(minimal)
    r.SetFlip(flip)
    r.SetInvert(invert)
    assert r.LoadImages(0, N - 1) == 0
    say(f"rsq complete={complete} minimal={minimal} flip={flip} invert={invert} axis={axis}")
    got = r.ProccessAndCopy()
    h = r.header
    want = R.process(raw, h.bh_a, h.bh_b, axis, D, 2 * math.pi / SY, flip, invert, complete,
                     minimal)
    assert got.shape == want.shape
    scale = np.abs(want).max()
    err = np.abs(got - want).max()
    say(f"  {got.shape}: max |gpu - oracle| = {err:.2e} (scale {scale:.3f})")
    assert err <= 1e-12 * scale
    # FP32 straight into a device buffer (the engine's stack layout)
    dev = DeviceBuffer(got.size * 4)
    r.process_to_device(dev.ptr.value, _lib.F32)
    f32 = np.empty(got.shape, np.float32)
    dev.download(f32)
    dev.close()
    ulp = np.spacing(np.abs(want).astype(np.float32))
    assert (np.abs(f32.astype(np.float64) - want) <= ulp).all()


def test_gpu_groups_do_not_change_the_result(tmp_path):
    """The full-fan rebinning takes its line integrals in groups of sinograms
    (64 MiB of FP64 scratch: 4 sinograms of 2000 x 1017 here, so 9 sinograms
    run in groups 4 + 4 + 1).  Every sinogram equals, bit for bit, the same
    sinogram processed alone, and the oracle to the tolerance above."""
    n, sy, sx, axis, dist = 9, 2000, 1024, 7, 125810 / 24
    raw = synthetic_raw(n, sy, sx, seed=5)
    path = str(tmp_path / "groups.rsq")
    write_rsq(path, raw, bh=(0, 1, 2, 0.1, 1.2, 2.3), dimz=n, nr_of_projections=2)
    for flip, minimal in ((False, False), (True, True)):
        r = Rsq(path)
        r.p.sin_x, r.p.sin_y = sx, sy
        r.SetAxisMove(axis)
        r.SetDistance(dist)
        r.SetDeltaFi(2 * math.pi / sy)
        r.SetCompleting(True)
        r.SetMinimal(minimal)
        r.SetFlip(flip)
        r.SetInvert(True)
        say(f"rsq groups: {n} sinograms of {sy}x{sx}, flip={flip} minimal={minimal}")
        assert r.LoadImages(0, n - 1) == 0
        whole = r.ProccessAndCopy()
        for k in (0, 3, 4, 8):
            assert r.LoadImages(k, k) == 0
            alone = r.ProccessAndCopy()
            assert np.array_equal(alone[0], whole[k]), k
        h = r.header
        want = R.process(raw[8:9], h.bh_a, h.bh_b, axis, dist, 2 * math.pi / sy, flip, True,
                         True, minimal)
        err = np.abs(whole[8:9] - want).max()
        say(f"  {whole.shape}: max |gpu - oracle| = {err:.2e}")
        assert err <= 1e-12 * np.abs(want).max()
        r.close()
