"""The slab pipeline of fdk_calculate_reconstruction: the volume backprojected
in chunks of whole slabs while the previous chunk's slabs go to the host on a
second stream; resident (whole volume in HBM) and streamed (two chunks in HBM)
volumes.

Every voxel is computed by the same kernel code from the same inputs whether
the volume is backprojected whole or in chunks, so the slabs handed out must
equal the one-launch volume bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tomokvas_amd import _lib
from tomokvas_amd.engine import VolumeSink

from gpu_util import engine, ogeom, say, scan

pytestmark = pytest.mark.gpu

_hip = None


def hip_runtime():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


def d2h(ptr, shape):
    hip_runtime()
    out = np.empty(shape, np.float32)
    assert _hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(out.nbytes),
                          2) == 0
    return out


def one_launch(g, s, storage):
    """No sink: the whole volume backprojected (in angle chunks of >= 100
    angles, at most 16), read from HBM."""
    fbp = engine(g, s, None, storage=storage)
    _lib.check(fbp.Initialize(), "Initialize")
    _lib.check(fbp.CalculateReconstruction(), "CalculateReconstruction")
    vol = d2h(fbp.device_volume_ptr(), (g.Nz, g.Ny, g.Nx))
    fbp.close()
    return vol


def piped(g, s, storage, mode, sink=None, collecting=False):
    sink = sink or VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
    fbp = engine(g, s, sink, storage=storage, collecting=collecting)
    _lib.check(fbp.SetVolumeMode(mode), "SetVolumeMode")
    _lib.check(fbp.Initialize(), "Initialize")
    chunk = fbp.GetVolumeChunk()
    _lib.check(fbp.CalculateReconstruction(), "CalculateReconstruction")
    times = fbp.stage_times_ms()
    streamed_ptr = fbp.device_volume_ptr()
    fbp.close()
    return sink, chunk, times, streamed_ptr


# (Nz, B): one or two chunks with a ragged last slab; chunk ends that are not
# slab ends; a streamed volume larger than its circular buffer (68 slices for
# B = 4, 70 for B = 6: the last chunks wrap and, for B = 6, slab [66, 72)
# straddles the buffer's end)
CASES = [(33, 10), (40, 6), (80, 4), (80, 6)]


@pytest.mark.parametrize("storage", [_lib.F32, _lib.F16])
@pytest.mark.parametrize("Nz,B", CASES)
def test_chunked_slabs_equal_one_launch(Nz, B, storage):
    g = scan(48, 40, 30, 48, 44, Nz, B=B)
    rng = np.random.default_rng(Nz * 100 + B)
    s = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    want = one_launch(g, s, storage)
    for mode in (_lib.VOLUME_RESIDENT, _lib.VOLUME_STREAMED):
        say(f"pipeline Nz={Nz} B={B} storage={storage} mode={mode}")
        sink, chunk, times, ptr = piped(g, s, storage, mode)
        nslab = -(-Nz // B)
        assert chunk >= 8
        assert [c[0] for c in sink.calls] == list(range(nslab))
        assert not any(c[1] for c in sink.calls)
        assert (ptr == 0) == (mode == _lib.VOLUME_STREAMED)
        np.testing.assert_array_equal(sink.volume, want)
        say(f"  chunks of <= {chunk} slices; BP {times[2]:.3f} ms, "
            f"exposed transfers {times[3]:.3f} ms")


@pytest.mark.parametrize("storage", [_lib.F32, _lib.F16])
def test_handout_angle_split_equals_one_pass(storage):
    # 700 projections: 7 angle chunks of 100, the first over the whole volume,
    # the last 6 through the slab pipeline in z-chunks -- the same launches
    # per voxel, in the same order, as the pass without a sink
    g = scan(32, 24, 700, 32, 32, 70, B=8)
    s = np.random.default_rng(7).standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    want = one_launch(g, s, storage)
    sink, chunk, times, _ = piped(g, s, storage, _lib.VOLUME_RESIDENT)
    say(f"  hand-out over 7 angle chunks: z-chunks of <= {chunk} slices, BP {times[2]:.3f} ms")
    assert [c[0] for c in sink.calls] == list(range(-(-g.Nz // g.B)))
    np.testing.assert_array_equal(sink.volume, want)


def test_streamed_collecting_two_angle_blocks():
    # two angle blocks accumulated through the sink with a streamed volume that
    # wraps its circular buffer, against the resident one-pass volume
    N = 40
    g = scan(48, 40, N, 48, 48, 80, B=4)
    s = O.project_ellipsoids(ogeom(g), O.shepp_logan(24.0)).astype(np.float32)
    full = one_launch(g, s, _lib.F32)
    sink = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
    for blk in range(2):
        gb = scan(48, 40, N // 2, 48, 48, 80, B=4, rot=blk * (N // 2) * g.dbeta)
        gb.dbeta = g.dbeta
        part = np.ascontiguousarray(s[:, blk * N // 2:(blk + 1) * N // 2, :])
        piped(gb, part, _lib.F32, _lib.VOLUME_STREAMED, sink=sink, collecting=bool(blk))
    assert sink.calls[-1] == (19, True)
    e = O.rel_rmse(sink.volume, full)
    say(f"  streamed two-block accumulation vs one pass rel RMSE {e:.3e}")
    assert e <= 1e-5


def test_streamed_volume_needs_a_sink():
    g = scan(32, 32, 8, 32, 32, 16, B=8)
    fbp = engine(g, np.zeros((32, 8, 32), np.float32), None)
    _lib.check(fbp.SetVolumeMode(_lib.VOLUME_STREAMED))
    _lib.check(fbp.Initialize())
    assert fbp.CalculateReconstruction() == _lib.FDK_ERR_INVALID_ARGUMENT
    assert fbp.SetVolumeMode(7) == _lib.FDK_ERR_INVALID_ARGUMENT
    fbp.close()


@pytest.mark.parametrize("storage", [_lib.F32, _lib.F16])
@pytest.mark.parametrize("with_sink", [False, True])
def test_reconstruct_host_overlaps_upload(storage, with_sink):
    # 360 angles: three chunks of 120, each uploaded while the previous one is
    # filtered and backprojected; the chunks' partial volumes are summed, as
    # the one-pass run without a sink also does
    g = scan(64, 48, 360, 48, 44, 40, B=8)
    rng = np.random.default_rng(9)
    s = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    want = one_launch(g, s, storage)            # three angle chunks as well
    sink = VolumeSink(g.Nx, g.Ny, g.Nz, g.B) if with_sink else None
    fbp = engine(g, None, sink, storage=storage)
    say(f"reconstruct_host storage={storage} sink={with_sink}")
    _lib.check(fbp.ReconstructHost(s), "ReconstructHost")
    got = sink.volume if with_sink else d2h(fbp.device_volume_ptr(), (g.Nz, g.Ny, g.Nx))
    times = fbp.stage_times_ms()
    fbp.close()
    e = O.rel_rmse(got, want)
    say(f"  rel RMSE vs one pass {e:.2e}; upload..BP {times[2]:.3f} ms, exposed {times[3]:.3f} ms")
    assert e <= 1e-6
    # a streamed volume takes one chunk (one launch over all angles)
    if with_sink:
        sink2 = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
        fbp = engine(g, None, sink2, storage=storage)
        _lib.check(fbp.SetVolumeMode(_lib.VOLUME_STREAMED))
        _lib.check(fbp.ReconstructHost(s), "ReconstructHost streamed")
        fbp.close()
        assert O.rel_rmse(sink2.volume, want) <= 1e-6


def test_reconstruct_host_collecting_over_angle_blocks():
    # two angle blocks of 250 projections (two chunks each), the second
    # collecting through the sink, against one pass over all 500
    N = 500
    g = scan(64, 48, N, 48, 48, 24, B=8)
    rng = np.random.default_rng(13)
    s = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    want = one_launch(g, s, _lib.F32)
    sink = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
    fbp = None
    for blk in range(2):
        gb = scan(64, 48, N // 2, 48, 48, 24, B=8, rot=blk * (N // 2) * g.dbeta)
        gb.dbeta = g.dbeta
        part = np.ascontiguousarray(s[:, blk * N // 2:(blk + 1) * N // 2, :])
        fbp = engine(gb, None, sink, collecting=bool(blk))
        say(f"reconstruct_host collecting block {blk}")
        _lib.check(fbp.ReconstructHost(part), "ReconstructHost")
        fbp.close()
    e = O.rel_rmse(sink.volume, want)
    say(f"  two collected blocks vs one pass rel RMSE {e:.2e}")
    assert e <= 1e-5


def test_buffers_kept_across_initialisations():
    # one context re-initialised for a larger, then a smaller, then the first
    # geometry (buffers grow once and are kept): each result equals a fresh
    # context's
    geoms = [scan(48, 40, 120, 40, 40, 24, B=8), scan(64, 56, 240, 56, 52, 40, B=8),
             scan(48, 40, 120, 40, 40, 24, B=8)]
    rng = np.random.default_rng(17)
    data = [rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32) for g in geoms]
    fresh = [one_launch(g, s, _lib.F16) for g, s in zip(geoms, data)]
    sink0 = VolumeSink(geoms[0].Nx, geoms[0].Ny, geoms[0].Nz, geoms[0].B)
    fbp = engine(geoms[0], data[0], sink0, storage=_lib.F16)
    for i, (g, s) in enumerate(zip(geoms, data)):
        sink = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
        g.apply(fbp, s, sink)
        say(f"re-initialised context, geometry {i}")
        _lib.check(fbp.Initialize(), "Initialize")
        _lib.check(fbp.CalculateReconstruction(), "CalculateReconstruction")
        np.testing.assert_array_equal(sink.volume, fresh[i])
    fbp.close()


@pytest.mark.parametrize("storage", [_lib.F32, _lib.F16])
def test_streamed_chunks_cut_their_angles(storage):
    # 240 angles: every chunk of a streamed volume (80 slices in a ring of 70,
    # slab 6) is backprojected in two angle launches, as the resident volume
    # is -- the same sums in the same order, so equal bit for bit
    g = scan(48, 40, 240, 48, 44, 80, B=6)
    rng = np.random.default_rng(23)
    s = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    want = one_launch(g, s, storage)
    say(f"streamed 240 angles storage={storage}")
    sink, chunk, times, ptr = piped(g, s, storage, _lib.VOLUME_STREAMED)
    assert ptr == 0
    np.testing.assert_array_equal(sink.volume, want)


def test_setters_after_initialize_require_it_again():
    # a size-changing setter after Initialize must not let CalculateReconstruction
    # run with buffers sized for the old geometry
    g = scan(48, 40, 30, 40, 40, 24, B=8)
    s = np.random.default_rng(29).standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    sink = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
    fbp = engine(g, s, sink)
    _lib.check(fbp.Initialize(), "Initialize")
    _lib.check(fbp.CalculateReconstruction(), "CalculateReconstruction")
    big = scan(48, 40, 30, 64, 64, 48, B=8)
    sink2 = VolumeSink(big.Nx, big.Ny, big.Nz, big.B)
    _lib.check(fbp.SetReconstructionProperties(sink2, big.Nx, big.cx, big.Ny, big.cy, big.Nz,
                                               big.cz, big.rot))
    assert fbp.CalculateReconstruction() == _lib.FDK_ERR_NOT_INITIALIZED
    _lib.check(fbp.SetCollecting(True))
    assert fbp.CalculateReconstruction() == _lib.FDK_ERR_NOT_INITIALIZED
    _lib.check(fbp.SetCollecting(False))
    _lib.check(fbp.SetVolumeMode(_lib.VOLUME_STREAMED))
    assert fbp.CalculateReconstruction() == _lib.FDK_ERR_NOT_INITIALIZED
    _lib.check(fbp.SetVolumeMode(_lib.VOLUME_RESIDENT))
    # initialising again keeps the projections already on the device
    _lib.check(fbp.Initialize(), "Initialize again")
    _lib.check(fbp.CalculateReconstruction(), "CalculateReconstruction again")
    fbp.close()
    want = one_launch(big, s, _lib.F32)
    np.testing.assert_array_equal(sink2.volume, want)


def test_calculate_waits_for_the_null_stream():
    """The context's streams are non-blocking: work the caller queued on the
    NULL stream before fdk_calculate_reconstruction -- here the noise kernel
    over the 9.4e8 samples of the 512^3 workload's projections, left running
    asynchronously -- is finished before the filter reads the stack, so the
    volume equals the one computed with the device synchronised in between."""
    from tomokvas_amd import geometry
    from tomokvas_amd.engine import FilteredBackProjection
    w = geometry.workloads()["cfg2"]
    g = w.geom
    L = _lib.load()
    say("512^3 from 900 x 1024^2: noise left running on the NULL stream, then the FDK")
    vols = []
    for sync in (True, False):
        fbp = FilteredBackProjection(0)
        g.apply(fbp, None, None)
        _lib.check(fbp.SetFilter(w.filter), "SetFilter")
        _lib.check(fbp.Initialize(), "Initialize")
        ells = np.ascontiguousarray(w.ellipsoids())
        _lib.check(L.fdk_phantom_project(fbp.handle, ells.ctypes.data, ells.shape[0], 0, g.Nang,
                                         fbp.device_sinogram_ptr(), None), "phantom")
        _lib.check(L.fdk_poisson_log(fbp.device_sinogram_ptr(), g.Nu * g.Nv * g.Nang, w.I0,
                                     w.seed, None), "poisson")
        if sync:
            assert hip_runtime().hipDeviceSynchronize() == 0
        _lib.check(L.fdk_calculate_reconstruction(fbp.handle, _lib.GET_BLOCK_FN(), None),
                   "calculate")
        vols.append(d2h(fbp.device_volume_ptr(), (g.Nz, g.Ny, g.Nx)))
        fbp.close()
    assert np.abs(vols[0]).max() > 0
    assert np.array_equal(vols[0], vols[1])
