"""GPU parity of the HIP FDK path against the CPU oracle (through the C ABI).

Bars: FP32 path vs the FP32 mirror <= 1e-5 relative RMSE (the FP32 parity bar);
stage-level FP32 kernels much tighter; FP64 instantiations vs the FP64 oracle
<= 1e-12; FP16/BF16 storage <= 1e-3 vs the FP64 oracle.
"""
import math

import numpy as np
import pytest

import oracle as O
from tomokvas_amd import _lib, geometry, phantoms
from tomokvas_amd.engine import VolumeSink

from gpu_util import engine, ogeom, run_engine, say, scan, stages, ulp_diff

pytestmark = pytest.mark.gpu

TOL_FDK = 1e-5      # FP32 GPU vs FP32 CPU mirror (relative RMSE)
TOL_F16 = 1e-3      # FP16/BF16 storage vs FP64 oracle (relative RMSE)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


@pytest.fixture(scope="module")
def cfg1():
    w = geometry.workloads()["cfg1"]
    g = w.geom
    sino = O.project_ellipsoids(ogeom(g), O.shepp_logan(g.fov_px / 2))
    return g, sino


# ------------------------------------------------------------------ engine, cfg1
def test_engine_fdk_cfg1_matches_mirror(cfg1):
    g, sino = cfg1
    say("engine FDK cfg1 (128^3, 180 x 128^2)")
    vol, times, sink = run_engine(g, sino.astype(np.float32))
    ref32 = O.fdk(ogeom(g), sino.astype(np.float32), kind=0)
    ref64 = O.fdk(ogeom(g), sino, kind=0, f64=True)
    e32 = O.rel_rmse(vol, ref32)
    e64 = O.rel_rmse(vol, ref64)
    say(f"  rel RMSE vs FP32 mirror {e32:.3e}, vs FP64 oracle {e64:.3e}, stage ms {times}")
    assert e32 <= TOL_FDK
    assert e64 <= TOL_FDK
    assert sink.calls == [(n, False) for n in range(O.slab_count(g.Nz, g.B))]


@pytest.mark.parametrize("storage,name", [(_lib.F16, "fp16"), (_lib.BF16, "bf16"),
                                          (_lib.F16_PACKED, "fp16-packed")])
def test_engine_low_precision_storage_cfg1(cfg1, storage, name):
    g, sino = cfg1
    vol, _, _ = run_engine(g, sino.astype(np.float32), storage=storage)
    ref64 = O.fdk(ogeom(g), sino, kind=0, f64=True)
    e = O.rel_rmse(vol, ref64)
    say(f"  {name} storage: rel RMSE vs FP64 oracle {e:.3e}")
    assert e <= TOL_F16


# ------------------------------------------------------------------ odd geometry
ODD = dict(Nu=37, Nv=23, Nang=7, Nx=19, Ny=23, Nz=17, D=60.0, tau=0.1, scale=1.3, B=16,
           cX=17.2, cZ=11.7, rot=0.3, flips=(1, 0, 1))


def test_engine_odd_geometry_and_dropped_slab():
    # (Nz-1) % B == 0: the reference would skip slice 16; every slice is produced
    g = scan(**ODD)
    rng = np.random.default_rng(1)
    sino = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    vol, _, sink = run_engine(g, sino)
    ref = O.fdk(ogeom(g), sino, kind=0)
    e = O.rel_rmse(vol, ref)
    say(f"  odd geometry rel RMSE {e:.3e}; slabs {sink.calls}")
    assert e <= TOL_FDK
    assert [c[0] for c in sink.calls] == [0, 1]
    assert O.slab_count_reference(g.Nz, g.B) == 1
    assert np.abs(vol[16]).max() > 0
    assert O.rel_rmse(vol[16], ref[16]) <= TOL_FDK


def test_engine_wide_detector():
    # 5000 detector columns: p = 16384, one row transform per workgroup of 1024
    # threads in 128 KiB of LDS
    g = scan(5000, 8, 6, 24, 24, 1, D=6000.0, scale=150.0)
    assert g.p == 16384
    sino = np.random.default_rng(5).standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    vol, _, _ = run_engine(g, sino)
    ref = O.fdk(ogeom(g), sino, kind=0)
    assert np.abs(ref).max() > 0
    e = O.rel_rmse(vol, ref)
    say(f"  Nu = 5000: FP32 vs mirror {e:.2e}")
    assert e <= TOL_FDK


# degenerate shapes: one projection, one detector row, one slice, one voxel,
# a one-column detector (p = 2)
DEGENERATE = {
    "one_projection": dict(Nu=33, Nv=20, Nang=1, Nx=24, Ny=24, Nz=12, D=200.0, tau=0.1, B=16),
    "one_row": dict(Nu=40, Nv=1, Nang=9, Nx=16, Ny=16, Nz=1, D=400.0, tau=0.1, B=16),
    "one_slice": dict(Nu=40, Nv=12, Nang=9, Nx=21, Ny=17, Nz=1, D=400.0, tau=0.1, B=16),
    "one_voxel": dict(Nu=5, Nv=5, Nang=3, Nx=1, Ny=1, Nz=1, D=100.0, tau=0.1, B=1),
    "one_column": dict(Nu=1, Nv=8, Nang=4, Nx=3, Ny=3, Nz=4, D=100.0, tau=0.1, scale=0.2, B=4),
}


@pytest.mark.parametrize("shape", list(DEGENERATE))
def test_engine_degenerate_shapes(shape):
    g = scan(**DEGENERATE[shape])
    sino = np.random.default_rng(3).standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    ref32 = O.fdk(ogeom(g), sino, kind=0)
    ref64 = O.fdk(ogeom(g), sino, kind=0, f64=True)
    assert np.abs(ref32).max() > 0
    vol, _, sink = run_engine(g, sino)
    e = O.rel_rmse(vol, ref32)
    # with a handful of projections nothing averages the storage's rounding of
    # each sample, so the bar is twice its unit roundoff (FP16 2^-11, BF16 2^-9)
    # rather than the many-projection volume bar
    errs, bars = {}, {"f16": 2 * 2.0 ** -11, "bf16": 2 * 2.0 ** -9}
    for st, name in ((_lib.F16, "f16"), (_lib.BF16, "bf16")):
        v, _, _ = run_engine(g, sino, storage=st)
        errs[name] = O.rel_rmse(v, ref64)
    say(f"  {shape}: FP32 vs mirror {e:.2e}, vs FP64 " +
        ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert e <= TOL_FDK
    assert all(errs[k] <= bars[k] for k in errs)
    # every slab is handed out once, in order
    assert [c[0] for c in sink.calls] == list(range(-(-g.Nz // g.B)))


def test_storage_is_scale_invariant_and_zero_safe():
    # FP16 stores value/scale per projection, BF16 and FP32 the values: inputs
    # scaled by 1e-30 or 1e30 reconstruct to the scaled volume with the same
    # relative error as unscaled ones (FP16's range alone would overflow or
    # flush them); zero projections give an exactly zero volume (no 0/0 scale)
    g = scan(**ODD)
    sino = np.random.default_rng(11).standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    ref64 = O.fdk(ogeom(g), sino.astype(np.float64), kind=0, f64=True)
    for st, name in ((_lib.F32, "f32"), (_lib.F16, "f16"), (_lib.BF16, "bf16")):
        base, _, _ = run_engine(g, sino, storage=st)
        e0 = O.rel_rmse(base, ref64)
        for k in (1e-30, 1e30):
            v, _, _ = run_engine(g, (sino * np.float32(k)).astype(np.float32), storage=st)
            assert np.all(np.isfinite(v)), (name, k)
            e = O.rel_rmse(v, ref64 * k)
            say(f"  {name} x{k:g}: rel RMSE {e:.2e} (unscaled {e0:.2e})")
            assert e <= 2 * e0 + 1e-6, (name, k, e, e0)
        z, _, _ = run_engine(g, np.zeros_like(sino), storage=st)
        assert not np.any(z), name


def test_engine_block_size_invariance(cfg1):
    g, sino = cfg1
    g = scan(64, 64, 30, 64, 64, 40, B=16)
    s = O.project_ellipsoids(ogeom(g), O.shepp_logan(32.0)).astype(np.float32)
    vols = []
    for B in (1, 7, 16, 39, 64):
        g.B = B
        v, _, _ = run_engine(g, s)
        vols.append(v)
    # the slab recentring rounds the float slice coordinates slightly
    # differently for each B (kernel.cu:398-400, :821); nothing else depends on B
    for v in vols[1:]:
        assert O.rel_rmse(v, vols[0]) <= 1e-6


def test_engine_collecting_two_angle_blocks():
    # two "sin blocks" (para_TC.cpp:117-155): angles [0,N/2) then [N/2,N) with the
    # rotation offset, the second accumulated through the slab callback
    N = 40
    g = scan(48, 40, N, 48, 48, 33, B=10)
    s = O.project_ellipsoids(ogeom(g), O.shepp_logan(24.0)).astype(np.float32)
    full, _, _ = run_engine(g, s)
    sink = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
    for blk in range(2):
        gb = scan(48, 40, N // 2, 48, 48, 33, B=10, rot=blk * (N // 2) * g.dbeta)
        gb.dbeta = g.dbeta
        part = np.ascontiguousarray(s[:, blk * N // 2:(blk + 1) * N // 2, :])
        fbp = engine(gb, part, sink, collecting=bool(blk))
        _lib.check(fbp.Initialize(), "Initialize")
        _lib.check(fbp.CalculateReconstruction(), "CalculateReconstruction")
        fbp.close()
    assert sink.calls[-1] == (3, True)
    # the FFT pads to p of the block's own detector, identical here; only the
    # float rounding of the angle offsets differs
    e = O.rel_rmse(sink.volume, full)
    say(f"  two-block accumulation vs one pass rel RMSE {e:.3e}")
    assert e <= 1e-5


def test_engine_errors_are_reported():
    g = scan(16, 16, 4, 16, 16, 16, D=4.0)        # D < Nu/2: invalid
    fbp = engine(g, np.zeros((16, 4, 16), np.float32))
    rc = fbp.Initialize()
    assert rc == _lib.FDK_ERR_INVALID_ARGUMENT
    assert "source distance" in _lib.last_error()
    fbp.close()
    fbp = engine(scan(16, 16, 4, 16, 16, 16))
    assert fbp.CalculateReconstruction() == _lib.FDK_ERR_NOT_INITIALIZED
    fbp.close()


# ------------------------------------------------------------------ stage level
def test_weight_kernel_matches_mirror(torch):
    g = scan(**ODD)
    st = stages(g)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    d = torch.from_numpy(x).cuda()
    st.weight(d)
    ref = O.weight(ogeom(g), x)
    u = ulp_diff(d.cpu().numpy(), ref)
    say(f"  weight max ulp difference {u.max()}")
    assert u.max() <= 1
    d64 = torch.from_numpy(x.astype(np.float64)).cuda()
    st.weight(d64)
    assert np.allclose(d64.cpu().numpy(), O.weight(ogeom(g), x.astype(np.float64), f64=True),
                       rtol=1e-15, atol=0)


@pytest.mark.parametrize("Nu", [1, 2, 3, 5, 9, 17, 37, 64, 100, 200, 300, 600, 1024, 1536, 2048,
                                3000, 4096, 4097, 6000, 8192])
def test_filter_kernel_all_lengths(torch, Nu):
    g = scan(Nu, 3, 5, 8, 8, 8, D=max(2500.0, Nu))
    for filt in (0, 1):
        st = stages(g, filt=filt)
        rng = np.random.default_rng(Nu)
        x = rng.standard_normal((g.Nv, g.Nang, g.Nu))
        Hf = O.filter_spectrum(filt, Nu, g.tau)
        ref = O.filter_rows(ogeom(g), x, Hf, f64=True)
        d = torch.from_numpy(x.astype(np.float32)).cuda()
        maxabs = torch.zeros(g.Nang, dtype=torch.int32, device="cuda")
        st.filter(d, maxabs=maxabs)
        y = d.cpu().numpy()
        e = O.rel_rmse(y, ref)
        assert e <= 2e-6, (Nu, filt, e)
        m = maxabs.cpu().numpy().view(np.float32)
        assert np.allclose(m, np.abs(y).max(axis=(0, 2)), rtol=0, atol=0)
        if g.p <= 8192:   # FP64 rows up to 4096 samples (p = 8192: 128 KiB of LDS)
            d64 = torch.from_numpy(x).cuda()
            st.filter(d64)
            assert O.rel_rmse(d64.cpu().numpy(), ref) <= 1e-13, (Nu, filt)


def test_filter_fused_weight_and_user_spectrum(torch):
    g = scan(37, 6, 9, 8, 8, 8)
    rng = np.random.default_rng(4)
    h = rng.standard_normal(g.p)
    st = stages(g, taps=h)
    x = rng.standard_normal((g.Nv, g.Nang, g.Nu))
    re, im = O.spectrum_complex(h)
    w = O.weight(ogeom(g), x, f64=True)
    ref = O.filter_rows_complex(ogeom(g), w, re, im, f64=True)
    d = torch.from_numpy(x).cuda()
    st.filter(d, weight=True)
    assert O.rel_rmse(d.cpu().numpy(), ref) <= 1e-13
    # adjoint flag multiplies by conj(H): <F a, b> = <a, F^T b>
    a = torch.from_numpy(rng.standard_normal((g.Nv, g.Nang, g.Nu))).cuda()
    b = torch.from_numpy(rng.standard_normal((g.Nv, g.Nang, g.Nu))).cuda()
    Fa = st.filter(a.clone())
    Ftb = st.filter(b.clone(), adjoint=True)
    lhs = float((Fa * b).sum())
    rhs = float((a * Ftb).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("Nu", [37, 1536])
def test_filter_fp32_fused_weight_complex_spectrum(torch, Nu):
    # FP32 rows (radix-8 passes over 512 threads at p = 4096): fused cosine
    # weight, a complex user spectrum, and its conjugate for the adjoint
    g = scan(Nu, 6, 9, 8, 8, 8, D=max(2500.0, Nu))
    rng = np.random.default_rng(40 + Nu)
    h = rng.standard_normal(g.p)
    st = stages(g, taps=h)
    x = rng.standard_normal((g.Nv, g.Nang, g.Nu))
    re, im = O.spectrum_complex(h)
    ref = O.filter_rows_complex(ogeom(g), O.weight(ogeom(g), x, f64=True), re, im, f64=True)
    d = torch.from_numpy(x.astype(np.float32)).cuda()
    st.filter(d, weight=True)
    e = O.rel_rmse(d.cpu().numpy(), ref)
    say(f"  FP32 fused filter Nu={Nu}: rel RMSE vs FP64 {e:.2e}")
    assert e <= 2e-6
    a = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    b = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    Fa = st.filter(torch.from_numpy(a).cuda()).cpu().numpy().astype(np.float64)
    Ftb = st.filter(torch.from_numpy(b).cuda(), adjoint=True).cpu().numpy().astype(np.float64)
    lhs, rhs = float((Fa * b).sum()), float((a * Ftb).sum())
    assert abs(lhs - rhs) <= 1e-5 * (np.abs(Fa * b).sum())


# "wide": voxels of 4 detector samples, so many block footprints exceed the
# backprojector's LDS tile and are gathered from HBM (and some miss the detector)
WIDE = dict(Nu=96, Nv=96, Nang=20, Nx=40, Ny=40, Nz=40, D=300.0, tau=0.1, scale=4.0)


@pytest.mark.parametrize("geo", ["odd", "cfg1", "wide"])
def test_backproject_kernel_matches_mirror(torch, geo):
    g = {"odd": lambda: scan(**ODD), "wide": lambda: scan(**WIDE),
         "cfg1": lambda: geometry.workloads()["cfg1"].geom}[geo]()
    st = stages(g)
    rng = np.random.default_rng(5)
    q = rng.standard_normal((g.Nv, g.Nang, g.Nu))
    ref32 = O.backproject(ogeom(g), q.astype(np.float32))
    ref64 = O.backproject(ogeom(g), q, f64=True)
    d = torch.from_numpy(q.astype(np.float32)).cuda()
    stack, scales = st.pack(d, _lib.F32)
    vol = st.backproject(stack, _lib.F32, scales).cpu().numpy()
    e32, e64 = O.rel_rmse(vol, ref32), O.rel_rmse(vol, ref64)
    say(f"  BP {geo}: vs mirror {e32:.3e}, vs FP64 {e64:.3e}")
    # the mirror samples at bit-identical positions; vs FP64 the float position
    # rounding (~1e-4 px at u ~ 100) dominates
    assert e32 <= 1e-6 and e64 <= 1e-5
    d64 = torch.from_numpy(q).cuda()
    c64, s64 = st.pack(d64, _lib.F64)
    v64 = st.backproject(c64, _lib.F64, s64).cpu().numpy()
    assert O.rel_rmse(v64, ref64) <= 1e-12
    # accumulate flag and angle sub-ranges: two halves add up to the whole
    h = g.Nang // 2
    c1, s1 = st.pack(torch.from_numpy(np.ascontiguousarray(q[:, :h]).astype(np.float32)).cuda(), _lib.F32)
    c2, s2 = st.pack(torch.from_numpy(np.ascontiguousarray(q[:, h:]).astype(np.float32)).cuda(), _lib.F32)
    v = st.backproject(c1, _lib.F32, s1, ang0=0, nang=h)
    st.backproject(c2, _lib.F32, s2, ang0=h, nang=g.Nang - h, vol=v, accumulate=True)
    # same samples and weights, the sum split in two: FP32 summation order only
    assert O.rel_rmse(v.cpu().numpy(), vol) <= 1e-6


@pytest.mark.parametrize("geo", ["odd", "wide"])
def test_fp16_scale_and_bf16_storage(torch, geo):
    g = scan(**(ODD if geo == "odd" else WIDE))
    st = stages(g)
    rng = np.random.default_rng(6)
    q = (rng.standard_normal((g.Nv, g.Nang, g.Nu)) * 1e4).astype(np.float32)
    d = torch.from_numpy(q).cuda()
    maxabs = torch.from_numpy(np.abs(q).max(axis=(0, 2)).astype(np.float32).view(np.int32)).cuda()
    ref = O.backproject(ogeom(g), q.astype(np.float64), f64=True)
    for dt, tol in ((_lib.F16, 1e-3), (_lib.BF16, 4e-3), (_lib.F16_PACKED, 1e-3)):
        stack, scales = st.pack(d, dt, maxabs)
        if dt in (_lib.F16, _lib.F16_PACKED):
            s = scales.cpu().numpy()
            assert np.allclose(s, np.abs(q).max(axis=(0, 2)) / 8192.0, rtol=1e-7)
        v = st.backproject(stack, dt, scales).cpu().numpy()
        e = O.rel_rmse(v, ref)
        say(f"  storage {dt}: rel RMSE {e:.3e}")
        assert e <= tol


@pytest.mark.parametrize("geo", ["odd", "cfg1", "wide"])
def test_low_precision_storage_interpolates_in_fp32(torch, geo):
    # FP16 and BF16 storage: the backprojector widens the stored samples to
    # FP32 and interpolates, weights and sums in FP32, so its volume is the
    # FP32 mirror's backprojection of the decoded stored samples up to FP32
    # rounding (the FP16 scale multiplies the weight instead of each sample);
    # the packed-FP16 kernel, with FP16 fractions and weights, is not
    g = {"odd": lambda: scan(**ODD), "wide": lambda: scan(**WIDE),
         "cfg1": lambda: geometry.workloads()["cfg1"].geom}[geo]()
    st = stages(g)
    rng = np.random.default_rng(9)
    q = (rng.standard_normal((g.Nv, g.Nang, g.Nu)) * 100).astype(np.float32)
    d = torch.from_numpy(q).cuda()
    m = np.abs(q).max(axis=(0, 2)).astype(np.float32)
    maxabs = torch.from_numpy(m.view(np.int32)).cuda()
    pitch = (g.Nu + 68 + 63) // 64 * 64
    n = g.Nang * g.Nv * pitch
    for dt, view in ((_lib.F16, np.float16), (_lib.BF16, np.uint16)):
        stack, scales = st.pack(d, dt, maxabs)
        raw = stack.cpu().numpy().view(view)[:n].reshape(g.Nang, g.Nv, pitch)[:, :, :g.Nu]
        if dt == _lib.F16:
            dec = raw.astype(np.float32) * scales.cpu().numpy()[:, None, None]
        else:
            dec = (raw.astype(np.uint32) << 16).view(np.float32)
        ref = O.backproject(ogeom(g), np.ascontiguousarray(dec.transpose(1, 0, 2)))
        v = st.backproject(stack, dt, scales).cpu().numpy()
        e = O.rel_rmse(v, ref)
        say(f"  {geo} storage {dt}: vs FP32 mirror of the stored samples {e:.3e}")
        assert e <= 1e-6
        if dt == _lib.F16:
            vp = st.backproject(stack, _lib.F16_PACKED, scales).cpu().numpy()
            ep = O.rel_rmse(vp, ref)
            say(f"  {geo} packed FP16 interpolation: {ep:.3e}")
            assert 1e-5 < ep <= 1e-3


def test_weight_and_pack_on_rows_not_16_byte_aligned(torch):
    # Nu = 96 rows take 16-byte accesses only when the array starts 16-byte
    # aligned; a view one float into a buffer takes the scalar path and gives
    # the same bits; a stored stack that is not 16-byte aligned is refused
    g = scan(**WIDE)
    st = stages(g)
    rng = np.random.default_rng(9)
    x = rng.standard_normal((g.Nv, g.Nang, g.Nu)).astype(np.float32)
    aligned = torch.from_numpy(x).cuda()
    buf = torch.empty(x.size + 1, dtype=torch.float32, device="cuda")
    off = buf[1:].view(x.shape)
    off.copy_(aligned)
    assert off.data_ptr() % 16 == 4
    st.weight(aligned)
    st.weight(off)
    assert torch.equal(aligned, off)
    m = torch.from_numpy(np.abs(x).max(axis=(0, 2)).astype(np.float32).view(np.int32)).cuda()
    s1, c1 = st.pack(aligned, _lib.F16, m)
    s2, c2 = st.pack(off, _lib.F16, m)
    n = st.stack_bytes(_lib.F16, g.Nang) - st.stack_bytes(_lib.F16, 0)   # the packed rows
    assert torch.equal(s1[:n], s2[:n]) and torch.equal(c1, c2)
    nb = st.stack_bytes(_lib.F16)
    bad = torch.empty(nb + 16, dtype=torch.uint8, device="cuda")[4:4 + nb]
    with pytest.raises(RuntimeError, match="16-byte aligned"):
        st.pack_into(aligned, _lib.F16, bad, c1.clone(), 0, m)


@pytest.mark.parametrize("geo", ["odd", "wide"])
def test_pack_writes_the_stored_stack_exactly(torch, geo):
    # [beta][v][pitch] rows: the samples (FP16: q * (1/scale) rounded once to
    # nearest even -- the kernel's fused multiply-convert rounds the exact
    # product, so the reference forms it exactly in FP64; BF16: rounded to
    # nearest even), zeros past column Nu - 1; Nu = 37 takes the scalar loads,
    # Nu = 96 the 16-byte ones
    g = scan(**(ODD if geo == "odd" else WIDE))
    st = stages(g)
    rng = np.random.default_rng(8)
    q = (rng.standard_normal((g.Nv, g.Nang, g.Nu)) * 100).astype(np.float32)
    d = torch.from_numpy(q).cuda()
    m = np.abs(q).max(axis=(0, 2)).astype(np.float32)
    maxabs = torch.from_numpy(m.view(np.int32)).cuda()
    qt = q.transpose(1, 0, 2)                                   # [beta][v][u]
    for dt in (_lib.F32, _lib.F16, _lib.BF16):
        stack, scales = st.pack(d, dt, maxabs)
        raw = stack.cpu().numpy()
        pitch = (g.Nu + 68 + 63) // 64 * 64
        n = g.Nang * g.Nv * pitch
        if dt == _lib.F32:
            got = raw.view(np.float32)[:n].reshape(g.Nang, g.Nv, pitch)
            want = qt
        elif dt == _lib.F16:
            sc = np.where(m > 0, m / np.float32(8192.0), np.float32(1)).astype(np.float32)
            assert np.array_equal(scales.cpu().numpy(), sc)
            inv = (np.float32(1) / sc).astype(np.float32)
            want = (qt.astype(np.float64) * inv.astype(np.float64)[:, None, None]).astype(np.float16)
            got = raw.view(np.float16)[:n].reshape(g.Nang, g.Nv, pitch)
        else:
            bits = qt.view(np.uint32).astype(np.uint64)
            rne = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16).astype(np.uint16)
            want = rne
            got = raw.view(np.uint16)[:n].reshape(g.Nang, g.Nv, pitch)
        assert np.array_equal(got[:, :, :g.Nu], want), dt
        assert not got[:, :, g.Nu:].any(), dt


def test_phantom_projector_matches_oracle(torch, cfg1):
    g, sino = cfg1
    st = stages(g)
    d = st.phantom(phantoms.shepp_logan(g.fov_px / 2))
    assert np.allclose(d.cpu().numpy(), sino.astype(np.float32), rtol=2e-6, atol=1e-6)
    g2 = scan(40, 30, 12, 16, 16, 16, D=500.0, tau=0.05)
    e = phantoms.random_ellipsoids(3, 64, 40.0)
    st2 = stages(g2)
    d2 = st2.phantom(e).cpu().numpy()
    ref = O.project_ellipsoids(ogeom(g2), O.random_ellipsoids(3, 64, 40.0))
    assert np.allclose(d2, ref, rtol=2e-6, atol=1e-6)


def test_poisson_noise_matches_oracle(torch):
    rng = np.random.default_rng(8)
    p = rng.uniform(0.0, 6.0, 200000).astype(np.float32)
    d = torch.from_numpy(p).cuda()
    from tomokvas_amd import _lib as L
    L.check(L.load().fdk_poisson_log(d.data_ptr(), d.numel(), 1e5, 1234, None))
    torch.cuda.synchronize()
    ref = O.poisson_log(p, 1e5, 1234)
    same = np.mean(d.cpu().numpy() == ref)
    say(f"  identical noisy samples: {same:.6f}")
    assert same >= 0.999


def test_poisson_noise_covers_more_than_2_32_samples(torch):
    # a 1800 x 2048^2 stack holds 7.5e9 samples; one dispatch covers at most
    # 2^32 - 1 threads, so the noise kernel must loop: samples past index 2^32
    # get their own draws (p = 0: -ln(n/I0) is exactly zero only when n = I0)
    from tomokvas_amd import _lib as L
    n = (1 << 32) + (1 << 21)
    d = torch.zeros(n, dtype=torch.float32, device="cuda")
    L.check(L.load().fdk_poisson_log(d.data_ptr(), n, 1e5, 99, None))
    torch.cuda.synchronize()
    tail = d[1 << 32:]
    zero_tail = float((tail == 0).float().mean())
    head_nz = float((d[:1 << 21] != 0).float().mean())
    rng = np.random.default_rng(12)
    idx = np.concatenate([rng.integers(0, 1 << 32, 200),
                          rng.integers(1 << 32, n, 400)]).astype(np.int64)
    got = d[torch.from_numpy(idx).cuda()].cpu().numpy()
    del d, tail
    ref = np.array([np.float32(-np.log(max(O.poisson_draw(1e5, 99, int(i)), 1.0) / 1e5))
                    for i in idx])
    same = np.mean(got == ref)
    say(f"  zero samples past 2^32: {zero_tail:.4%}; noisy in the first 2^21: {head_nz:.4%}; "
        f"sampled draws identical to the oracle's: {same:.4f}")
    assert zero_tail < 0.01 and head_nz > 0.99
    assert same >= 0.99 and np.abs(got - ref).max() <= 1e-6


# ------------------------------------------------------------------ full size (cfg2)
def test_cfg2_full_size_properties(torch):
    """512^3 from 900 x 1024^2, 64 ellipsoids + Poisson, Shepp-Logan filter."""
    w = geometry.workloads()["cfg2"]
    g = w.geom
    say("cfg2 full size: generate, filter, backproject")
    st = stages(g, filt=w.filter)
    sino = st.phantom(w.ellipsoids())
    st.poisson_log(sino, w.I0, w.seed)
    raw = sino.clone()
    maxabs = torch.zeros(g.Nang, dtype=torch.int32, device="cuda")
    st.filter(sino, weight=True, maxabs=maxabs)
    stack, scales = st.pack(sino, _lib.F32)
    v32 = st.backproject(stack, _lib.F32, scales)
    c16, s16 = st.pack(sino, _lib.F16, maxabs)
    v16 = st.backproject(c16, _lib.F16, s16)
    torch.cuda.synchronize()
    e16 = O.rel_rmse(v16.cpu().numpy(), v32.cpu().numpy())
    say(f"  FP16 vs FP32 rel RMSE {e16:.3e}")
    assert e16 <= TOL_F16
    # filter: 256 random rows against the FP64 oracle (rows are independent)
    rng = np.random.default_rng(10)
    rows = rng.choice(g.Nv * g.Nang, 256, replace=False)
    flat_in = raw.view(-1, g.Nu)[torch.from_numpy(rows).cuda()].cpu().numpy().astype(np.float64)
    flat_out = sino.view(-1, g.Nu)[torch.from_numpy(rows).cuda()].cpu().numpy()
    vs = rows // g.Nang
    gr = ogeom(g)
    Hf = O.filter_spectrum(w.filter, g.Nu, g.tau)
    ref_rows = np.zeros_like(flat_in)
    for i, v in enumerate(vs):
        one = O.Geometry(**{**g.oracle_dict(), "Nv": 1, "Nang": 1, "cZ": g.cZ - v})
        wrow = O.weight(one, flat_in[i].reshape(1, 1, -1), f64=True)
        ref_rows[i] = O.filter_rows(one, wrow, Hf, f64=True).ravel()
    ef = O.rel_rmse(flat_out, ref_rows)
    # the FP32 mirror on the same rows: the ramp filter removes the large DC of
    # log-transformed counts, so FP32 input rounding is amplified for both
    mir = np.stack([O.filter_rows(O.Geometry(**{**g.oracle_dict(), "Nv": 1, "Nang": 1,
                                                 "cZ": g.cZ - v}),
                                  O.weight(O.Geometry(**{**g.oracle_dict(), "Nv": 1, "Nang": 1,
                                                         "cZ": g.cZ - v}),
                                           flat_in[i].astype(np.float32).reshape(1, 1, -1)),
                                  Hf).ravel() for i, v in enumerate(vs)])
    em = O.rel_rmse(mir, ref_rows)
    say(f"  filter rows vs FP64 oracle rel RMSE {ef:.3e} (FP32 mirror {em:.3e})")
    assert ef <= max(2e-6, 2 * em)
    # backprojection: 2000 random voxels against the FP64 oracle on the same input
    host_f = sino.cpu().numpy().astype(np.float64)
    pts = np.stack([rng.integers(0, g.Nx, 2000), rng.integers(0, g.Ny, 2000),
                    rng.integers(0, g.Nz, 2000)], axis=1).astype(np.int32)
    ref_pts = O.backproject_points(gr, host_f, pts, f64=True)
    got = v32.cpu().numpy()[pts[:, 2], pts[:, 1], pts[:, 0]]
    eb = O.rel_rmse(got, ref_pts)
    say(f"  BP voxel sample vs FP64 oracle rel RMSE {eb:.3e}")
    assert eb <= 1e-5


# ------------------------------------------------------------------ committed fixture
def test_engine_matches_golden_fixture(torch):
    import os
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz"))
    keys = ["Nu", "Nang", "Nv", "cX", "cZ", "dbeta", "D", "tau", "Nx", "Ny", "Nz", "cx", "cy",
            "cz", "rot", "scale", "flipX", "flipY", "flipZ", "B"]
    d = dict(zip(keys, fx["fdk_geom"]))
    g = scan(int(d["Nu"]), int(d["Nv"]), int(d["Nang"]), int(d["Nx"]), int(d["Ny"]), int(d["Nz"]),
             D=d["D"], tau=d["tau"], scale=d["scale"], B=int(d["B"]))
    vol, _, _ = run_engine(g, fx["fdk_sino"])
    e = O.rel_rmse(vol, fx["fdk_vol64"])
    say(f"  golden 32^3 FDK: FP32 engine vs FP64 fixture {e:.3e}")
    assert e <= TOL_FDK
    # FP64 stage path reproduces the fixture to FP64 rounding
    st = stages(g)
    s = torch.from_numpy(fx["fdk_sino"].astype(np.float64)).cuda()
    st.filter(s, weight=True)
    stack, scales = st.pack(s, _lib.F64)
    v64 = st.backproject(stack, _lib.F64, scales).cpu().numpy()
    assert O.rel_rmse(v64, fx["fdk_vol64"]) <= 1e-12


# ------------------------------------------------------------------ full size (cfg3)
def test_cfg3_precision_form(torch):
    """1024^3 from 1800 x 2048^2 (64 seeded ellipsoids + Poisson, Shepp-Logan
    filter) through fdk_calculate_reconstruction with FP32, FP16 and BF16
    storage; FP16/BF16 volumes against the FP32 volume.  Three 64^3 blocks are
    checked against the CPU FP32 mirror and the FP64 oracle run on each block's
    detector rows only (weight and filter act on rows independently):

      - a seeded block near the isocentre;
      - a block of the top slab inside the phantom (72 % of it material),
        whose upper slices project past the last detector row for about half
        the angles, so the backprojector's range-tested gathers carry signal;
      - a top-slab block straddling the edge of the reconstruction circle
        (38 % of it masked to zero), which in this phantom is air.

    The first two meet the bars (FP32 GPU vs mirror 1e-5, 16-bit storage vs
    FP64 1e-3 relative RMSE).  The air block's own norm is ~1e-3 of theirs, so
    relative errors there mean little: it is held to exact zeros where masked,
    to the FP32 mirror's own distance from FP64, and to absolute RMS errors no
    larger than the central block's."""
    import ctypes as C
    from tomokvas_amd import _hip
    from tomokvas_amd.engine import FilteredBackProjection
    w = geometry.workloads()["cfg3"]
    g = w.geom
    L = _lib.load()
    say("cfg3 full size: 1024^3 from 1800 x 2048^2, FP32 / FP16 / BF16 storage")
    fbp = FilteredBackProjection(0)
    g.apply(fbp, None, None)
    _lib.check(fbp.SetFilter(w.filter), "SetFilter")
    _lib.check(fbp.SetStorage(_lib.F32), "SetStorage")
    _lib.check(fbp.Initialize(), "Initialize")
    ells = np.ascontiguousarray(w.ellipsoids())
    _lib.check(L.fdk_phantom_project(fbp.handle, ells.ctypes.data, ells.shape[0], 0, g.Nang,
                                     fbp.device_sinogram_ptr(), None), "phantom")
    _lib.check(L.fdk_poisson_log(fbp.device_sinogram_ptr(), g.Nu * g.Nv * g.Nang, w.I0, w.seed,
                                 None), "poisson")
    hip = _hip.hip()
    n = 64
    # (x0, y0, z0): the seeded central block; a top-slab block inside the
    # phantom; a top-slab block at radius ~0.99 of the reconstruction circle
    # (62 % of its columns inside)
    origins = [geometry.sub_cube(g, w.seed, n)[:3], (448, 512, g.Nz - n),
               (848, 827, g.Nz - n)]
    cases = []
    for x0, y0, z0 in origins:
        lo, hi = geometry.detector_rows_of_block(g, x0, y0, z0, n)
        raw = np.empty((hi - lo, g.Nang, g.Nu), np.float32)
        assert hip.hipMemcpy(raw.ctypes.data, C.c_void_p(fbp.device_sinogram_ptr()
                                                         + lo * g.Nang * g.Nu * 4),
                             raw.nbytes, 2) == 0
        cases.append(dict(origin=(x0, y0, z0), rows=(lo, hi), raw=raw, blocks={}))
    nvol = g.Nx * g.Ny * g.Nz
    ref = torch.empty(nvol, dtype=torch.float32, device="cuda")
    for st, name in ((_lib.F32, "f32"), (_lib.F16, "f16"), (_lib.BF16, "bf16"),
                     (_lib.F16_PACKED, "f16-packed")):
        _lib.check(fbp.SetStorage(st), "SetStorage")
        _lib.check(L.fdk_calculate_reconstruction(fbp.handle, _lib.GET_BLOCK_FN(), None), name)
        t = fbp.stage_times_ms()
        dvol = fbp.device_volume_ptr()
        for c in cases:
            x0, y0, z0 = c["origin"]
            blk = np.empty((n, g.Ny, g.Nx), np.float32)
            assert hip.hipMemcpy(blk.ctypes.data, C.c_void_p(dvol + z0 * g.Nx * g.Ny * 4),
                                 blk.nbytes, 2) == 0
            c["blocks"][name] = blk[:, y0:y0 + n, x0:x0 + n].reshape(-1)
        if st == _lib.F32:
            assert hip.hipMemcpy(C.c_void_p(ref.data_ptr()), C.c_void_p(dvol), nvol * 4, 3) == 0
            say(f"  {name}: filter {t[0]:.1f} ms, pack {t[1]:.1f} ms, BP {t[2]:.1f} ms")
        else:
            out3 = (C.c_double * 3)()
            _lib.check(L.fdk_compare(C.c_void_p(dvol), C.c_void_p(ref.data_ptr()), nvol, out3,
                                     None), "compare")
            e = math.sqrt(out3[0] / out3[1])
            say(f"  {name}: filter {t[0]:.1f} ms, pack {t[1]:.1f} ms, BP {t[2]:.1f} ms; "
                f"full volume vs FP32 rel RMSE {e:.3e}")
            assert e <= TOL_F16, name
    fbp.close()
    del ref
    Hf = O.filter_spectrum(w.filter, g.Nu, g.tau)
    for i, c in enumerate(cases):
        # the block on the CPU: FP32 mirror and FP64 oracle from the same raw rows
        (x0, y0, z0), (lo, hi), blocks = c["origin"], c["rows"], c["blocks"]
        og = O.Geometry(**g.rows_oracle_dict(lo, hi))
        zz, yy, xx = np.meshgrid(np.arange(z0, z0 + n), np.arange(y0, y0 + n),
                                 np.arange(x0, x0 + n), indexing="ij")
        pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1).astype(np.int32)
        raw = c.pop("raw")
        # row positions on the full detector (row0): the FP32 mirror samples where
        # the GPU's FP32 path does
        m32 = O.backproject_points(og, O.filter_rows(og, O.weight(og, raw), Hf), pts, row0=lo)
        raw64 = raw.astype(np.float64)
        del raw
        o64 = O.backproject_points(og, O.filter_rows(og, O.weight(og, raw64, f64=True), Hf,
                                                     f64=True), pts, f64=True, row0=lo)
        del raw64
        zero = o64 == 0
        e_mirror = O.rel_rmse(m32, o64)
        e = {k: O.rel_rmse(v, o64) for k, v in blocks.items()}
        e_f32_mirror = O.rel_rmse(blocks["f32"], m32)
        # absolute RMS errors
        arms = lambda a, b: math.sqrt(np.mean((a.astype(np.float64) - b) ** 2))  # noqa: E731
        c["abs"] = {k: arms(v, o64) for k, v in blocks.items()}
        c["abs"]["f32 vs mirror"] = arms(blocks["f32"], m32.astype(np.float64))
        say(f"  block x{x0} y{y0} z{z0} n{n} (rows [{lo},{hi}), {zero.mean():.0%} masked, "
            f"RMS {math.sqrt(np.mean(o64 ** 2)):.3e}): "
            f"FP32 vs mirror {e_f32_mirror:.3e}; vs FP64 oracle: mirror {e_mirror:.3e}, " +
            ", ".join(f"{k} {v:.3e}" for k, v in e.items()))
        say("    absolute RMS errors: " + ", ".join(f"{k} {v:.3e}" for k, v in c["abs"].items()))
        # masked voxels are exactly zero in every storage type
        for k, v in blocks.items():
            assert not v[zero].any(), k
        # FP32 rounding of the log-transformed input, amplified by the ramp
        # filter, sets both FP32 paths' distance to FP64 (the mirror's included)
        assert e["f32"] <= max(TOL_FDK, 2 * e_mirror)
        if i < 2:
            assert e_f32_mirror <= TOL_FDK
            assert e["f16"] <= TOL_F16 and e["bf16"] <= TOL_F16 and e["f16-packed"] <= TOL_F16
        else:
            for k, v in c["abs"].items():
                assert v <= cases[0]["abs"][k], k
    # the air block is partly outside the circle
    assert 0 < (cases[2]["blocks"]["f32"] == 0).mean() < 1


def test_set_gpu_id_moves_every_stream(cfg1):
    # fdk_set_gpu_id makes every stream and event again on the (new) device;
    # the slab pipeline (copy stream) and fdk_reconstruct_host (copy and prep
    # streams) after it must give the results of a context that never moved
    g, sino = cfg1
    s32 = sino.astype(np.float32)
    out = []
    for move in (False, True):
        fbp = engine(g, None, None)
        if move:
            _lib.check(fbp.SetGPUid(0), "SetGPUid")
        sink = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
        g.apply(fbp, s32, sink)
        _lib.check(fbp.Initialize(), "Initialize")
        _lib.check(fbp.CalculateReconstruction(), "CalculateReconstruction")
        a = sink.volume.copy()
        sink2 = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
        g.apply(fbp, None, sink2)
        _lib.check(fbp.ReconstructHost(s32), "ReconstructHost")
        out.append((a, sink2.volume.copy()))
        fbp.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert O.rel_rmse(out[1][1], out[1][0]) <= 1e-5
