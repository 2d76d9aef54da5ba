"""GPU tests of the differentiable form: adjoint identities, gradcheck, and the
gradient against the CPU oracle's transcription (through the C ABI)."""
import numpy as np
import pytest

import oracle as O
from tomokvas_amd import _lib, geometry, phantoms

from gpu_util import ogeom, say, scan, stages

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def small():
    # gradcheck size: 32^3 volume, 32 x 32 detector, 24 projections (p = 64)
    return scan(32, 32, 24, 32, 32, 32, D=100.0, tau=0.2)


@pytest.mark.parametrize("geo", ["small", "odd"])
def test_dot_product_adjoint_fp64(torch, geo):
    g = small() if geo == "small" else scan(37, 23, 7, 19, 23, 17, D=60.0, tau=0.1, scale=1.3,
                                            cX=17.2, cZ=11.7, rot=0.3, flips=(1, 0, 1))
    st = stages(g)
    gen = torch.Generator(device="cuda").manual_seed(1)
    y = torch.randn(st.sino_shape(), dtype=torch.float64, device="cuda", generator=gen)
    x = torch.randn(st.vol_shape, dtype=torch.float64, device="cuda", generator=gen)
    stack, scales = st.pack(y, _lib.F64)
    BPy = st.backproject(stack, _lib.F64, scales)
    FPx = st.forward_project(x)
    lhs = float((BPy * x).sum())
    rhs = float((y * FPx).sum())
    rel = abs(lhs - rhs) / abs(lhs)
    say(f"  <BP y, x> = {lhs:.15e}, <y, BP^T x> = {rhs:.15e}, rel {rel:.2e}")
    assert rel <= 1e-6
    # against the CPU transcription of the transposed splat
    ref = O.forward_project(ogeom(g), x.cpu().numpy(), f64=True)
    assert O.rel_rmse(FPx.cpu().numpy(), ref) <= 1e-12


# geometries of the FP32 adjoint checks; "wide" has voxels of 4 detector
# samples, so many blocks splat straight into HBM instead of through LDS
ADJ_GEOMS = {
    "small": lambda: small(),
    "odd": lambda: scan(37, 23, 7, 19, 23, 17, D=60.0, tau=0.1, scale=1.3, cX=17.2, cZ=11.7,
                        rot=0.3, flips=(1, 0, 1)),
    "wide": lambda: scan(96, 96, 20, 40, 40, 40, D=300.0, tau=0.1, scale=4.0),
}


@pytest.mark.parametrize("geo", list(ADJ_GEOMS))
def test_forward_project_fp32_matches_fp64(torch, geo):
    g = ADJ_GEOMS[geo]()
    st = stages(g)
    gen = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn(st.vol_shape, dtype=torch.float64, device="cuda", generator=gen)
    y = torch.randn(st.sino_shape(), dtype=torch.float64, device="cuda", generator=gen)
    f64 = st.forward_project(x).cpu().numpy()
    f32 = st.forward_project(x.float()).double().cpu().numpy()
    # the CPU FP32 mirror samples at bit-identical positions; only the order of
    # the atomic sums differs
    ref32 = O.forward_project(ogeom(g), x.float().cpu().numpy())
    e32 = O.rel_rmse(f32, ref32)
    e64 = O.rel_rmse(f32, f64)
    # FP32 adjoint vs FP32 backprojector: <BP y, x> = <y, BP^T x>
    stack, scales = st.pack(y.float(), _lib.F32)
    bpy = st.backproject(stack, _lib.F32, scales).double()
    lhs = float((bpy * x).sum())
    rhs = float((y * torch.from_numpy(f32).cuda()).sum())
    rel = abs(lhs - rhs) / abs(lhs)
    say(f"  FP32 adjoint {geo}: vs FP32 mirror {e32:.2e}, vs FP64 {e64:.2e}, "
        f"dot-product rel {rel:.2e}")
    assert e32 <= 1e-5 and rel <= 1e-5
    if geo != "small":
        # "small" has exact 45-degree angles: in FP32 cos == sin, so U == 1 and
        # the outermost slices land exactly on the detector's first and last
        # rows, which FP64 rounding puts just outside (a full-sample jump with
        # random data); elsewhere FP32 and FP64 differ by position rounding
        assert e64 <= 1e-5


def test_full_operator_adjoint_fp64(torch):
    from fdk_autograd import fdk, ramlak_taps
    g = small()
    st = stages(g)
    taps = torch.from_numpy(ramlak_taps(st)).cuda()
    gen = torch.Generator(device="cuda").manual_seed(2)
    p = torch.randn(st.sino_shape(), dtype=torch.float64, device="cuda", generator=gen,
                    requires_grad=True)
    x = torch.randn(st.vol_shape, dtype=torch.float64, device="cuda", generator=gen)
    Ap = fdk(p, taps, st)
    (ATx,) = torch.autograd.grad(Ap, p, grad_outputs=x)
    lhs = float((Ap.detach() * x).sum())
    rhs = float((p.detach() * ATx).sum())
    rel = abs(lhs - rhs) / abs(lhs)
    say(f"  <A p, x> vs <p, A^T x>: rel {rel:.2e}")
    assert rel <= 1e-6


def test_ramlak_taps_reproduce_reference_filter(torch):
    # learnable taps initialised to Ram-Lak give the reference's FDK exactly
    from fdk_autograd import fdk, ramlak_taps
    g = small()
    st = stages(g)
    sino = O.project_ellipsoids(ogeom(g), O.shepp_logan(16.0))
    taps = torch.from_numpy(ramlak_taps(st)).cuda()
    vol = fdk(torch.from_numpy(sino).cuda(), taps, st).cpu().numpy()
    ref = O.fdk(ogeom(g), sino, kind=0, f64=True)
    assert O.rel_rmse(vol, ref) <= 1e-12


def test_gradcheck_fp64(torch):
    from fdk_autograd import fdk, ramlak_taps
    g = small()
    st = stages(g)
    gen = torch.Generator(device="cuda").manual_seed(3)
    p = torch.randn(st.sino_shape(), dtype=torch.float64, device="cuda", generator=gen,
                    requires_grad=True)
    taps = torch.from_numpy(ramlak_taps(st)).cuda().requires_grad_(True)
    # the adjoint's detector splats are atomics: the FP64 summation order varies
    # between runs, hence a small nondeterminism tolerance
    ok = torch.autograd.gradcheck(lambda a, b: fdk(a, b, st), (p, taps), eps=1e-6, atol=1e-5,
                                  rtol=1e-4, fast_mode=True, nondet_tol=1e-9)
    assert ok


def test_gradient_matches_cpu_transcription(torch):
    from fdk_autograd import fdk, ramlak_taps
    g = small()
    st = stages(g)
    og = ogeom(g)
    ells = O.random_ellipsoids(7, 16, 32.0)
    sino = O.project_ellipsoids(og, ells)
    x_true = O.voxelize_ellipsoids(og, ells)
    h0 = ramlak_taps(st)
    rng = np.random.default_rng(4)
    h = h0 * (1 + 0.05 * rng.standard_normal(h0.size))
    p = torch.from_numpy(sino).cuda().requires_grad_(True)
    taps = torch.from_numpy(h).cuda().requires_grad_(True)
    vol = fdk(p, taps, st)
    loss = ((vol - torch.from_numpy(x_true).cuda()) ** 2).sum()
    loss.backward()
    # CPU transcription of the same chain rule
    re, im = O.spectrum_complex(h)
    q = O.weight(og, sino, f64=True)
    y = O.filter_rows_complex(og, q, re, im, f64=True)
    x = O.backproject(og, y, f64=True)
    gx = 2 * (x - x_true)
    gy = O.forward_project(og, gx, f64=True)
    gp = O.weight(og, O.filter_rows_complex(og, gy, re, -im, f64=True), f64=True)
    gh = O.tap_gradient(og, q, gy)
    ep = O.rel_rmse(p.grad.cpu().numpy(), gp)
    eh = O.rel_rmse(taps.grad.cpu().numpy(), gh)
    say(f"  dL/dp rel {ep:.2e}, dL/dh rel {eh:.2e}")
    assert ep <= 1e-10 and eh <= 1e-10


def test_tap_gradient_kernel_fp32_and_fp64(torch):
    for Nu in (5, 32, 100, 256):
        g = scan(Nu, 6, 9, 8, 8, 8, D=max(2500.0, Nu))
        st = stages(g)
        rng = np.random.default_rng(Nu)
        q = rng.standard_normal(st.sino_shape())
        G = rng.standard_normal(st.sino_shape())
        ref = O.tap_gradient(ogeom(g), q, G)
        d64 = st.tap_gradient(torch.from_numpy(q).cuda(), torch.from_numpy(G).cuda())
        d32 = st.tap_gradient(torch.from_numpy(q.astype(np.float32)).cuda(),
                              torch.from_numpy(G.astype(np.float32)).cuda())
        assert O.rel_rmse(d64, ref) <= 1e-12, Nu
        assert O.rel_rmse(d32, ref) <= 1e-5, Nu


def test_tap_gradient_odd_row_counts(torch):
    # an odd number of detector rows leaves the last row pair without its
    # second row (read as row 0 and masked to zero); one row is the extreme
    for Nu, Nv, Nang in ((37, 3, 3), (64, 1, 1), (100, 1, 5)):
        g = scan(Nu, Nv, Nang, 8, 8, 8, D=2500.0)
        st = stages(g)
        rng = np.random.default_rng(Nu + Nv)
        q = rng.standard_normal(st.sino_shape())
        G = rng.standard_normal(st.sino_shape())
        ref = O.tap_gradient(ogeom(g), q, G)
        d64 = st.tap_gradient(torch.from_numpy(q).cuda(), torch.from_numpy(G).cuda())
        d32 = st.tap_gradient(torch.from_numpy(q.astype(np.float32)).cuda(),
                              torch.from_numpy(G.astype(np.float32)).cuda())
        assert O.rel_rmse(d64, ref) <= 1e-12, (Nu, Nv, Nang)
        assert O.rel_rmse(d32, ref) <= 1e-5, (Nu, Nv, Nang)


def test_chunked_recomputation_matches_full(torch):
    """cfg4 geometry reduced to 64^3 / 64^2 / 60 projections, chunks of 7."""
    from fdk_autograd import fdk, ramlak_taps
    g = scan(64, 64, 60, 64, 64, 64, D=2500.0, tau=0.02)
    st = stages(g)
    ells = phantoms.random_ellipsoids(5, 64, 64.0)
    sino = st.phantom(ells)
    x_true = torch.zeros(st.vol_shape, device="cuda")
    taps = torch.from_numpy(ramlak_taps(st).astype(np.float32)).cuda()
    grads = []
    for chunk in (None, 7):
        p = sino.clone().requires_grad_(True)
        t = taps.clone().requires_grad_(True)
        vol = fdk(p, t, st, chunk=chunk)
        ((vol - x_true) ** 2).sum().backward()
        grads.append((p.grad.cpu().numpy(), t.grad.cpu().numpy()))
    e_p = O.rel_rmse(grads[1][0], grads[0][0])
    e_h = O.rel_rmse(grads[1][1], grads[0][1])
    say(f"  chunked vs full: dp {e_p:.2e}, dh {e_h:.2e}")
    # FP32 with atomic splats: summation order differs between the two runs
    assert e_p <= 1e-5 and e_h <= 1e-5


def test_cfg4_full_size_step(torch):
    """cfg4: 256^3, 360 x 256^2, learnable 512 taps; one forward+backward in FP32,
    gradient against the FP64 instantiation on the same input."""
    from fdk_autograd import fdk, ramlak_taps
    w = geometry.workloads()["cfg4"]
    st = stages(w.geom)
    assert st.p == 512
    sino = st.phantom(w.ellipsoids())
    taps64 = torch.from_numpy(ramlak_taps(st)).cuda()
    out = {}
    for dt in (torch.float64, torch.float32):
        p = sino.to(dt).clone().requires_grad_(True)
        t = taps64.to(dt).clone().requires_grad_(True)
        vol = fdk(p, t, st, chunk=50)
        loss = (vol ** 2).sum()
        loss.backward()
        out[dt] = (p.grad.double().cpu().numpy(), t.grad.double().cpu().numpy())
    ep = O.rel_rmse(out[torch.float32][0], out[torch.float64][0])
    eh = O.rel_rmse(out[torch.float32][1], out[torch.float64][1])
    say(f"  cfg4 FP32 vs FP64 gradient: dp {ep:.2e}, dh {eh:.2e}")
    assert ep <= 1e-4 and eh <= 1e-4


def test_cfg5_fp16_forward_chunked_backward(torch):
    """cfg5 at full size: 1024^3 from 1200 x 1536^2, learnable 4096 taps.  One
    step with the FP16-stored forward and 50-projection chunk recomputation in
    the backward pass, against the FP32 forward with the whole stack at once;
    loss = sum of squares over the central 256^3 crop of the volume."""
    import time
    from fdk_autograd import fdk, ramlak_taps
    w = geometry.workloads()["cfg5"]
    st = stages(w.geom)
    assert st.p == 4096
    sino = st.phantom(w.ellipsoids())
    taps = torch.from_numpy(ramlak_taps(st).astype(np.float32)).cuda()
    Nz, Ny, Nx = st.vol_shape
    crop = (slice(Nz // 2 - 128, Nz // 2 + 128), slice(Ny // 2 - 128, Ny // 2 + 128),
            slice(Nx // 2 - 128, Nx // 2 + 128))
    out = {}
    for name, storage, chunk in (("f16_chunk50", _lib.F16, 50), ("f32_whole", _lib.F32, None)):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        t0 = time.perf_counter()
        p = sino.clone().requires_grad_(True)
        t = taps.clone().requires_grad_(True)
        vol = fdk(p, t, st, storage=storage, chunk=chunk)
        loss = (vol[crop] ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        peak = (torch.cuda.max_memory_allocated() - base) / 2**30
        say(f"  cfg5 {name}: step {dt:.2f} s, peak HBM above the inputs {peak:.1f} GiB")
        out[name] = (p.grad.double(), t.grad.double().cpu().numpy())
        del p, t, vol, loss
    gp16, gh16 = out["f16_chunk50"]
    gp32, gh32 = out["f32_whole"]
    ep = float(torch.sqrt(((gp16 - gp32) ** 2).sum() / (gp32 ** 2).sum()))
    eh = O.rel_rmse(gh16, gh32)
    say(f"  cfg5 gradient, FP16 chunked vs FP32 whole: dp {ep:.2e}, dh {eh:.2e}")
    # the FP16 stack changes the forward volume by ~1e-4 relative, which the
    # crop loss's gradient inherits
    assert ep <= 1e-3 and eh <= 1e-3
