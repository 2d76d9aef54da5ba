"""Pins of the CPU oracle (oracle/fdk_oracle.c) -- CPU only.

The reference ships no tests, fixtures or golden vectors and cannot be built
here (CUDA 7.0 + cuFFT + Win32), so every oracle function is pinned by
something other than its own formula:

  padded_length        closed-form values of kernel.cu:570-573
  ramlak / taps        Kak & Slaney closed form, float vs double agreement
  spectrum_abs         exact values Hf[0] = tau*sum(t), Hf[p/4] = 1/(8 tau),
                       Hf[p/2] = tau*|alternating sum|, numpy.fft cross-check,
                       ramp shape ~ k/(2 p tau)
  filter (FFT)         direct O(Nu^2) convolution; numpy.fft; impulse response;
                       linearity
  weight               numpy closed form; value 1 on the central ray; symmetry
  axis tables / slabs  centring identity aZ = g - (Nz-1)/2; reference slab count
                       (and its dropped last slice) on hand-worked cases
  backproject          constant projections -> 2 pi / p at the isocentre;
                       single-impulse footprint from hand-computed U, T;
                       90-degree rotation symmetry; axis flips; uniform ball FDK
                       reproduces the density; FP32 mirror vs FP64
  forward_project      dot-product adjoint test against backproject (FP64)
  tap_gradient         finite differences of the filtered output
  phantom projector    exact sphere chords 2 sqrt(r^2 - d^2); Python generator
                       identical to the C generator
  philox / poisson     Random123 known-answer vectors; Poisson mean/variance
"""
import math

import numpy as np
import pytest

import oracle as O
from tomokvas_amd import phantoms


def geom(N=32, Nang=24, Nv=None, Nz=None, scale=1.0, tau=0.02, D=2500.0, B=16, Nx=None,
         Ny=None, **kw):
    Nv = N if Nv is None else Nv
    Nz = N if Nz is None else Nz
    Nx = N if Nx is None else Nx
    Ny = N if Ny is None else Ny
    # float-valued parameters, as the engine's setters receive them
    return O.Geometry(Nu=N, Nang=Nang, Nv=Nv, cX=(N - 1) / 2, cZ=(Nv - 1) / 2,
                      dbeta=float(np.float32(2 * np.pi / Nang)), D=float(np.float32(D)),
                      tau=float(np.float32(tau)),
                      Nx=Nx, Ny=Ny, Nz=Nz, cx=(Nx - 1) / 2, cy=(Ny - 1) / 2,
                      scale=scale, B=B, **kw)


# ----------------------------------------------------------------- filter construction
def test_padded_length_closed_form():
    # p = 2 * nextpow2(Nu) with a floor of 2 before doubling (kernel.cu:570-573)
    cases = {1: 4, 2: 4, 3: 8, 4: 8, 5: 16, 24: 64, 32: 64, 33: 128, 128: 256, 256: 512,
             900: 2048, 1024: 2048, 1536: 4096, 2048: 4096}
    for nu, p in cases.items():
        assert O.padded_length(nu) == p


def test_ramlak_closed_form():
    tau = 0.05
    assert O.ramlak(0, tau) == pytest.approx(1 / (4 * tau * tau), rel=1e-15)
    for n in (2, -2, 4, 10):
        assert O.ramlak(n, tau) == 0.0
    for n in (1, -1, 3, -7, 101):
        assert O.ramlak(n, tau) == pytest.approx(-1 / (n * n * np.pi ** 2 * tau * tau), rel=1e-15)
    # the reference rounds the double result to float (kernel.cu:792 returns float)
    for n in range(-9, 10):
        assert O.ramlak_ref(n, np.float32(tau)) == np.float32(O.ramlak(n, float(np.float32(tau))))


def test_taps_layout():
    # t[i] = h(i - Nu/2)/2 for i < Nu, zero padded to p (kernel.cu:783-790)
    Nu, tau = 24, 0.3
    t = O.taps(0, Nu, tau)
    assert t.size == 64
    assert np.all(t[Nu:] == 0)
    assert t[Nu // 2] == pytest.approx(1 / (8 * tau * tau))
    assert t[Nu // 2 + 1] == pytest.approx(-1 / (2 * np.pi ** 2 * tau * tau))
    sl = O.taps(1, Nu, tau)
    assert sl[Nu // 2] == pytest.approx(1 / (np.pi ** 2 * tau * tau))   # 2/(pi^2 tau^2)/2


@pytest.mark.parametrize("Nu,tau", [(24, 0.3), (32, 1.0), (128, 1.0), (256, 0.02), (2048, 0.005)])
def test_spectrum_exact_values(Nu, tau):
    t = O.taps(0, Nu, tau)
    p = t.size
    Hf = O.spectrum_abs(t, tau)
    assert Hf.size == p // 2 + 1
    assert Hf[0] == pytest.approx(tau * t.sum(), rel=1e-10, abs=1e-12 / tau)
    assert Hf[p // 4] == pytest.approx(1 / (8 * tau), rel=1e-12)
    alt = np.sum(t * (-1.0) ** np.arange(p))
    assert Hf[p // 2] == pytest.approx(tau * abs(alt), rel=1e-12)
    # independent transform: numpy's FFT of the same taps
    ref = tau * np.abs(np.fft.rfft(t))
    assert np.max(np.abs(Hf - ref)) <= 1e-10 * np.max(ref)
    # ramp shape: |k| / (2 p tau) away from DC (truncation error ~1/k)
    k = np.arange(p // 16, p // 2 + 1)
    assert np.all(np.abs(Hf[k] / (k / (2 * p * tau)) - 1) < 0.03)


def test_spectrum_complex_matches_numpy():
    rng = np.random.default_rng(1)
    h = rng.standard_normal(64)
    re, im = O.spectrum_complex(h)
    ref = np.fft.fft(h)
    assert np.allclose(re, ref.real, atol=1e-12) and np.allclose(im, ref.imag, atol=1e-12)


# ----------------------------------------------------------------- filtering
@pytest.mark.parametrize("Nu", [24, 32, 40])
def test_filter_equals_direct_convolution(Nu):
    # for these Nu the centred kernel's DFT is real and positive, so tau*|DFT|
    # filtering is the linear convolution with g(n) = h(n)/2, times tau*p
    g = geom(N=Nu, Nang=5, Nv=3)
    rng = np.random.default_rng(Nu)
    sino = rng.standard_normal(g.sino_shape)
    Hf = O.filter_spectrum(0, Nu, g.tau)
    fft = O.filter_rows(g, sino, Hf, f64=True)
    direct = O.filter_direct(g, sino, 0)
    assert np.max(np.abs(fft - direct)) <= 1e-11 * np.max(np.abs(direct))


def test_filter_matches_numpy_and_is_linear():
    g = geom(N=37, Nang=4, Nv=3)
    p = g.p
    rng = np.random.default_rng(7)
    a = rng.standard_normal(g.sino_shape)
    b = rng.standard_normal(g.sino_shape)
    Hf = O.filter_spectrum(1, g.Nu, g.tau)
    fa = O.filter_rows(g, a, Hf, f64=True)
    # independent: numpy rfft/irfft (irfft normalises by 1/p -> multiply back)
    ref = np.fft.irfft(np.fft.rfft(a, n=p, axis=-1) * Hf, n=p, axis=-1)[..., : g.Nu] * p
    assert np.max(np.abs(fa - ref)) <= 1e-11 * np.max(np.abs(ref))
    fab = O.filter_rows(g, 2 * a - 3 * b, Hf, f64=True)
    fb = O.filter_rows(g, b, Hf, f64=True)
    assert np.max(np.abs(fab - (2 * fa - 3 * fb))) <= 1e-10 * np.max(np.abs(fab))
    # FP32 mirror agrees with the FP64 restatement
    f32 = O.filter_rows(g, a.astype(np.float32), Hf)
    assert O.rel_rmse(f32, fa) < 1e-6


def test_filter_impulse_response():
    g = geom(N=32, Nang=1, Nv=1)
    x = np.zeros(g.sino_shape)
    x[0, 0, 10] = 1.0
    Hf = O.filter_spectrum(0, g.Nu, g.tau)
    y = O.filter_rows(g, x, Hf, f64=True)[0, 0]
    expect = np.array([g.tau * g.p * O.ramlak(m - 10, g.tau) / 2 if -16 <= m - 10 <= 15 else 0.0
                       for m in range(32)])
    assert np.max(np.abs(y - expect)) <= 1e-10 * np.max(np.abs(expect))


def test_filter_complex_spectrum_of_user_taps():
    # H = DFT(h) for spatial taps h: output = p * circular convolution of the padded row
    g = geom(N=16, Nang=2, Nv=2)
    p = g.p
    rng = np.random.default_rng(3)
    h = rng.standard_normal(p)
    x = rng.standard_normal(g.sino_shape)
    re, im = O.spectrum_complex(h)
    y = O.filter_rows_complex(g, x, re, im, f64=True)
    xp = np.zeros(x.shape[:-1] + (p,))
    xp[..., :16] = x
    ref = np.zeros_like(xp)
    for m in range(p):
        for j in range(p):
            ref[..., m] += h[j] * xp[..., (m - j) % p]
    assert np.allclose(y, p * ref[..., :16], rtol=1e-10, atol=1e-10)


# ----------------------------------------------------------------- weight
def test_weight_closed_form():
    g = geom(N=20, Nang=3, Nv=11)
    x = np.ones(g.sino_shape)
    w = O.weight(g, x, f64=True)
    u = np.arange(g.Nu) - g.cX
    v = np.arange(g.Nv) - g.cZ
    ref = g.D / np.sqrt(u[None, :] ** 2 + g.D ** 2 + v[:, None] ** 2)
    assert np.allclose(w[:, 0, :], ref, rtol=1e-14)
    assert np.allclose(w, w[:, :1, :])             # independent of the angle
    assert np.allclose(w, w[::-1, :, ::-1])        # symmetric about the centre
    g2 = geom(N=21, Nang=1, Nv=21)                 # integer centre -> weight 1 there
    w2 = O.weight(g2, np.ones(g2.sino_shape), f64=True)
    assert w2[10, 0, 10] == 1.0
    w32 = O.weight(g, x.astype(np.float32))
    assert np.max(np.abs(w32 - w)) < 1e-7


# ----------------------------------------------------------------- slabs / coordinates
def test_reference_slab_count_and_dropped_slice():
    # nr_of_Z_blocks = ceil((Nz-1)/B) (kernel.cu:398-399)
    assert O.slab_count_reference(128, 30) == 5 and O.slab_count(128, 30) == 5
    assert O.slab_count_reference(1024, 30) == 35 and O.slab_count(1024, 30) == 35
    assert O.slab_count_reference(512, 30) == 18
    # (Nz-1) % B == 0: the reference visits one slab too few (slice 16 of 17 is lost)
    assert O.slab_count_reference(17, 16) == 1 and O.slab_count(17, 16) == 2
    assert O.slab_count_reference(1, 16) == 0 and O.slab_count(1, 16) == 1


@pytest.mark.parametrize("Nz,B", [(128, 30), (17, 16), (33, 16), (1, 16), (1023, 30), (100, 7)])
def test_slab_centring_identity(Nz, B):
    # dZ_n = -cz + (n - nZ/2) B + dcorZ gives aZ = scale*(g - (Nz-1)/2 - cz)
    g = geom(N=8, Nz=Nz, B=B, scale=1.5)
    _, _, az = O.axis_tables(g)
    gz = np.arange(Nz)
    exact = 1.5 * (gz - (Nz - 1) / 2)
    assert np.max(np.abs(az - exact)) <= 4e-6 * max(1.0, np.max(np.abs(exact)))


def test_axis_flip_tables():
    g = geom(N=9, Nz=9, flipX=1, flipZ=1)
    ax, ay, az = O.axis_tables(g)
    g0 = geom(N=9, Nz=9)
    bx, by, bz = O.axis_tables(g0)
    assert np.array_equal(ax, -bx) and np.array_equal(ay, by) and np.array_equal(az, -bz)


# ----------------------------------------------------------------- backprojection
def test_backproject_constant_projection_closed_form():
    # Q = 1 everywhere: voxel at the isocentre has U = 1 for every angle, so the
    # value is Nang * dbeta / p = 2 pi / p (float dbeta)
    g = geom(N=17, Nang=12, Nz=17)
    q = np.ones(g.sino_shape)
    vol = O.backproject(g, q, f64=True)
    assert vol[8, 8, 8] == pytest.approx(12 * g.dbeta / g.p, rel=1e-12)
    # a voxel off-centre: sum of U^2 over the angles, U = D/(D - S)
    s, c = O.sincos(g)
    aX, aY = 3.0, -2.0
    S = aX * c + aY * s
    assert vol[8, 8 - 2, 8 + 3] == pytest.approx(g.dbeta / g.p * np.sum((g.D / (g.D - S)) ** 2),
                                                 rel=1e-12)
    # outside the cylinder |(aX,aY)| < Nu/2 the value is 0
    assert vol[8, 0, 0] == 0.0


def test_backproject_single_impulse_footprint():
    # one projection (beta = 0), one unit sample at (u0, v0); hand-computed U, T
    g = geom(N=16, Nang=1, Nz=16, D=40.0, tau=0.1)
    g.dbeta = 0.5
    q = np.zeros(g.sino_shape)
    u0, v0 = 9, 8
    q[v0, 0, u0] = 1.0
    vol = O.backproject(g, q, f64=True)
    # beta = 0: S = aX, T = aY, U = D/(D - aX); u = aY*U + cX, v = aZ*U + cZ
    ref = np.zeros_like(vol)
    for z in range(16):
        for y in range(16):
            for x in range(16):
                aX, aY, aZ = x - 7.5, y - 7.5, z - 7.5
                if aX * aX + aY * aY >= 64:
                    continue
                U = 40.0 / (40.0 - aX)
                u, v = aY * U + 7.5, aZ * U + 7.5
                wu = max(0.0, 1 - abs(u - u0))
                wv = max(0.0, 1 - abs(v - v0))
                ref[z, y, x] = U * U * wu * wv * 0.5 / g.p
    assert np.max(np.abs(vol - ref)) <= 1e-12
    assert np.count_nonzero(ref) > 10


def test_backproject_rotation_and_flip_symmetry():
    rng = np.random.default_rng(5)
    g = geom(N=16, Nang=8, Nz=6)
    q = rng.standard_normal(g.sino_shape)
    v0 = O.backproject(g, q, f64=True)
    # rotating the data by 2 angles (90 degrees) rotates the volume by 90 deg about z
    q2 = np.roll(q, -2, axis=1)
    v2 = O.backproject(g, q2, f64=True)
    # angle shift +90: (x, y) -> (y, -x) in centred coordinates
    # (the float angle step makes 8*dbeta differ from 2*pi by ~1e-7 rad)
    assert np.allclose(v2, np.rot90(v0, k=1, axes=(1, 2)), rtol=0, atol=1e-5 * np.abs(v0).max())
    gf = geom(N=16, Nang=8, Nz=6, flipX=1)
    vf = O.backproject(gf, q, f64=True)
    assert np.allclose(vf, v0[:, :, ::-1], atol=1e-12 * np.abs(v0).max())


def test_fdk_uniform_ball_reproduces_density():
    N, Nang = 48, 90
    g = geom(N=N, Nang=Nang, Nz=N)
    ells = O.random_ellipsoids(0, 1, 1.0)     # one ball of radius 14 px, density 0.5
    for k in range(3):
        ells[0].c[k] = 0.0
        ells[0].ax[k] = 14.0
    for k in range(9):
        ells[0].R[k] = float(np.eye(3).ravel()[k])
    ells[0].rho = 0.5
    sino = O.project_ellipsoids(g, ells)
    vol = O.fdk(g, sino, kind=0, f64=True)
    c = N // 2
    core = vol[c - 3:c + 3, c - 6:c + 6, c - 6:c + 6]
    assert abs(core.mean() - 0.5) < 0.02
    assert abs(vol[c, c, 2]) < 0.05      # outside the ball


def test_mirror_vs_fp64_shepp_logan():
    g = geom(N=32, Nang=40, Nz=32)
    sino = O.project_ellipsoids(g, O.shepp_logan(16.0))
    v32 = O.fdk(g, sino.astype(np.float32))
    v64 = O.fdk(g, sino, f64=True)
    assert O.rel_rmse(v32, v64) < 1e-5


def test_backproject_points_matches_volume():
    rng = np.random.default_rng(9)
    g = geom(N=12, Nang=7, Nz=10)
    q = rng.standard_normal(g.sino_shape)
    vol = O.backproject(g, q, f64=True)
    pts = np.array([[0, 0, 0], [5, 6, 3], [11, 2, 9], [6, 6, 5]], np.int32)
    vals = O.backproject_points(g, q, pts, f64=True)
    assert np.allclose(vals, [vol[z, y, x] for x, y, z in pts], rtol=0, atol=1e-15)


# ----------------------------------------------------------------- adjoint pieces
def test_forward_project_is_adjoint_of_backproject():
    rng = np.random.default_rng(11)
    g = geom(N=14, Nang=9, Nv=12, Nz=11, D=60.0)
    y = rng.standard_normal(g.sino_shape)
    x = rng.standard_normal(g.vol_shape)
    lhs = np.vdot(O.backproject(g, y, f64=True), x)
    rhs = np.vdot(y, O.forward_project(g, x, f64=True))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_tap_gradient_finite_differences():
    g = geom(N=8, Nang=3, Nv=2)
    p = g.p
    rng = np.random.default_rng(2)
    q = rng.standard_normal(g.sino_shape)
    G = rng.standard_normal(g.sino_shape)
    h = rng.standard_normal(p)

    def loss(hh):
        re, im = O.spectrum_complex(hh)
        return np.sum(G * O.filter_rows_complex(g, q, re, im, f64=True))

    dh = O.tap_gradient(g, q, G)
    eps = 1e-6
    for j in (0, 1, 5, p - 1):
        e = np.zeros(p)
        e[j] = eps
        fd = (loss(h + e) - loss(h - e)) / (2 * eps)
        assert dh[j] == pytest.approx(fd, rel=1e-6, abs=1e-6)


# ----------------------------------------------------------------- phantoms
def test_sphere_chord_lengths():
    g = geom(N=40, Nang=4, Nv=40, D=1000.0, tau=0.01)
    ells = O.random_ellipsoids(0, 1, 1.0)
    for k in range(3):
        ells[0].c[k] = 0.0
        ells[0].ax[k] = 10.0
    for k in range(9):
        ells[0].R[k] = float(np.eye(3).ravel()[k])
    ells[0].rho = 2.0
    sino = O.project_ellipsoids(g, ells)
    s, c = O.sincos(g)
    for (v, b, u) in [(20, 0, 20), (19, 1, 25), (10, 2, 16), (30, 3, 27), (5, 0, 5)]:
        src = np.array([g.D * c[b], g.D * s[b], 0.0])
        det = np.array([-(u - g.cX) * s[b], (u - g.cX) * c[b], v - g.cZ])
        d = det - src
        d /= np.linalg.norm(d)
        t = -src @ d
        dist2 = src @ src - t * t
        chord = 2 * math.sqrt(max(0.0, 100.0 - dist2))
        assert sino[v, b, u] == pytest.approx(2.0 * chord * g.tau, rel=1e-10, abs=1e-12)


def test_python_ellipsoid_generator_matches_c():
    for seed in (0, 1, 20251202):
        py = phantoms.random_ellipsoids(seed, 16, 1024.0)
        c = O.ellipsoids_to_numpy(O.random_ellipsoids(seed, 16, 1024.0))
        assert np.array_equal(py[:, :16], c[:, :16])
        R = py[:, 6:15].reshape(-1, 3, 3)
        assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-12)
        r = np.linalg.norm(py[:, :3], axis=1)
        assert np.all(r <= 0.4 * 1024.0)
        assert np.all((py[:, 3:6] >= 0.02 * 1024) & (py[:, 3:6] <= 0.25 * 1024))
        assert np.all((py[:, 15] >= 0.1) & (py[:, 15] <= 1.0))
    sl = phantoms.shepp_logan(64.0)
    slc = O.ellipsoids_to_numpy(O.shepp_logan(64.0))
    assert np.allclose(sl[:, :16], slc[:, :16], rtol=0, atol=1e-12)


# ----------------------------------------------------------------- noise
def test_philox_known_answers():
    # Random123 philox4x32-10 known-answer vectors
    assert O.philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox4x32([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e,
                                                               0xa20bc7c6, 0x6d5451fd]
    assert O.philox4x32([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                        [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420,
                                                      0x24126ea1]


@pytest.mark.parametrize("lam", [3.0, 25.0, 1e3, 1e5])
def test_poisson_moments(lam):
    n = 20000
    draws = np.array([O.poisson_draw(lam, 77, i) for i in range(n)])
    assert np.all(draws == np.round(draws)) and np.all(draws >= 0)
    se = math.sqrt(lam / n)
    assert abs(draws.mean() - lam) < 5 * se
    assert abs(draws.var() / lam - 1) < 0.06


def test_poisson_log_noise_level():
    p = np.full(50000, 2.0, np.float32)
    out = O.poisson_log(p, 1e5, 5)
    lam = 1e5 * math.exp(-2.0)
    # delta-method std of -log(n/I0) is 1/sqrt(lam)
    assert abs(out.mean() - 2.0) < 5 / math.sqrt(lam * p.size) + 1e-4
    assert abs(out.std() * math.sqrt(lam) - 1) < 0.03
