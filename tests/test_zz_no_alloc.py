"""fdk_tap_gradient_noalloc (the allocation-free tap gradient, scratch in a
file-scope device array) against fdk_tap_gradient through the C ABI: FP32 and
FP64, a geometry whose partial spectra exceed the array's share (fewer
workgroups), two calls in a row with the second on fewer angles."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tomokvas_amd import _lib

from gpu_util import ogeom, say, scan, stages

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def both(st, torch, q, G, nang):
    L = _lib.load()
    code = _lib.F64 if q.dtype == torch.float64 else _lib.F32
    ref = np.zeros(st.p, np.float64)
    got = np.zeros(st.p, np.float64)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(L.fdk_tap_gradient(st.fbp.handle, code, q.data_ptr(), G.data_ptr(), nang,
                                  ref.ctypes.data, s), "fdk_tap_gradient")
    _lib.check(L.fdk_tap_gradient_noalloc(st.fbp.handle, code, q.data_ptr(), G.data_ptr(), nang,
                                          got.ctypes.data, s), "fdk_tap_gradient_noalloc")
    return ref, got


@pytest.mark.parametrize("geo", ["small", "capped64", "capped32"])
def test_noalloc_tap_gradient_matches_the_allocating_one(torch, geo):
    f64, f32 = (torch.float64, 1e-12), (torch.float32, 2e-6)
    if geo == "small":
        g, kinds = scan(37, 6, 9, 8, 8, 8, D=2500.0), (f64, f32)
    elif geo == "capped64":
        # p = 2048: the FP64 partials of the 1280 workgroups a call of 64 x 40
        # rows takes exceed the 32 MiB array, so the allocation-free call runs 1022
        g, kinds = scan(1024, 64, 40, 8, 8, 8, D=10000.0), (f64,)
    else:
        # p = 4096 (FP32 only): 1280 workgroups' partials exceed it, 1021 run
        g, kinds = scan(2048, 64, 40, 8, 8, 8, D=10000.0), (f32,)
    st = stages(g)
    rng = np.random.default_rng(17)
    for dt, tol in kinds:
        for nang in (g.Nang, g.Nang // 2):        # the second call on fewer angles
            qn = rng.standard_normal((g.Nv, nang, g.Nu))
            Gn = rng.standard_normal((g.Nv, nang, g.Nu))
            q = torch.from_numpy(qn).to("cuda", dt).contiguous()
            G = torch.from_numpy(Gn).to("cuda", dt).contiguous()
            ref, got = both(st, torch, q, G, nang)
            e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            say(f"  {geo} {dt} nang={nang}: no-alloc vs allocating rel {e:.2e}")
            assert e <= tol
            if geo == "small" and dt == torch.float64:
                og = ogeom(g)
                gd = dict(g.oracle_dict(), Nang=nang)
                oref = O.tap_gradient(O.Geometry(**gd), qn, Gn)
                assert O.rel_rmse(got, oref) <= 1e-12
