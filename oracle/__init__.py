"""CPU oracle for the FDK hot path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/libfdk_oracle.so`` (built from ``fdk_oracle.c`` by
``oracle/Makefile``): the FP32 mirror and FP64 restatement of the reference's
FBPG/kernel.cu plus analytic phantoms.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg import this package; the product
(``tomokvas_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, fields

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfdk_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle library in place (make)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class CGeom(C.Structure):
    _fields_ = [
        ("Nu", C.c_int), ("Nang", C.c_int), ("Nv", C.c_int),
        ("cX", C.c_float), ("cZ", C.c_float), ("dbeta", C.c_float),
        ("D", C.c_float), ("tau", C.c_float),
        ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
        ("cx", C.c_float), ("cy", C.c_float), ("cz", C.c_float),
        ("rot", C.c_float), ("scale", C.c_float),
        ("flipX", C.c_int), ("flipY", C.c_int), ("flipZ", C.c_int),
        ("B", C.c_int),
    ]


class CEllipsoid(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("ax", C.c_double * 3),
                ("R", C.c_double * 9), ("rho", C.c_double)]


@dataclass
class Geometry:
    """Engine-level geometry, as FilteredBackProjection's setters receive it."""
    Nu: int
    Nang: int
    Nv: int
    cX: float
    cZ: float
    dbeta: float
    D: float
    tau: float
    Nx: int
    Ny: int
    Nz: int
    cx: float
    cy: float
    cz: float = 0.0
    rot: float = 0.0
    scale: float = 1.0
    flipX: int = 0
    flipY: int = 0
    flipZ: int = 0
    B: int = 16

    def c(self) -> CGeom:
        return CGeom(*[getattr(self, f.name) for f in fields(self)])

    @property
    def p(self) -> int:
        return padded_length(self.Nu)

    @property
    def sino_shape(self):
        return (self.Nv, self.Nang, self.Nu)

    @property
    def vol_shape(self):
        return (self.Nz, self.Ny, self.Nx)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    G = C.POINTER(CGeom)
    fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    E = C.POINTER(CEllipsoid)
    sig = {
        "ofdk_padded_length": (C.c_int, [C.c_int]),
        "ofdk_ramlak_ref": (C.c_float, [C.c_int, C.c_float]),
        "ofdk_ramlak": (C.c_double, [C.c_int, C.c_double]),
        "ofdk_shepplogan": (C.c_double, [C.c_int, C.c_double]),
        "ofdk_fill_taps": (None, [C.c_int, C.c_int, C.c_double, C.c_int, dp]),
        "ofdk_spectrum_abs": (None, [dp, C.c_int, C.c_double, dp]),
        "ofdk_spectrum_complex": (None, [dp, C.c_int, dp, dp]),
        "ofdk_sincos": (None, [G, dp, dp]),
        "ofdk_slab_count_reference": (C.c_int, [C.c_int, C.c_int]),
        "ofdk_slab_count": (C.c_int, [C.c_int, C.c_int]),
        "ofdk_axis_tables": (None, [G, fp, fp, fp]),
        "ofdk_weight_f32": (None, [G, fp]),
        "ofdk_weight_f64": (None, [G, dp]),
        "ofdk_filter_f32": (None, [G, dp, fp]),
        "ofdk_filter_f64": (None, [G, dp, dp]),
        "ofdk_filter_complex_f32": (None, [G, dp, dp, fp]),
        "ofdk_filter_complex_f64": (None, [G, dp, dp, dp]),
        "ofdk_filter_direct_f64": (None, [G, C.c_int, dp]),
        "ofdk_backproject_f32": (None, [G, fp, fp]),
        "ofdk_backproject_f64": (None, [G, dp, dp]),
        "ofdk_backproject_points_f32": (None, [G, fp, C.c_int, C.c_int, ip, fp]),
        "ofdk_backproject_points_f64": (None, [G, dp, C.c_int, C.c_int, ip, dp]),
        "ofdk_forward_project_f32": (None, [G, fp, fp]),
        "ofdk_forward_project_f64": (None, [G, dp, dp]),
        "ofdk_accumulate_f32": (None, [C.c_size_t, fp, fp]),
        "ofdk_fdk_f32": (None, [G, C.c_int, fp, fp]),
        "ofdk_fdk_f64": (None, [G, C.c_int, dp, dp]),
        "ofdk_tap_gradient_f64": (None, [G, dp, dp, dp]),
        "ofdk_shepp_logan": (C.c_int, [C.c_double, E]),
        "ofdk_random_ellipsoids": (C.c_int, [C.c_uint64, C.c_int, C.c_double, E]),
        "ofdk_project_ellipsoids": (None, [G, E, C.c_int, dp]),
        "ofdk_voxelize_ellipsoids": (None, [G, E, C.c_int, dp]),
        "ofdk_philox4x32": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32)]),
        "ofdk_poisson_log": (None, [C.c_size_t, C.c_double, C.c_uint64, fp]),
        "ofdk_poisson_draw": (C.c_double, [C.c_double, C.c_uint64, C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _g(geom: Geometry):
    return C.byref(geom.c())


# ---------------------------------------------------------------- filter helpers
def padded_length(Nu: int) -> int:
    return _load().ofdk_padded_length(Nu)


def ramlak_ref(n: int, tau: float) -> float:
    return _load().ofdk_ramlak_ref(n, tau)


def ramlak(n: int, tau: float) -> float:
    return _load().ofdk_ramlak(n, tau)


def shepplogan(n: int, tau: float) -> float:
    return _load().ofdk_shepplogan(n, tau)


def taps(kind: int, Nu: int, tau: float) -> np.ndarray:
    p = padded_length(Nu)
    out = np.zeros(p, np.float64)
    _load().ofdk_fill_taps(kind, Nu, tau, p, out)
    return out


def spectrum_abs(t: np.ndarray, tau: float) -> np.ndarray:
    t = np.ascontiguousarray(t, np.float64)
    out = np.zeros(t.size // 2 + 1, np.float64)
    _load().ofdk_spectrum_abs(t, t.size, tau, out)
    return out


def spectrum_complex(h: np.ndarray):
    h = np.ascontiguousarray(h, np.float64)
    re = np.zeros(h.size, np.float64)
    im = np.zeros(h.size, np.float64)
    _load().ofdk_spectrum_complex(h, h.size, re, im)
    return re, im


def filter_spectrum(kind: int, Nu: int, tau: float) -> np.ndarray:
    return spectrum_abs(taps(kind, Nu, tau), tau)


# ---------------------------------------------------------------- geometry
def sincos(geom: Geometry):
    s = np.zeros(geom.Nang, np.float64)
    c = np.zeros(geom.Nang, np.float64)
    _load().ofdk_sincos(_g(geom), s, c)
    return s, c


def slab_count_reference(Nz: int, B: int) -> int:
    return _load().ofdk_slab_count_reference(Nz, B)


def slab_count(Nz: int, B: int) -> int:
    return _load().ofdk_slab_count(Nz, B)


def axis_tables(geom: Geometry):
    ax = np.zeros(geom.Nx, np.float32)
    ay = np.zeros(geom.Ny, np.float32)
    az = np.zeros(geom.Nz, np.float32)
    _load().ofdk_axis_tables(_g(geom), ax, ay, az)
    return ax, ay, az


# ---------------------------------------------------------------- stages
def _arr(a, dt):
    return np.ascontiguousarray(a, dtype=dt).copy()


def weight(geom, sino, f64=False):
    dt = np.float64 if f64 else np.float32
    out = _arr(sino, dt)
    (_load().ofdk_weight_f64 if f64 else _load().ofdk_weight_f32)(_g(geom), out)
    return out


def filter_rows(geom, sino, Hf, f64=False):
    dt = np.float64 if f64 else np.float32
    out = _arr(sino, dt)
    Hf = np.ascontiguousarray(Hf, np.float64)
    (_load().ofdk_filter_f64 if f64 else _load().ofdk_filter_f32)(_g(geom), Hf, out)
    return out


def filter_rows_complex(geom, sino, Hre, Him, f64=False):
    dt = np.float64 if f64 else np.float32
    out = _arr(sino, dt)
    f = _load().ofdk_filter_complex_f64 if f64 else _load().ofdk_filter_complex_f32
    f(_g(geom), np.ascontiguousarray(Hre, np.float64), np.ascontiguousarray(Him, np.float64), out)
    return out


def filter_direct(geom, sino, kind=0):
    out = _arr(sino, np.float64)
    _load().ofdk_filter_direct_f64(_g(geom), kind, out)
    return out


def backproject(geom, sino, f64=False):
    dt = np.float64 if f64 else np.float32
    s = np.ascontiguousarray(sino, dt)
    vol = np.zeros(geom.vol_shape, dt)
    (_load().ofdk_backproject_f64 if f64 else _load().ofdk_backproject_f32)(_g(geom), s, vol)
    return vol


def backproject_points(geom, sino, xyz, f64=False, row0=0):
    """Backprojection at the voxels xyz (x, y, global z).  row0 > 0: sino holds
    detector rows [row0, row0 + geom.Nv) of a taller detector, geom.cZ relative
    to row0; row positions are those of the taller detector (bit-identical)."""
    dt = np.float64 if f64 else np.float32
    s = np.ascontiguousarray(sino, dt)
    pts = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
    out = np.zeros(pts.shape[0], dt)
    f = _load().ofdk_backproject_points_f64 if f64 else _load().ofdk_backproject_points_f32
    f(_g(geom), s, int(row0), pts.shape[0], pts, out)
    return out


def forward_project(geom, vol, f64=False):
    dt = np.float64 if f64 else np.float32
    v = np.ascontiguousarray(vol, dt)
    sino = np.zeros(geom.sino_shape, dt)
    (_load().ofdk_forward_project_f64 if f64 else _load().ofdk_forward_project_f32)(_g(geom), v, sino)
    return sino


def accumulate(partial, slab):
    out = _arr(slab, np.float32)
    _load().ofdk_accumulate_f32(out.size, np.ascontiguousarray(partial, np.float32), out)
    return out


def fdk(geom, sino, kind=0, f64=False):
    dt = np.float64 if f64 else np.float32
    s = _arr(sino, dt)
    vol = np.zeros(geom.vol_shape, dt)
    (_load().ofdk_fdk_f64 if f64 else _load().ofdk_fdk_f32)(_g(geom), kind, s, vol)
    return vol


def tap_gradient(geom, q, G):
    dh = np.zeros(geom.p, np.float64)
    _load().ofdk_tap_gradient_f64(_g(geom), np.ascontiguousarray(q, np.float64),
                                  np.ascontiguousarray(G, np.float64), dh)
    return dh


# ---------------------------------------------------------------- phantoms
def _ell_array(n):
    return (CEllipsoid * n)()


def shepp_logan(half_fov: float):
    arr = _ell_array(10)
    _load().ofdk_shepp_logan(half_fov, arr)
    return arr


def random_ellipsoids(seed: int, n: int, fov: float):
    arr = _ell_array(n)
    _load().ofdk_random_ellipsoids(seed, n, fov, arr)
    return arr


def ellipsoids_to_numpy(arr) -> np.ndarray:
    """(n, 19) float64: centre(3), semi-axes(3), R(9 row-major), rho, pad(3)."""
    out = np.zeros((len(arr), 19), np.float64)
    for i, e in enumerate(arr):
        out[i, 0:3] = list(e.c)
        out[i, 3:6] = list(e.ax)
        out[i, 6:15] = list(e.R)
        out[i, 15] = e.rho
    return out


def project_ellipsoids(geom, ells):
    sino = np.zeros(geom.sino_shape, np.float64)
    _load().ofdk_project_ellipsoids(_g(geom), ells, len(ells), sino)
    return sino


def voxelize_ellipsoids(geom, ells):
    vol = np.zeros(geom.vol_shape, np.float64)
    _load().ofdk_voxelize_ellipsoids(_g(geom), ells, len(ells), vol)
    return vol


def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    _load().ofdk_philox4x32(c, k, o)
    return list(o)


def poisson_log(p, I0, seed):
    out = _arr(p, np.float32)
    _load().ofdk_poisson_log(out.size, I0, seed, out)
    return out


def poisson_draw(lam, seed, idx):
    return _load().ofdk_poisson_draw(lam, seed, idx)


# ---------------------------------------------------------------- metrics
def rel_rmse(a, ref) -> float:
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.sqrt(np.mean(ref * ref))
    return float(np.sqrt(np.mean((a - ref) ** 2)) / (den if den > 0 else 1.0))
