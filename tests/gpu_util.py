"""Helpers shared by the GPU tests (not collected: no test_ prefix)."""
from __future__ import annotations

import sys

import numpy as np
# torch ships its own HIP runtime; it has to be the first one loaded into the
# process (the library then binds to the same one), else torch finds no device.
# Every GPU test module imports this helper, so running one file alone keeps
# the order the full suite gets from its collection.
import torch  # noqa: F401

import oracle as O
from tomokvas_amd import _lib, geometry
from tomokvas_amd.engine import FilteredBackProjection, VolumeSink
from tomokvas_amd.stages import Stages


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scan(Nu, Nv, Nang, Nx, Ny, Nz, D=2500.0, tau=0.02, scale=1.0, B=16, **kw):
    """Engine geometry with float-valued parameters (as the setters receive them)."""
    f = lambda x: float(np.float32(x))  # noqa: E731
    g = geometry.ScanGeometry(Nu=Nu, Nv=Nv, Nang=Nang, tau=f(tau), D=f(D), Nx=Nx, Ny=Ny, Nz=Nz,
                              scale=f(scale), B=B, **kw)
    for name in ("cX", "cZ", "cx", "cy", "cz", "rot", "dbeta"):
        setattr(g, name, f(getattr(g, name)))
    return g


def ogeom(g: geometry.ScanGeometry) -> O.Geometry:
    return O.Geometry(**g.oracle_dict())


def engine(g, sino=None, sink=None, storage=_lib.F32, filt=0, taps=None, collecting=False,
           gpu=0):
    fbp = FilteredBackProjection(gpu)
    g.apply(fbp, sino, sink)
    _lib.check(fbp.SetCollecting(collecting), "SetCollecting")
    _lib.check(fbp.SetStorage(storage), "SetStorage")
    if taps is not None:
        _lib.check(fbp.SetFilter(_lib.FILTER_USER, taps), "SetFilter")
    else:
        _lib.check(fbp.SetFilter(filt), "SetFilter")
    return fbp


def run_engine(g, sino, storage=_lib.F32, filt=0):
    sink = VolumeSink(g.Nx, g.Ny, g.Nz, g.B)
    fbp = engine(g, sino, sink, storage=storage, filt=filt)
    _lib.check(fbp.Initialize(), "Initialize")
    _lib.check(fbp.CalculateReconstruction(), "CalculateReconstruction")
    times = fbp.stage_times_ms()
    fbp.close()
    return sink.volume.copy(), times, sink


def stages(g, filt=0, taps=None):
    fbp = engine(g, None, None, filt=filt, taps=taps)
    return Stages(fbp)


def ulp_diff(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7fffffff), a)
    b = np.where(b < 0, -(b & 0x7fffffff), b)
    return np.abs(a - b)
