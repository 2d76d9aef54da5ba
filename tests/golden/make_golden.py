"""Writes the committed fixtures in tests/golden/ from the oracle's FP64
restatement, cross-checking each against an independent computation first.

    python tests/golden/make_golden.py

fixtures.npz:
  spectrum_ramlak_p{P}, spectrum_shepplogan_p{P}   tau*|DFT(taps)|, Nu = P/2, tau = 1
  az_N{Nz}_B{B}                                    float voxel z-coordinates (slab recentring)
  fdk_sino, fdk_vol64                              32^3 Shepp-Logan FDK (24 x 32^2), FP64
  fdk_geom                                         its geometry (float parameters)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

GEOM = dict(Nu=32, Nang=24, Nv=32, cX=15.5, cZ=15.5, dbeta=float(np.float32(2 * np.pi / 24)),
            D=100.0, tau=float(np.float32(0.2)), Nx=32, Ny=32, Nz=32, cx=15.5, cy=15.5, cz=0.0,
            rot=0.0, scale=1.0, flipX=0, flipY=0, flipZ=0, B=16)
AZ_CASES = [(128, 30), (17, 16), (1023, 30), (100, 7)]


def build() -> dict:
    out = {}
    for p in (64, 256, 512, 2048, 4096):
        for kind, name in ((0, "ramlak"), (1, "shepplogan")):
            t = O.taps(kind, p // 2, 1.0)
            hf = O.spectrum_abs(t, 1.0)
            ref = np.abs(np.fft.rfft(t))          # independent transform
            assert np.max(np.abs(hf - ref)) < 1e-12 * np.max(ref)
            out[f"spectrum_{name}_p{p}"] = hf
    for nz, b in AZ_CASES:
        g = O.Geometry(**{**GEOM, "Nz": nz, "B": b})
        az = O.axis_tables(g)[2]
        assert np.max(np.abs(az - (np.arange(nz) - (nz - 1) / 2))) < 1e-5 * nz
        out[f"az_N{nz}_B{b}"] = az
    g = O.Geometry(**GEOM)
    sino = O.project_ellipsoids(g, O.shepp_logan(16.0)).astype(np.float32)
    vol = O.fdk(g, sino.astype(np.float64), kind=0, f64=True)
    direct = O.backproject(g, O.filter_direct(g, O.weight(g, sino.astype(np.float64), f64=True), 0),
                           f64=True)
    assert O.rel_rmse(vol, direct) < 1e-12      # FFT filter == direct convolution here
    out["fdk_sino"] = sino
    out["fdk_vol64"] = vol
    out["fdk_geom"] = np.array([GEOM[k] for k in GEOM], np.float64)
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **build())
    print("wrote", os.path.join(HERE, "fixtures.npz"))
