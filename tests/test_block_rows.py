"""Voxel-block subsampling used by the large-volume checks -- CPU only.

geometry.sub_cube picks a seeded n^3 block and geometry.detector_rows_of_block
the detector rows its backprojection reads; weight and row filter act on rows
independently, so the FDK of those rows alone (cZ shifted by the first row)
must reproduce the block of the full FDK.  Pinned here against the FP64
oracle's full reconstruction on small scans, with the axis flips, a z offset
and a start angle, and shown to be tight: a range with rows cut past its
two-row margin changes the block.  With the row positions evaluated on the
full detector (backproject_points' row0) the FP32 mirror on those rows is
bit-identical to the full-detector mirror, as the GPU's FP32 positions are.
"""
import numpy as np
import pytest

import oracle as O
from tomokvas_amd import geometry
from tomokvas_amd.geometry import ScanGeometry


def _points(x0, y0, z0, n):
    zz, yy, xx = np.meshgrid(np.arange(z0, z0 + n), np.arange(y0, y0 + n),
                             np.arange(x0, x0 + n), indexing="ij")
    return np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1).astype(np.int32)


def _fdk_block(g, sino, pts, lo=0, hi=None):
    """FP64 oracle FDK of detector rows [lo, hi) at the given voxels."""
    hi = g.Nv if hi is None else hi
    og = O.Geometry(**g.rows_oracle_dict(lo, hi))
    rows = np.ascontiguousarray(sino[lo:hi], np.float64)
    Hf = O.filter_spectrum(0, g.Nu, g.tau)
    f = O.filter_rows(og, O.weight(og, rows, f64=True), Hf, f64=True)
    return O.backproject_points(og, f, pts, f64=True, row0=lo)


CASES = {
    "plain": dict(flips=(0, 0, 0), block=(12, 10, 14)),
    "flip-xz": dict(flips=(1, 0, 1), block=(4, 18, 2), cz=3.5, rot=0.3),
    "flip-y": dict(flips=(0, 1, 0), block=(20, 3, 22), cz=-2.0, rot=-1.1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_block_rows_reproduce_full_fdk(name):
    c = CASES[name]
    n = 8
    g = ScanGeometry(Nu=48, Nv=80, Nang=36, tau=0.05, D=60.0, Nx=32, Ny=32, Nz=32,
                     scale=1.0, cz=c.get("cz", 0.0), rot=c.get("rot", 0.0),
                     flips=c["flips"], B=16)
    x0, y0, z0 = c["block"]
    rng = np.random.default_rng(7)
    sino = rng.standard_normal((g.Nv, g.Nang, g.Nu))
    pts = _points(x0, y0, z0, n)
    full = _fdk_block(g, sino, pts)
    lo, hi = geometry.detector_rows_of_block(g, x0, y0, z0, n)
    # a proper subset of the detector, so the check below means something
    assert 0 < lo and hi < g.Nv and hi - lo >= n, (lo, hi)
    sub = _fdk_block(g, sino, pts, lo, hi)
    scale = np.abs(full).max()
    assert scale > 0
    assert np.abs(sub - full).max() <= 1e-12 * scale
    # the two-row margin is all the slack there is: cutting a third row on
    # either side drops samples the block reads
    for lo2, hi2 in ((lo + 3, hi), (lo, hi - 4)):
        cut = _fdk_block(g, sino, pts, lo2, hi2)
        assert np.abs(cut - full).max() > 1e-3 * scale, (lo2, hi2)


def test_block_rows_bit_identical_in_fp32_with_row0():
    # the FP32 mirror on the block's rows alone, with row positions evaluated on
    # the full detector (row0), reproduces the full-detector mirror bit for bit:
    # the weight depends on v - cZ (an exact difference either way), the filter
    # acts on each row alone, and the shifted positions are exact
    g = ScanGeometry(Nu=48, Nv=300, Nang=36, tau=0.05, D=60.0, Nx=32, Ny=32, Nz=32,
                     scale=1.0, cZ=211.3, rot=0.4, flips=(0, 1, 1), B=16)
    x0, y0, z0, n = 9, 13, 6, 8
    rng = np.random.default_rng(3)
    sino = (rng.standard_normal((g.Nv, g.Nang, g.Nu)) + 4.0).astype(np.float32)
    pts = _points(x0, y0, z0, n)
    lo, hi = geometry.detector_rows_of_block(g, x0, y0, z0, n)
    assert 0 < lo and hi < g.Nv
    og = O.Geometry(**g.oracle_dict())
    osub = O.Geometry(**g.rows_oracle_dict(lo, hi))
    Hf = O.filter_spectrum(1, g.Nu, g.tau)
    wfull = O.weight(og, sino)
    wsub = O.weight(osub, np.ascontiguousarray(sino[lo:hi]))
    assert np.array_equal(wsub, wfull[lo:hi])
    ffull = O.filter_rows(og, wfull, Hf)
    fsub = O.filter_rows(osub, wsub, Hf)
    assert np.array_equal(fsub, ffull[lo:hi])
    full = O.backproject_points(og, ffull, pts)
    sub = O.backproject_points(osub, fsub, pts, row0=lo)
    assert np.array_equal(sub, full)
    # without row0 the positions are rounded on the short detector instead:
    # equal to FP32 rounding, not bit for bit
    shifted = O.backproject_points(osub, fsub, pts)
    assert not np.array_equal(shifted, full)
    assert np.abs(shifted - full).max() <= 1e-4 * np.abs(full).max()
    # FP64: the same either way to rounding
    f64 = O.filter_rows(og, O.weight(og, sino.astype(np.float64), f64=True), Hf, f64=True)
    o_full = O.backproject_points(og, f64, pts, f64=True)
    o_sub = O.backproject_points(osub, np.ascontiguousarray(f64[lo:hi]), pts, f64=True, row0=lo)
    assert np.abs(o_sub - o_full).max() <= 1e-12 * np.abs(o_full).max()


def test_sub_cube_inside_volume_and_near_isocentre():
    for w in geometry.workloads().values():
        g = w.geom
        n = min(64, g.Nx, g.Ny, g.Nz)
        half = int(256 / g.scale)
        for seed in range(20):
            x0, y0, z0, m = geometry.sub_cube(g, seed, n)
            assert m == n
            for o, N, c in ((x0, g.Nx, g.Nx // 2), (y0, g.Ny, g.Ny // 2), (z0, g.Nz, g.Nz // 2)):
                assert 0 <= o <= N - n
                assert abs(o + n // 2 - c) <= half
        assert geometry.sub_cube(g, 3, n) == geometry.sub_cube(g, 3, n)
