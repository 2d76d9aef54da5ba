"""Committed fixtures (tests/golden/fixtures.npz, written by make_golden.py from
the FP64 restatement after independent cross-checks) -- CPU only."""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FX = np.load(os.path.join(HERE, "golden", "fixtures.npz"))
KEYS = ["Nu", "Nang", "Nv", "cX", "cZ", "dbeta", "D", "tau", "Nx", "Ny", "Nz", "cx", "cy", "cz",
        "rot", "scale", "flipX", "flipY", "flipZ", "B"]


def fixture_geom():
    vals = FX["fdk_geom"]
    d = {}
    for k, v in zip(KEYS, vals):
        d[k] = int(v) if k in ("Nu", "Nang", "Nv", "Nx", "Ny", "Nz", "flipX", "flipY", "flipZ", "B") else float(v)
    return O.Geometry(**d)


@pytest.mark.parametrize("p", [64, 256, 512, 2048, 4096])
def test_spectra_fixtures(p):
    for kind, name in ((0, "ramlak"), (1, "shepplogan")):
        hf = O.spectrum_abs(O.taps(kind, p // 2, 1.0), 1.0)
        assert np.array_equal(hf, FX[f"spectrum_{name}_p{p}"])
        assert FX[f"spectrum_{name}_p{p}"][p // 4] == pytest.approx(1 / 8) or name == "shepplogan"


def test_slab_coordinate_fixtures():
    for key in FX.files:
        if not key.startswith("az_"):
            continue
        nz, b = (int(s[1:]) for s in key[3:].split("_"))
        g = fixture_geom()
        g.Nz, g.B = nz, b
        assert np.array_equal(O.axis_tables(g)[2], FX[key])


def test_fdk_fixture():
    g = fixture_geom()
    sino = FX["fdk_sino"]
    assert np.array_equal(sino, O.project_ellipsoids(g, O.shepp_logan(16.0)).astype(np.float32))
    v64 = O.fdk(g, sino.astype(np.float64), kind=0, f64=True)
    assert O.rel_rmse(v64, FX["fdk_vol64"]) < 1e-14
    v32 = O.fdk(g, sino, kind=0)
    assert O.rel_rmse(v32, FX["fdk_vol64"]) < 1e-5
