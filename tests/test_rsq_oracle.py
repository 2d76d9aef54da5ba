"""Pins of the RSQ preprocessing oracle (oracle/rsq.py), CPU only.

The reference's rsq_load.cpp is not built by its project and has no caller,
so no output of it exists: the oracle is pinned by the header's byte layout
(field offsets derived by hand from FBPG/rsq_load.h:13-49), closed forms of
the flat/dark + beam-hardening formula, hand-computed sizes, and the physics of
fan-beam redundancy (the rebinned half detector must reproduce the analytic
full-detector sinogram of off-centre disks).
"""
import math
import struct

import numpy as np
import pytest

from oracle import rsq as R


def make_header(fields: dict, bh: str) -> bytearray:
    buf = bytearray(R.DATA_OFFSET)
    # hand-derived offsets in ima_data_type: check[16] at 0, then ints
    offs = {"data_type": 16, "nr_of_bytes": 20, "nr_of_blocks": 24, "dimx_p": 44,
            "dimy_p": 48, "dimz_p": 52, "mu_scaling": 88, "nr_of_samples": 92,
            "nr_of_projections": 96, "scandist_um": 100, "reference_line_um": 120,
            "energy": 168, "intensity": 172, "data_offset": 508}
    buf[0:16] = b"CTDATA-HEADER_V1"
    for k, v in fields.items():
        struct.pack_into("<i", buf, offs[k], v)
    b = bh.encode()
    buf[R.BH_OFFSET:R.BH_OFFSET + len(b)] = b
    return buf


def test_header_layout_and_beam_hardening_line():
    assert struct.calcsize(R.HEADER_FMT) == 512
    buf = make_header({"dimx_p": 1024, "dimz_p": 496, "nr_of_projections": 248,
                       "mu_scaling": 4096, "energy": 70000, "data_offset": 6},
                      "Corr at 0 1000 2000:  0.5 2.0 3.5")
    h = R.parse_header(bytes(buf))
    assert h["check"] == b"CTDATA-HEADER_V1"
    assert (h["dimx_p"], h["dimz_p"], h["nr_of_projections"]) == (1024, 496, 248)
    assert (h["mu_scaling"], h["energy"], h["data_offset"]) == (4096, 70000, 6)
    # the end points give y = A x + B: A = (3.5 - 0.5) / 2000, B = 0.5
    assert h["bh_x"] == [0, 1000, 2000] and h["bh_y"] == [0.5, 2.0, 3.5]
    assert h["bh_a"] == 3.0 / 2000 and h["bh_b"] == 0.5


def test_sizes_known_answers():
    # Scanco defaults (rsq_load.cpp:11-22): 1024 x 2000, axis 21, D = 125810/24
    D, dfi = 125810. / 24., 2 * math.pi / 2000
    assert R.sizes(1024, 2000, 21, False, False, D, dfi) == (1003, 2000)
    assert R.sizes(1024, 2000, 21, True, False, D, dfi) == (2005, 2000)
    # pi + 2 atan(1002.5 / 5242.083) = 3.5195 rad = 1120.3 rows -> 1121
    assert R.sizes(1024, 2000, 21, True, True, D, dfi) == (2005, 1121)


def test_flat_field_closed_forms():
    sin_x = 4
    dark = np.array([100, 100, 7, 3], np.uint16)
    flat = np.array([10100, 300, 7 + 4096, 65535], np.uint16)
    # rows with I - dark = (flat - dark) * {1, 1/2, 1/4}
    rows = [dark + (flat - dark) // k for k in (1, 2, 4)]
    block = np.stack([dark, flat] + rows).astype(np.uint16)
    A, B = 2.0, 0.25
    v = R.flat_field(block, A, B, invert=False)
    want = np.array([[B] * sin_x, [B - A * math.log(2)] * sin_x, [B - A * math.log(4)] * sin_x])
    assert np.allclose(v, want, rtol=0, atol=1e-15)
    assert np.array_equal(R.flat_field(block, A, B, invert=True), -v)


def fan_sinogram(n_ang, cols, c0, D, Rsrc, disks, sense):
    """Line integrals of disks (x, y, r, mu) along the ray of fan angle
    gamma = atan(t / D), t = column - c0, from the source at Rsrc (cos b, sin b):
    direction angle b + pi - sense * gamma."""
    b = 2 * np.pi * np.arange(n_ang)[:, None] / n_ang
    g = np.arctan((np.arange(cols)[None, :] - c0) / D)
    sx, sy = Rsrc * np.cos(b), Rsrc * np.sin(b)
    th = b + np.pi - sense * g
    ux, uy = np.cos(th), np.sin(th)
    out = np.zeros((n_ang, cols))
    for (x, y, r, mu) in disks:
        d = (x - sx) * uy - (y - sy) * ux
        out += 2 * mu * np.sqrt(np.clip(r * r - d * d, 0, None))
    return out


DISKS = [(40.0, -25.0, 30.0, 1.0), (-55.0, 10.0, 22.0, 0.7), (5.0, 60.0, 15.0, 1.3)]


@pytest.mark.parametrize("axis_move", [0, 5, 21])
def test_half_to_full_fan_reproduces_the_full_detector(axis_move):
    sin_x, sin_y, D, Rsrc = 160, 1440, 600.0, 420.0
    cx = 2 * sin_x - 1 - 2 * axis_move
    c0 = sin_x - 1 - axis_move
    full = fan_sinogram(sin_y, cx, c0, D, Rsrc, DISKS, sense=1)
    half = full[:, c0 - axis_move:]                 # measured columns t >= -axis_move
    assert half.shape == (sin_y, sin_x)
    dfi = 2 * math.pi / sin_y
    out = R.half_to_full_fan(half, axis_move, D, dfi, flip=False, sin_y_cpl=sin_y)
    assert np.array_equal(out[:, c0:], full[:, c0:])   # the measured half, copied
    err = np.abs(out[:, :c0] - full[:, :c0])
    scale = np.abs(full).max()
    # linear interpolation between angle rows: smooth except at disk edges
    assert np.sqrt(np.mean(err ** 2)) < 2e-3 * scale and err.max() < 0.1 * scale
    # the opposite rotation sense is not what the rebinning assumes: O(1) errors
    other = fan_sinogram(sin_y, cx, c0, D, Rsrc, DISKS, sense=-1)
    bad = R.half_to_full_fan(other[:, c0 - axis_move:], axis_move, D, dfi, False, sin_y)
    assert np.sqrt(np.mean((bad[:, :c0] - other[:, :c0]) ** 2)) > 0.05 * scale


def test_flip_and_minimal_select_and_reverse_rows():
    sin_x, sin_y, D, axis_move = 96, 720, 400.0, 7
    rng = np.random.default_rng(3)
    ff = rng.standard_normal((sin_y, sin_x))
    dfi = 2 * math.pi / sin_y
    full = R.half_to_full_fan(ff, axis_move, D, dfi, False, sin_y)
    assert np.array_equal(R.half_to_full_fan(ff, axis_move, D, dfi, True, sin_y), full[::-1])
    _, m = R.sizes(sin_x, sin_y, axis_move, True, True, D, dfi)
    assert m < sin_y
    assert np.array_equal(R.half_to_full_fan(ff, axis_move, D, dfi, False, m), full[:m])
    assert np.array_equal(R.half_to_full_fan(ff, axis_move, D, dfi, True, m), full[-m:][::-1])
    cut = R.cut_half(ff, axis_move, flip=True)
    assert np.array_equal(cut, ff[::-1, axis_move:])


def test_process_chains_flat_field_and_rebinning():
    sin_x, sin_y, n = 48, 360, 3
    rng = np.random.default_rng(5)
    dark = rng.integers(90, 110, sin_x)
    flat = rng.integers(30000, 31000, sin_x)
    counts = rng.integers(200, 29000, (n, sin_y, sin_x))
    raw = np.concatenate([np.broadcast_to(dark, (n, 1, sin_x)),
                          np.broadcast_to(flat, (n, 1, sin_x)), counts], axis=1).astype(np.uint16)
    args = dict(A=1.5, B=-0.1, axis_move=4, distance=200.0, delta_fi=2 * math.pi / sin_y,
                flip=False, invert=True)
    out = R.process(raw, complete=True, minimal=False, **args)
    for k in range(n):
        ff = R.flat_field(raw[k], 1.5, -0.1, True)
        assert np.array_equal(out[k], R.half_to_full_fan(ff, 4, 200.0, args["delta_fi"], False,
                                                          sin_y))
    cut = R.process(raw, complete=False, minimal=False, **args)
    assert cut.shape == (n, sin_y, sin_x - 4)
