"""CPU restatement of the reference's Scanco RSQ preprocessing -- TEST
INFRASTRUCTURE ONLY (imported by tests/ alone; the product never does).

Follows FBPG/rsq_load.cpp (class ``rsq``; the file is not part of the
reference's build and has no caller, so no output of it exists to compare
with).  Each function names the lines it restates.  Pinned in
tests/test_rsq_oracle.py by: the header layout against the field offsets of
``ima_data_type`` (FBPG/rsq_load.h:13-49, 512 bytes); the flat/dark +
beam-hardening formula against closed forms; the cut and the half->full-fan
rebinning against an analytic fan-beam sinogram of off-centre disks (the
rebinned half must reproduce the measured full sinogram, which fixes the
column mirror, the axis offset, the +-pi branch and the sign of the 2*gamma
term).  The fan->parallel path (``Half2Para``/``Fan2Para``) is not restated:
it is not built here.
"""
from __future__ import annotations

import math
import struct

import numpy as np

PI = 3.14159265358979323846

# ima_data_type (FBPG/rsq_load.h:13-49): 16 chars, 7 ints, 21 ints, 40 chars,
# 2 ints, 83 ints, data_offset
HEADER_FMT = "<16s7i21i40s2i83ii"
HEADER_FIELDS = (
    ["check", "data_type", "nr_of_bytes", "nr_of_blocks", "patient_index", "scanner_id",
     "creation_date0", "creation_date1", "dimx_p", "dimy_p", "dimz_p", "dimx_um", "dimy_um",
     "dimz_um", "slice_thickness_um", "slice_increment_um", "slice_1_pos_um", "min_data_value",
     "max_data_value", "mu_scaling", "nr_of_samples", "nr_of_projections", "scandist_um",
     "scanner_type", "sampletime_us", "index_measurement", "site", "reference_line_um",
     "recon_alg", "name", "energy", "intensity"] + [f"fill{i}" for i in range(83)] +
    ["data_offset"])
BH_OFFSET, BH_BYTES = 3432, 152        # rsq_load.cpp:494-496
DATA_OFFSET = 3584                     # rsq_load.cpp:12 (header_offset default)


def parse_header(buf: bytes) -> dict:
    """LoadHeader (rsq_load.cpp:489-552): the 512-byte header and the
    beam-hardening line "Corr at X0 X1 X2:  Y0 Y1 Y2" at byte 3432, whose end
    points give y = A x + B."""
    vals = struct.unpack_from(HEADER_FMT, buf, 0)
    h = dict(zip(HEADER_FIELDS, vals))
    bh = buf[BH_OFFSET:BH_OFFSET + BH_BYTES].split(b"\0")[0].decode("latin-1")
    head, tail = bh.split(":", 1)
    xs = [int(t) for t in head.split()[2:5]]
    ys = [float(t) for t in tail.split()[:3]]
    a = (ys[2] - ys[0]) / (xs[2] - xs[0])
    h.update(bh_x=xs, bh_y=ys, bh_a=a, bh_b=ys[0] - a * xs[0])
    return h


def sizes(sin_x, sin_y, axis_move, complete, minimal, distance, delta_fi):
    """CalculateSizeOfSinogram (rsq_load.cpp:82-105), parallel branch left out."""
    if complete:
        cx = 2 * sin_x - 1 - 2 * axis_move
        if minimal:
            return cx, int(math.ceil((PI + 2 * math.atan(cx / 2. / distance)) / delta_fi))
        return cx, sin_y
    return sin_x - axis_move, sin_y


def flat_field(block: np.ndarray, A: float, B: float, invert: bool) -> np.ndarray:
    """FlatFieldCorrection (rsq_load.cpp:299-325): row 0 dark, row 1 flat,
    rows 2.. the projections; A*log((I - dark)/(flat - dark)) + B, negated
    when inverting.  block: uint16 [sin_y + 2][sin_x] -> float64 [sin_y][sin_x]."""
    d = block[0].astype(np.int64)
    f = block[1].astype(np.int64)
    num = (block[2:].astype(np.int64) - d).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = num / (f - d).astype(np.float64)
        v = A * np.log(v) + B
    return -v if invert else v


def cut_half(ff: np.ndarray, axis_move: int, flip: bool) -> np.ndarray:
    """CutHalfSin (rsq_load.cpp:393-410): drop the first axis_move columns;
    flip reverses the row order."""
    out = ff[:, axis_move:].copy()
    return out[::-1].copy() if flip else out


def half_to_full_fan(ff: np.ndarray, axis_move: int, distance: float, delta_fi: float,
                     flip: bool, sin_y_cpl: int) -> np.ndarray:
    """Half2FullFan (rsq_load.cpp:327-391): output columns
    [sin_x-1-axis_move, 2 sin_x-2-2 axis_move] are the measured half (columns
    axis_move.. of the input); column sin_x-1-axis_move+t for t < 0 is the
    redundant ray: input column -t+axis_move at angle index
    (f*dfi - 2 atan(t/D) - pi)/dfi (+pi instead when outside [0, sin_y-1]),
    interpolated linearly between its ceil and floor rows (a ceil of sin_y wraps
    to row 0).  flip computes the last sin_y_cpl angles and writes them in
    reverse order."""
    sin_y, sin_x = ff.shape
    cx = 2 * sin_x - 1 - 2 * axis_move
    c0 = sin_x - 1 - axis_move                  # output column of t = 0
    start, stop = (sin_y - sin_y_cpl, sin_y) if flip else (0, sin_y_cpl)
    out = np.empty((sin_y_cpl, cx), np.float64)
    t = np.arange(-sin_x + 1 + axis_move, 0)    # negative offsets
    at = np.arctan(t / distance)
    col = -t + axis_move
    for f in range(start, stop):
        r = sin_y_cpl - 1 - f + start if flip else f
        out[r, c0:] = ff[f, axis_move:]
        fi_rad = f * delta_fi
        nf = (fi_rad - 2 * at - PI) / delta_fi
        bad = (nf > sin_y - 1) | (nf < 0)
        nf = np.where(bad, (fi_rad - 2 * at + PI) / delta_fi, nf)
        a = np.ceil(nf).astype(np.int64)
        b = np.floor(nf).astype(np.int64)
        assert (b >= 0).all() and (b <= sin_y - 1).all() and (a <= sin_y).all(), \
            "angle outside the scan (the reference reads out of bounds here)"
        Av = ff[np.where(a > sin_y - 1, a - sin_y, a), col]
        Bv = ff[b, col]
        out[r, :c0] = np.where(a != b, Av + (Bv - Av) * (-nf + a), Av)
    return out


def process(raw: np.ndarray, A: float, B: float, axis_move: int, distance: float,
            delta_fi: float, flip: bool, invert: bool, complete: bool,
            minimal: bool) -> np.ndarray:
    """ProccessAndCopy (rsq_load.cpp:270-297) over raw sinograms
    uint16 [n][sin_y + 2][sin_x] -> float64 [n][sin_y_cpl][sin_x_cpl]."""
    n, sy2, sin_x = raw.shape
    sin_y = sy2 - 2
    _, sy_cpl = sizes(sin_x, sin_y, axis_move, complete, minimal, distance, delta_fi)
    outs = []
    for k in range(n):
        ff = flat_field(raw[k], A, B, invert)
        if complete:
            outs.append(half_to_full_fan(ff, axis_move, distance, delta_fi, flip, sy_cpl))
        else:
            outs.append(cut_half(ff, axis_move, flip))
    return np.stack(outs)
