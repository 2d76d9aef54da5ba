"""Synthetic Scanco RSQ files for the RSQ tests (not collected)."""
import struct

import numpy as np

from oracle import rsq as R


def write_rsq(path, raw: np.ndarray, bh=(0, 1000, 2000, 0.0, 0.75, 1.5), dimz=0,
              nr_of_projections=0, header_offset=R.DATA_OFFSET, tail=0):
    """Header (ima_data_type fields at their byte offsets), the beam-hardening
    line at byte 3432, raw sinograms uint16 [n][sin_y+2][sin_x] at
    header_offset, and `tail` spare bytes."""
    buf = bytearray(max(header_offset, R.BH_OFFSET + R.BH_BYTES))
    buf[0:16] = b"CTDATA-HEADER_V1"
    n, sy2, sx = raw.shape
    for off, v in ((44, sx), (48, sx), (52, dimz), (92, sx), (96, nr_of_projections),
                   (508, header_offset // 512 - 1)):
        struct.pack_into("<i", buf, off, v)
    line = "Corr at %d %d %d:  %r %r %r" % bh
    buf[R.BH_OFFSET:R.BH_OFFSET + len(line)] = line.encode()
    with open(path, "wb") as f:
        f.write(bytes(buf[:header_offset]))
        f.write(raw.astype("<u2").tobytes())
        f.write(b"\0" * tail)


def synthetic_raw(n, sin_y, sin_x, seed):
    """Dark ~100, flat ~30000, counts between them."""
    rng = np.random.default_rng(seed)
    dark = rng.integers(90, 110, (n, 1, sin_x))
    flat = rng.integers(29000, 31000, (n, 1, sin_x))
    frac = np.exp(-rng.uniform(0.05, 3.0, (n, sin_y, sin_x)))
    counts = dark + np.round((flat - dark) * frac).astype(np.int64)
    return np.concatenate([dark, flat, counts], axis=1).astype(np.uint16)
