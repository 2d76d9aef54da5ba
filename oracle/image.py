"""CPU restatement of the reference's 16-bit slice scaling -- TEST
INFRASTRUCTURE ONLY (imported by tests/ alone).

ConvertToTiff (FBPG/image_file.cpp:104-110): u = (unsigned short)
(65535 * (x - min) / (max - min)) in FP64 with the volume's range; pinned in
tests/test_tiff.py by the closed forms min -> 0, max -> 65535 and the
midpoint -> 32767.  Values the reference leaves undefined (outside the range,
NaN, max == min) follow these rules: clamp, 0, all zeros.
"""
import numpy as np


def to_u16(x: np.ndarray, vmin: float, vmax: float) -> np.ndarray:
    if vmax == vmin:
        return np.zeros(np.shape(x), np.uint16)
    v = 65535 * (np.asarray(x, np.float32).astype(np.float64) - vmin) / (vmax - vmin)
    v = np.where(np.isnan(v), 0.0, np.clip(v, 0.0, 65535.0))
    return np.trunc(v).astype(np.uint16)
