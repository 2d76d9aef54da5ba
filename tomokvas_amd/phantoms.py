"""Synthetic objects for benchmark and test input (no reference counterpart).

Ellipsoids are rows of 19 float64: centre (3), semi-axes (3), rotation (9, rows =
ellipsoid axes in world coordinates), density, 3 unused; lengths in virtual
detector pixels of the reconstruction frame.  The random generator is
splitmix64-driven so every run and machine reproduces the same object.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea & Flood 2014)."""

    def __init__(self, seed: int):
        self.s = seed & _M64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53


def _rot_z(deg: float) -> np.ndarray:
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])


# 3D Shepp-Logan head phantom (10 ellipsoids, original densities):
# rho, semi-axes a b c, centre x y z, rotation about z in degrees; unit cube [-1, 1].
SHEPP_LOGAN_3D = [
    (1.00, .6900, .9200, .810, 0.00, 0.0000, 0.00, 0),
    (-0.98, .6624, .8740, .780, 0.00, -0.0184, 0.00, 0),
    (-0.02, .1100, .3100, .220, 0.22, 0.0000, 0.00, -18),
    (-0.02, .1600, .4100, .280, -0.22, 0.0000, 0.00, 18),
    (0.01, .2100, .2500, .410, 0.00, 0.3500, -0.15, 0),
    (0.01, .0460, .0460, .050, 0.00, 0.1000, 0.25, 0),
    (0.01, .0460, .0460, .050, 0.00, -0.1000, 0.25, 0),
    (0.01, .0460, .0230, .050, -0.08, -0.6050, 0.00, 0),
    (0.01, .0230, .0230, .020, 0.00, -0.6060, 0.00, 0),
    (0.01, .0230, .0460, .020, 0.06, -0.6050, 0.00, 0),
]


def shepp_logan(half_fov: float) -> np.ndarray:
    out = np.zeros((10, 19), np.float64)
    for i, (rho, a, b, c, x, y, z, deg) in enumerate(SHEPP_LOGAN_3D):
        out[i, 0:3] = np.array([x, y, z]) * half_fov
        out[i, 3:6] = np.array([a, b, c]) * half_fov
        out[i, 6:15] = _rot_z(deg).ravel()
        out[i, 15] = rho
    return out


def random_ellipsoids(seed: int, n: int, fov: float) -> np.ndarray:
    """n ellipsoids: centres uniform in the ball of radius 0.4*fov (0.8*fov
    across), semi-axes U(0.02, 0.25)*fov, density U(0.1, 1.0), uniformly
    random rotation (Shoemake 1992 quaternion)."""
    rng = SplitMix64(seed)
    rad = 0.4 * fov
    out = np.zeros((n, 19), np.float64)
    for i in range(n):
        while True:
            c = [(2.0 * rng.uniform() - 1.0) * rad for _ in range(3)]
            if c[0] * c[0] + c[1] * c[1] + c[2] * c[2] <= rad * rad:
                break
        ax = []
        for k in range(3):
            out[i, k] = c[k]
            ax.append((0.02 + 0.23 * rng.uniform()) * fov)
        out[i, 3:6] = ax
        out[i, 15] = 0.1 + 0.9 * rng.uniform()
        u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
        a, b = math.sqrt(1 - u1), math.sqrt(u1)
        qw, qx = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
        qy, qz = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
        R = [
            1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw),
            2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw),
            2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy),
        ]
        out[i, 6:15] = R
    return out
