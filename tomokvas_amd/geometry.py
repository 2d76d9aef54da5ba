"""Scan geometries in the engine's convention and the benchmark workloads.

The engine works on a virtual detector through the isocentre (FBPG/kernel.cu:
105-111: S, T, U = D/(D-S), u = T*U + cX); lengths are virtual pixels, tau is
the virtual pitch (the reference's "sampling", passed in cm:
FBPG/para_TC.cpp:139).  A flat detector of pitch d at source-detector distance
SDD with source-isocentre distance SOD maps to tau = d*SOD/SDD and D = SOD/tau.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, asdict

import numpy as np

from . import phantoms


@dataclass
class ScanGeometry:
    Nu: int
    Nv: int
    Nang: int
    tau: float              # virtual pixel pitch, cm
    D: float                # source-isocentre distance, virtual px
    Nx: int
    Ny: int
    Nz: int
    scale: float = 1.0      # voxel size in virtual px
    cX: float = None
    cZ: float = None
    cx: float = None
    cy: float = None
    cz: float = 0.0
    rot: float = 0.0
    B: int = 30             # slices per slab (FBPG/setting.xml:3)
    flips: tuple = (0, 0, 0)
    dbeta: float = None

    def __post_init__(self):
        if self.cX is None:
            self.cX = (self.Nu - 1) / 2.0
        if self.cZ is None:
            self.cZ = (self.Nv - 1) / 2.0
        if self.cx is None:
            self.cx = (self.Nx - 1) / 2.0
        if self.cy is None:
            self.cy = (self.Ny - 1) / 2.0
        if self.dbeta is None:
            self.dbeta = float(np.float32(2.0 * math.pi / self.Nang))

    @property
    def p(self) -> int:
        p = 2
        while p < self.Nu:
            p *= 2
        return 2 * p

    @property
    def fov_px(self) -> float:
        return self.Nx * self.scale

    @property
    def voxel_updates(self) -> int:
        return self.Nx * self.Ny * self.Nz * self.Nang

    def apply(self, fbp, sinogram=None, sink=None) -> None:
        """Drive the engine setters in the order of FBPG/para_TC.cpp:130-142."""
        from ._lib import check
        check(fbp.SetBlockSize(self.B), "SetBlockSize")
        check(fbp.SetAxisDirection(*self.flips), "SetAxisDirection")
        check(fbp.SetSinogramProperties(sinogram, self.Nu, self.cX, self.Nang, self.Nv, self.cZ,
                                        self.dbeta, self.D, self.tau), "SetSinogramProperties")
        check(fbp.SetReconstructionProperties(sink, self.Nx, self.cx, self.Ny, self.cy, self.Nz,
                                              self.cz, self.rot), "SetReconstructionProperties")
        check(fbp.SetScaleXYZ(self.scale), "SetScaleXYZ")

    def oracle_dict(self) -> dict:
        """Keyword arguments of oracle.Geometry."""
        return dict(Nu=self.Nu, Nang=self.Nang, Nv=self.Nv, cX=self.cX, cZ=self.cZ,
                    dbeta=self.dbeta, D=self.D, tau=self.tau, Nx=self.Nx, Ny=self.Ny,
                    Nz=self.Nz, cx=self.cx, cy=self.cy, cz=self.cz, rot=self.rot,
                    scale=self.scale, flipX=self.flips[0], flipY=self.flips[1],
                    flipZ=self.flips[2], B=self.B)

    def rows_oracle_dict(self, lo: int, hi: int) -> dict:
        """oracle.Geometry arguments for detector rows [lo, hi) alone: the float
        cZ the engine receives, less lo (an exact float difference), so the
        weight of every row equals the full detector's bit for bit."""
        d = self.oracle_dict()
        d.update(Nv=hi - lo, cZ=float(np.float32(self.cZ)) - lo)
        return d


def flat_detector(Nu: int, Nv: int, Nang: int, pitch_mm: float, sod_mm: float, sdd_mm: float,
                  Nx: int, Ny: int, Nz: int, scale: float, **kw) -> ScanGeometry:
    tau_mm = pitch_mm * sod_mm / sdd_mm
    return ScanGeometry(Nu=Nu, Nv=Nv, Nang=Nang, tau=tau_mm / 10.0, D=sod_mm / tau_mm,
                        Nx=Nx, Ny=Ny, Nz=Nz, scale=scale, **kw)


@dataclass
class Workload:
    name: str
    geom: ScanGeometry
    phantom: str            # "shepp-logan" or "ellipsoids"
    filter: int             # 0 Ram-Lak, 1 Shepp-Logan
    seed: int = 20251202
    n_ellipsoids: int = 64
    I0: float = None        # Poisson counting noise when set
    extra: dict = field(default_factory=dict)

    def ellipsoids(self) -> np.ndarray:
        if self.phantom == "shepp-logan":
            return phantoms.shepp_logan(self.geom.fov_px / 2.0)
        return phantoms.random_ellipsoids(self.seed, self.n_ellipsoids, self.geom.fov_px)

    def describe(self) -> dict:
        g = self.geom
        return {"workload": self.name, "detector": f"{g.Nu}x{g.Nv}", "projections": g.Nang,
                "volume": f"{g.Nx}x{g.Ny}x{g.Nz}", "scale": g.scale, "tau_cm": g.tau,
                "D_px": g.D, "phantom": self.phantom, "filter": ["ram-lak", "shepp-logan"][self.filter],
                "I0": self.I0, "slab": g.B}


def workloads() -> dict:
    """The five benchmark configurations (flat detector, SOD 500 mm, SDD 1000 mm)."""
    w = {}
    w["cfg1"] = Workload("cfg1 plain FP32 FDK 128^3, 180x128^2, Shepp-Logan phantom, Ram-Lak",
                         flat_detector(128, 128, 180, 0.4, 500, 1000, 128, 128, 128, 1.0),
                         "shepp-logan", 0)
    w["cfg2"] = Workload("cfg2 plain FP32 FDK 512^3, 900x1024^2, 64 ellipsoids + Poisson, Shepp-Logan filter",
                         flat_detector(1024, 1024, 900, 0.1, 500, 1000, 512, 512, 512, 2.0),
                         "ellipsoids", 1, I0=1e5)
    w["cfg3"] = Workload("cfg3 precision form 1024^3, 1800x2048^2, 64 ellipsoids + Poisson, Shepp-Logan filter",
                         flat_detector(2048, 2048, 1800, 0.1, 500, 1000, 1024, 1024, 1024, 2.0),
                         "ellipsoids", 1, I0=1e5)
    w["cfg4"] = Workload("cfg4 differentiation form 256^3, 360x256^2, 64 ellipsoids, learnable 512 taps",
                         flat_detector(256, 256, 360, 0.4, 500, 1000, 256, 256, 256, 1.0),
                         "ellipsoids", 0)
    w["cfg5"] = Workload("cfg5 differentiation + FP16 1024^3, 1200x1536^2, 64 ellipsoids",
                         flat_detector(1536, 1536, 1200, 0.1, 500, 1000, 1024, 1024, 1024, 1.5),
                         "ellipsoids", 0)
    return w


def sub_cube(g: ScanGeometry, seed: int, n: int = 64):
    """Seeded n^3 voxel block (x0, y0, z0, n) whose centre lies within 256
    virtual pixels of the isocentre: the voxel subsample that large volumes
    are checked on against the FP64 oracle."""
    rng = np.random.default_rng(seed)
    half = int(256 / g.scale)
    cx, cy, cz = g.Nx // 2, g.Ny // 2, g.Nz // 2
    x0 = int(np.clip(cx + rng.integers(-half, half) - n // 2, 0, g.Nx - n))
    y0 = int(np.clip(cy + rng.integers(-half, half) - n // 2, 0, g.Ny - n))
    z0 = int(np.clip(cz + rng.integers(-half, half) - n // 2, 0, g.Nz - n))
    return x0, y0, z0, n


def detector_rows_of_block(g: ScanGeometry, x0: int, y0: int, z0: int, n: int):
    """Detector row range [v_lo, v_hi) holding every sample the backprojection
    of the block x0..x0+n-1, y0.., z0.. reads, over all angles (two rows of
    margin each side); the cosine weight and the row filter act on rows
    independently, so these rows alone reproduce the block."""
    # voxel coordinates with the axis flips applied (kernel.cu:82-88, 398-400,
    # 501-506, 821)
    sx, sy, sz = (-1.0 if f else 1.0 for f in g.flips)
    ax = sx * g.scale * (np.arange(x0, x0 + n) - g.cx)
    ay = sy * g.scale * (np.arange(y0, y0 + n) - g.cy)
    az = sz * g.scale * (np.array([z0, z0 + n - 1]) - (g.Nz - 1) / 2 - g.cz)
    b = np.arange(g.Nang) * g.dbeta + g.rot
    X, Y = np.meshgrid(ax, ay)
    S = X.ravel()[:, None] * np.cos(b)[None, :] + Y.ravel()[:, None] * np.sin(b)[None, :]
    U = g.D / (g.D - S)
    v = np.concatenate([(az[0] * U).ravel(), (az[1] * U).ravel()]) + g.cZ
    lo = max(0, int(math.floor(v.min())) - 2)
    hi = min(g.Nv, int(math.ceil(v.max())) + 3)
    return lo, hi
