"""Write a synthetic scan as a .CRA header and raw projection file.

The projections are exact line integrals of a workload's ellipsoid phantom
(with its Poisson noise when it has one), computed on the GPU by the engine's
phantom projector, so that ``tomokvas-fbp <dir>/<name>.CRA`` reconstructs the
phantom.  The header is chosen so that the reconstruction program derives the
workload's geometry: sample spacing = tau, distance = D * tau (D px), centre
(Nu-1)/2, SCALE = 1/scale with X/Z DIMENSION = N * scale (the program's clamps
may still shorten z, FBPG/mucat_file.cpp:150-160).

    python -m tomokvas_amd.make_cra --workload cfg2 /tmp/scan cfg2
"""
from __future__ import annotations

import argparse
import ctypes as C
import math
import os
import time

import numpy as np

from . import _lib, geometry
from ._hip import hip
from .engine import FilteredBackProjection


def write_cra_files(folder: str, name: str, g: geometry.ScanGeometry, stack: np.ndarray,
                    source: str) -> str:
    """Writes <folder>/<name>.raw (the projections [beta][v][u] as float32, the
    order the reader expects, FBPG/mucat_file.cpp:236-269) from `stack`
    [v][beta][u] and the .CRA header from which the reconstruction program
    derives the geometry g; returns the header's path."""
    os.makedirs(folder, exist_ok=True)
    assert stack.shape == (g.Nv, g.Nang, g.Nu)
    raw = f"{name}.raw"
    with open(os.path.join(folder, raw), "wb") as f:          # [beta][v][u]
        step = max(1, (256 << 20) // (g.Nu * g.Nv * 4))
        for b0 in range(0, g.Nang, step):
            f.write(np.ascontiguousarray(stack[:, b0:b0 + step, :].transpose(1, 0, 2))
                    .astype(np.float32).tobytes())
    spacing_um = g.tau * 1e4
    lines = [
        f"FILE NAME:              {raw}",
        f"SOURCE NAME:            {source}",
        f"SAMPLE SPACING:         {spacing_um!r}",
        f"DISTANCE:               {g.D * spacing_um * 1e-4!r}",
        f"DETECTOR X PIXELS:      {g.Nu}",
        f"DETECTOR Z PIXELS:      {g.Nv}",
        f"NUMBER OF PROJECTIONS:  {g.Nang}",
        f"X DIMENSION:            {int(round(g.Nx * g.scale))}",
        f"Z DIMENSION:            {int(round(g.Nz * g.scale))}",
        f"ROTATE ANGLE:           {math.degrees(g.rot)!r}",
        f"CENTRE:                 {g.cX!r}",
        "NUMBER OF BLOCKS:       1",
        f"SCALE:                  {1.0 / g.scale!r}",
    ]
    path = os.path.join(folder, f"{name}.CRA")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def write_cra(folder: str, name: str, workload: str = "cfg1", seed: int | None = None) -> str:
    """A workload's synthetic scan: its projections made on the GPU, written
    by write_cra_files."""
    w = geometry.workloads()[workload]
    g = w.geom
    fbp = FilteredBackProjection(0)
    g.apply(fbp, None, None)
    _lib.check(fbp.Initialize(), "Initialize")
    L = _lib.load()
    ells = np.ascontiguousarray(w.ellipsoids())
    d_sino = fbp.device_sinogram_ptr()
    _lib.check(L.fdk_phantom_project(fbp.handle, ells.ctypes.data, ells.shape[0], 0, g.Nang,
                                     C.c_void_p(d_sino), None), "phantom")
    if w.I0:
        _lib.check(L.fdk_poisson_log(C.c_void_p(d_sino), g.Nu * g.Nv * g.Nang, w.I0,
                                     w.seed if seed is None else seed, None), "poisson")
    hip().hipDeviceSynchronize()
    stack = np.empty((g.Nv, g.Nang, g.Nu), np.float32)            # [v][beta][u]
    assert hip().hipMemcpy(C.c_void_p(stack.ctypes.data), C.c_void_p(d_sino), stack.nbytes,
                           2) == 0
    fbp.close()
    return write_cra_files(folder, name, g, stack, f"synthetic {workload}")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="cfg1")
    ap.add_argument("folder")
    ap.add_argument("name")
    a = ap.parse_args()
    t0 = time.perf_counter()
    path = write_cra(a.folder, a.name, a.workload)
    size = os.path.getsize(os.path.join(a.folder, a.name + ".raw"))
    print(f"wrote {path} and {size / 1e9:.2f} GB of projections in "
          f"{time.perf_counter() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
