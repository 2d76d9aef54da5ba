"""Build the gfx950 FDK library in place: tomokvas_amd/libtomokvas_fdk.so, and
the command line program tomokvas_amd/tomokvas-fbp linked against it.

Every HIP/C++ source under tomokvas_amd/csrc is compiled by hipcc for gfx950
(objects under tomokvas_amd/_build/) and linked into one shared library that
exports the C ABI of include/tomokvas/fdk.h.

    python -m tomokvas_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtomokvas_fdk.so")
ARCH = "gfx950"

SOURCES = [
    "fdk_host.cpp",
    "fdk_engine.cpp",
    "fdk_filter.hip",
    "fdk_backproject.hip",
    "fdk_adjoint.hip",
    "fdk_phantom.hip",
    "fdk_probe.hip",
    "fdk_io.cpp",
    "fdk_driver.cpp",
    "fdk_rsq.hip",
    "fdk_tiff.hip",
    "fdk_tiff_host.cpp",
]
# the command line program, linked against the library (host code only)
MAIN = "fdk_main.cpp"
EXE = os.path.join(HERE, "tomokvas-fbp")

# -packed-fp32-ops: no packed FP32 add/mul/fma instructions (some of their forms
# have given wrong results, without an error, on MI355X).
# -munsafe-fp-atomics: float/double atomic adds map to the L2 atomic units.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [hipcc(), *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, MAIN), "-o", EXE + ".tmp", "-L", HERE, "-l:" + os.path.basename(LIB),
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{MAIN} failed:\n{r.stderr[-6000:]}")
    os.replace(EXE + ".tmp", EXE)
    if verbose:
        print(f"built {LIB} and {EXE}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
