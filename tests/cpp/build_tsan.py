"""Builds tests/cpp/tsan_host: the library's threaded host code (the .CRA
reader's parallel reads, the .bin writer's I/O thread, the reconstruction
program's loader thread, the TIFF writer threads) compiled by g++ with
-fsanitize=thread, over the CPU stand-ins of tsan_stubs.cpp.  Host code only:
no GPU code is compiled."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "tomokvas_amd", "csrc")
HOST_SOURCES = ["fdk_io.cpp", "fdk_driver.cpp", "fdk_host.cpp", "fdk_tiff_host.cpp"]


def build(out: str) -> str:
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    srcs += [os.path.join(HERE, "tsan_stubs.cpp"), os.path.join(HERE, "tsan_host.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-ffp-contract=off",
           "-Wno-unknown-pragmas", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", out, "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ThreadSanitizer build failed:\n{r.stderr[-4000:]}")
    return out
