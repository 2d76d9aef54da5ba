"""Compiles tests/cpp/para_tc_host.cpp against the drop-in C++ class and links
it to tomokvas_amd/libtomokvas_fdk.so (run from __graft_entry__.build(); the
GPU test runs the program this leaves next to its source)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "para_tc_host.cpp")
EXE = os.path.join(HERE, "para_tc_host")


def build(out: str = EXE) -> str:
    pkg = os.path.join(ROOT, "tomokvas_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", out + ".tmp", "-L", pkg, "-l:libtomokvas_fdk.so",
           "-Wl,-rpath,$ORIGIN/../../tomokvas_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"para_tc_host.cpp failed to build:\n{r.stderr[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build())
