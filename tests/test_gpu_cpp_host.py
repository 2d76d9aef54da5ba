"""The drop-in C++ class (include/tomokvas/FilteredBackProjection.hpp) in a
running program: tests/cpp/para_tc_host.cpp follows FBPG/para_TC.cpp:116-153
(one engine per angle block, SetGPUid after construction, SetCollecting for
every block after the first, the projection buffer freed after Initialize,
slabs through a bin_siz-like GetBlock(int, bool) sink).  Its volume must equal
the oracle's one-pass FDK of all angles."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

from gpu_util import ogeom, say, scan

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "para_tc_host")


@pytest.mark.parametrize("blocks", [1, 3])
def test_para_tc_call_sequence_through_the_cpp_class(tmp_path, blocks):
    assert os.path.exists(EXE), "tests/cpp/para_tc_host is built by __graft_entry__.build()"
    g = scan(64, 48, 90, 56, 60, 41, D=400.0, tau=0.02, scale=1.1, B=16)
    sino = O.project_ellipsoids(ogeom(g), O.shepp_logan(28.0)).astype(np.float32)
    src, dst = tmp_path / "in.raw", tmp_path / "out.raw"
    sino.tofile(src)
    cmd = [EXE, str(src), str(dst), str(g.Nu), str(g.Nv), str(g.Nang), str(g.Nx), str(g.Ny),
           str(g.Nz), str(g.B), repr(g.D), repr(g.tau), repr(g.scale), str(blocks)]
    say(f"para_tc_host, {blocks} angle block(s): {' '.join(cmd[3:])}")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    say(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0, r.stderr
    nslab = -(-g.Nz // g.B)
    assert f"slabs handed out {blocks * nslab}, collecting {(blocks - 1) * nslab}" in r.stdout
    vol = np.fromfile(dst, np.float32).reshape(g.Nz, g.Ny, g.Nx)
    ref = O.fdk(ogeom(g), sino, kind=0)
    e = O.rel_rmse(vol, ref)
    say(f"  rel RMSE vs the oracle's FP32 mirror (one pass) {e:.3e}")
    # angle blocks differ from one pass by FP32 summation order and the float
    # rotation offset of each block
    assert e <= 1e-5
