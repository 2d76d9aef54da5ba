#!/usr/bin/env python3
"""FDK benchmark on one MI355X: backprojection GUPS and FDK seconds, FP16-stored
vs FP32 filtered projections, for the 1024^3 workload (1800 x 2048^2).

One "step" = one pass of the hot path over the resident projection stack:
cosine weight + ramp filter (one LDS FFT kernel), repack into the storage type,
backprojection of the whole volume (fdk_calculate_reconstruction without a slab
sink: the input is already in HBM and the volume stays there).  Stage times come
from HIP events recorded on the engine's stream around each launch.

Prints ONE JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tomokvas_amd import _lib, geometry  # noqa: E402
from tomokvas_amd.engine import FilteredBackProjection  # noqa: E402

METRIC = ("backprojection GUPS + FDK seconds @1024³: FP16-store vs FP32, "
          "rel-RMSE≤1e-3")
L2_PEAK_GBS = 8 * 16 * 128 * 2.4  # 8 XCD x 16 channels x 128 B/clk x 2.4 GHz = 39321.6 GB/s
HBM_PEAK_GBS = 8000.0
# non-packed VALU issue: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (one wave64 instruction
# per SIMD every 4 clocks)
VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9
NAMES = {_lib.F32: "f32", _lib.F16: "f16", _lib.BF16: "bf16", _lib.F16_PACKED: "f16-packed"}
CODES = {v: k for k, v in NAMES.items()}
CELL_BYTES = {_lib.F32: 4, _lib.F16: 2, _lib.BF16: 2, _lib.F16_PACKED: 2}   # bytes per stored sample


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hip():
    _lib.load()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    h.hipHostFree.argtypes = [C.c_void_p]
    return h


def end_to_end(fbp, g, hipm, storage, steps):
    """The projections handed over in host memory (page-locked), as the
    engine's boundary receives them: fdk_initialize's upload followed by the
    FDK (the reference's order) against fdk_reconstruct_host, which uploads
    angle chunks behind the filtering and backprojection.  Both include the
    engine's allocations; the volume stays in HBM."""
    L = _lib.load()
    n = g.Nu * g.Nv * g.Nang
    host = C.c_void_p()
    assert hipm.hipHostMalloc(C.byref(host), n * 4, 0) == 0
    assert hipm.hipMemcpy(host, C.c_void_p(fbp.device_sinogram_ptr()), n * 4, 2) == 0
    arr = np.ctypeslib.as_array(C.cast(host, C.POINTER(C.c_float)), (n,))
    _lib.check(fbp.SetStorage(storage), "SetStorage")
    nullcb = _lib.GET_BLOCK_FN()
    nvol = g.Nx * g.Ny * g.Nz
    ref = C.c_void_p()
    assert hipm.hipMalloc(C.byref(ref), nvol * 4) == 0
    serial, fused = [], []
    for _ in range(steps + 1):
        g.apply(fbp, arr, None)
        t0 = time.perf_counter()
        _lib.check(fbp.Initialize(), "Initialize")
        _lib.check(L.fdk_calculate_reconstruction(fbp.handle, nullcb, None), "calculate")
        hipm.hipDeviceSynchronize()
        serial.append(time.perf_counter() - t0)
    assert hipm.hipMemcpy(ref, C.c_void_p(fbp.device_volume_ptr()), nvol * 4, 3) == 0
    for _ in range(steps + 1):
        g.apply(fbp, None, None)
        t0 = time.perf_counter()
        _lib.check(L.fdk_reconstruct_host(fbp.handle, host, nullcb, None), "reconstruct_host")
        hipm.hipDeviceSynchronize()
        fused.append(time.perf_counter() - t0)
    out3 = (C.c_double * 3)()
    _lib.check(L.fdk_compare(C.c_void_p(fbp.device_volume_ptr()), ref, nvol, out3, None),
               "compare")
    hipm.hipFree(ref)
    hipm.hipHostFree(host)
    ts, tf = float(np.mean(serial[1:])), float(np.mean(fused[1:]))
    return {"storage": NAMES[storage], "host": "page-locked", "projection_bytes": n * 4,
            "upload_then_fdk_s": ts, "overlapped_s": tf,
            "end_to_end_gups": g.voxel_updates / tf / 1e9,
            "rel_rmse_overlapped_vs_serial": math.sqrt(out3[0] / out3[1])}


def rsq_leg(hipm, steps):
    """Scanco RSQ preprocessing on the GPU (fdk_rsq_process): one detector block
    of 248 sinograms of 2000 projections x 1024 columns (the scanner's layout,
    1.02 GB of 16-bit counts) rebinned to full fan (2000 x 2005) in FP32.  The
    algorithmic bytes are each raw sample read once and each output sample
    written once; the work is one FP64 log per raw sample and the rebinning of
    the redundant half from the FP64 line integrals."""
    L = _lib.load()
    p = _lib.RsqParams()
    _lib.check(L.fdk_rsq_default_params(C.byref(p)), "rsq params")
    p.complete, p.invert = 1, 1
    n = 248
    sx_cpl, sy_cpl = 2 * p.sin_x - 1 - 2 * p.axis_move, p.sin_y
    rng = np.random.default_rng(3)
    raw = rng.integers(200, 30000, (n, p.sin_y + 2, p.sin_x), dtype=np.uint16)
    raw[:, 0, :] = 100
    raw[:, 1, :] = 31000
    d_raw, d_out = C.c_void_p(), C.c_void_p()
    out_bytes = n * sy_cpl * sx_cpl * 4
    assert hipm.hipMalloc(C.byref(d_raw), raw.nbytes) == 0
    assert hipm.hipMalloc(C.byref(d_out), out_bytes) == 0
    assert hipm.hipMemcpy(d_raw, C.c_void_p(raw.ctypes.data), raw.nbytes, 1) == 0
    ts = []
    for _ in range(steps + 1):
        hipm.hipDeviceSynchronize()
        t0 = time.perf_counter()
        _lib.check(L.fdk_rsq_process(C.byref(p), 1.1, 0.1, d_raw, n, _lib.F32, d_out, None), "rsq")
        hipm.hipDeviceSynchronize()
        ts.append(time.perf_counter() - t0)
    hipm.hipFree(d_raw)
    hipm.hipFree(d_out)
    t = float(np.mean(ts[1:]))
    alg = raw.nbytes + out_bytes
    return {"sinograms": n, "raw_bytes": raw.nbytes, "output_bytes": out_bytes, "ms": 1e3 * t,
            "algorithmic_gbs": alg / t / 1e9, "frac_hbm": alg / t / 1e9 / HBM_PEAK_GBS}


def tiff_leg(fbp, g, hipm, folder):
    """TIFF export of the device volume (fdk_tiff_export_device): min/max and
    16-bit scaling on the GPU, one file per slice into `folder`."""
    L = _lib.load()
    os.makedirs(folder, exist_ok=True)
    nvol = g.Nx * g.Ny * g.Nz
    dvol = C.c_void_p(fbp.device_volume_ptr())
    mm = (C.c_double * 2)()
    hipm.hipDeviceSynchronize()
    t0 = time.perf_counter()
    _lib.check(L.fdk_volume_minmax(dvol, nvol, mm, None), "minmax")
    t_mm = time.perf_counter() - t0
    d_u16 = C.c_void_p()
    assert hipm.hipMalloc(C.byref(d_u16), nvol * 2) == 0
    hipm.hipDeviceSynchronize()
    t0 = time.perf_counter()
    _lib.check(L.fdk_volume_to_u16(dvol, nvol, mm[0], mm[1], d_u16, None), "to_u16")
    hipm.hipDeviceSynchronize()
    t_u16 = time.perf_counter() - t0
    hipm.hipFree(d_u16)
    t0 = time.perf_counter()
    _lib.check(L.fdk_tiff_export_device(dvol, g.Nx, g.Ny, g.Nz, folder.encode(), 0, mm), "export")
    t_all = time.perf_counter() - t0
    return {"slices": g.Nz, "minmax_ms": 1e3 * t_mm, "minmax_gbs": 4 * nvol / t_mm / 1e9,
            "to_u16_ms": 1e3 * t_u16, "to_u16_gbs": 6 * nvol / t_u16 / 1e9,
            "export_s": t_all, "files_bytes": 2 * nvol}


def slab_pipeline(fbp, g, hipm, storage, steps):
    """FDK with every slab handed to the host (page-locked memory, the writer's
    case): the backprojection runs in chunks while the previous chunk's slabs
    leave on a second stream.  Returns the stage times and the transfer time
    left exposed after the last backprojection."""
    L = _lib.load()
    plane = g.Nx * g.Ny
    nslab = -(-g.Nz // g.B)
    host = C.c_void_p()
    assert hipm.hipHostMalloc(C.byref(host), nslab * g.B * plane * 4, 0) == 0
    base = host.value
    cb = _lib.GET_BLOCK_FN(lambda user, n, collecting: base + n * g.B * plane * 4)
    _lib.check(fbp.SetStorage(storage), "SetStorage")
    _lib.check(L.fdk_calculate_reconstruction(fbp.handle, cb, None), "slab warmup")
    walls, stages = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        _lib.check(L.fdk_calculate_reconstruction(fbp.handle, cb, None), "slab step")
        walls.append(time.perf_counter() - t0)
        stages.append(fbp.stage_times_ms())
    # the last slabs must hold the device volume
    last = np.empty(plane, np.float32)
    z = g.Nz - 1
    assert hipm.hipMemcpy(last.ctypes.data, C.c_void_p(fbp.device_volume_ptr() + z * plane * 4),
                          plane * 4, 2) == 0
    got = np.ctypeslib.as_array(C.cast(C.c_void_p(base + z * plane * 4),
                                       C.POINTER(C.c_float)), (plane,))
    same = bool(np.array_equal(got, last))
    # the same bytes in one device-to-host copy: what the slab transfers cost
    # when they do not overlap the backprojection
    vol_bytes = plane * g.Nz * 4
    hipm.hipDeviceSynchronize()
    t0 = time.perf_counter()
    assert hipm.hipMemcpy(C.c_void_p(base), C.c_void_p(fbp.device_volume_ptr()), vol_bytes, 2) == 0
    t_copy = time.perf_counter() - t0
    hipm.hipHostFree(host)
    st = np.array(stages).mean(axis=0)
    chunk = C.c_int()
    _lib.check(L.fdk_get_volume_chunk(fbp.handle, C.byref(chunk)), "chunk")
    return {"storage": NAMES[storage], "host": "page-locked", "chunk_slices": chunk.value,
            "filter_ms": float(st[0]), "pack_ms": float(st[1]),
            "bp_ms": float(st[2]), "exposed_transfer_ms": float(st[3]),
            "wall_ms": 1e3 * float(np.mean(walls)), "volume_bytes": vol_bytes,
            "d2h_gbs": vol_bytes / t_copy / 1e9, "serial_transfer_ms": 1e3 * t_copy,
            "slabs_equal_device_volume": same}


def read_profile(workload: str, storage: str):
    """(HBM bytes, VALU instructions per voxel-update) of one backprojection step
    (its launches summed), measured by rocprofv3 --pmc (profiles/bp_traffic_
    <workload>.txt, lines 'storage fetch_bytes write_bytes valu_per_update');
    (None, None) if absent."""
    path = os.path.join(ROOT, "profiles", f"bp_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        parts = line.split()
        if len(parts) == 4 and parts[0] == storage and not line.startswith("#"):
            return float(parts[1]) + float(parts[2]), float(parts[3])
    return None, None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def block_check(fbp, w, g, block, block_vols, hipm, threads, reps):
    """CPU baseline and FP64 errors on a seeded voxel block of a workload, from
    its detector rows only: the FP32 OpenMP mirror (weight + filter of the rows,
    backprojection of the block; median of `reps`) and the FP64 oracle, and each
    GPU volume's block (block_vols: storage name -> z-slab [n][Ny][Nx]) against
    both.  Returns (timing dict, errors dict)."""
    import oracle as O
    os.environ["OMP_NUM_THREADS"] = str(threads)
    x0, y0, z0, n = block
    lo, hi = geometry.detector_rows_of_block(g, x0, y0, z0, n)
    rows = hi - lo
    raw = np.empty((rows, g.Nang, g.Nu), np.float32)
    src = fbp.device_sinogram_ptr() + lo * g.Nang * g.Nu * 4
    rc = hipm.hipMemcpy(raw.ctypes.data, C.c_void_p(src), raw.nbytes, 2)
    assert rc == 0, "hipMemcpy D2H failed"
    og = O.Geometry(**g.rows_oracle_dict(lo, hi))
    zz, yy, xx = np.meshgrid(np.arange(z0, z0 + n), np.arange(y0, y0 + n),
                             np.arange(x0, x0 + n), indexing="ij")
    pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1).astype(np.int32)
    Hf = O.filter_spectrum(w.filter, g.Nu, g.tau)
    tf, tb = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        f32 = O.filter_rows(og, O.weight(og, raw), Hf)
        t1 = time.perf_counter()
        m32 = O.backproject_points(og, f32, pts, row0=lo)
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t2 - t1)
    t_f, t_b = float(np.median(tf)), float(np.median(tb))
    t2 = time.perf_counter()
    f64 = O.filter_rows(og, O.weight(og, raw.astype(np.float64), f64=True), Hf, f64=True)
    o64 = O.backproject_points(og, f64, pts, f64=True, row0=lo)
    t3 = time.perf_counter()
    del f32, f64, raw
    upd = pts.shape[0] * g.Nang
    gups = upd / t_b / 1e9
    bp_full = g.voxel_updates / (gups * 1e9)
    filt_full = t_f * (g.Nv / rows)
    timing = {
        "gups": gups, "block": [x0, y0, z0, n], "detector_rows": [lo, hi],
        "voxel_updates": upd, "bp_seconds": t_b, "filter_seconds": t_f,
        "fp64_oracle_seconds": t3 - t2,
        "extrapolated_full_fdk_seconds": bp_full + filt_full,
        "extrapolated_full_bp_seconds": bp_full,
        "extrapolated_full_filter_seconds": filt_full,
    }
    errs = {"mirror_f32_vs_fp64": O.rel_rmse(m32, o64)}
    for st, vol in block_vols.items():
        sub = vol[:, y0:y0 + n, x0:x0 + n].reshape(-1)
        errs[f"gpu_{st}_vs_fp64_oracle_block"] = O.rel_rmse(sub, o64)
        if st == "f32":
            errs["gpu_f32_vs_mirror_block"] = O.rel_rmse(sub, m32)
    return timing, errs


def full_filter_stage(fbp, w, g, hipm, rows_per_slice=128):
    """Seconds the FP32 mirror takes to weight and filter every detector row of
    the resident projections, in slices of rows_per_slice detector rows (each
    copied to the host untimed, then weighted and filtered as a whole)."""
    import oracle as O
    Hf = O.filter_spectrum(w.filter, g.Nu, g.tau)
    buf = np.empty((rows_per_slice, g.Nang, g.Nu), np.float32)
    total = 0.0
    for lo in range(0, g.Nv, rows_per_slice):
        hi = min(g.Nv, lo + rows_per_slice)
        part = buf[:hi - lo]
        src = fbp.device_sinogram_ptr() + lo * g.Nang * g.Nu * 4
        assert hipm.hipMemcpy(part.ctypes.data, C.c_void_p(src), part.nbytes, 2) == 0
        og = O.Geometry(**g.rows_oracle_dict(lo, hi))
        t0 = time.perf_counter()
        f = O.filter_rows(og, O.weight(og, part), Hf)
        total += time.perf_counter() - t0
        del f
    return total


def cpu_leg(fbp, w, g, block_vols, hipm, threads):
    """CPU baseline: the FP32 OpenMP mirror (timed) and the FP64 oracle on a
    seeded 64^3 block of the workload (its detector rows only), the mirror's
    weight + filter of every detector row (the full filter stage), and the cfg1
    workload (128^3 from 180 x 128^2) reconstructed in full by the mirror."""
    import oracle as O
    block = geometry.sub_cube(g, w.seed)
    log(f"cpu leg: block {block}")
    bt, errs = block_check(fbp, w, g, block, block_vols, hipm, threads, 3)
    t_filter_full = full_filter_stage(fbp, w, g, hipm)
    log(f"cpu leg: full filter stage {t_filter_full:.1f} s")
    x0, y0, z0, n = block
    # cfg1 in full: the mirror's FDK of the Shepp-Logan projections
    w1 = geometry.workloads()["cfg1"]
    g1 = w1.geom
    o1 = O.Geometry(**g1.oracle_dict())
    s1 = O.project_ellipsoids(o1, O.shepp_logan(g1.fov_px / 2)).astype(np.float32)
    t4 = time.perf_counter()
    O.fdk(o1, s1, kind=0)
    t5 = time.perf_counter()
    out = {
        "value": bt["gups"], "unit": "GUPS", "cores": threads, "kind": "port",
        "nproc": os.cpu_count(), "cpu_model": cpu_model(),
        "sample": (f"FP32 OpenMP mirror (gcc -O3 -march=native -fopenmp, {threads} threads), "
                   f"backprojection of a seeded {n}^3 voxel block (x{x0} y{y0} z{z0}) over all "
                   f"{g.Nang} projections = {bt['voxel_updates']:.3e} voxel-updates, median of "
                   f"3; weight+filter of its {bt['detector_rows'][1] - bt['detector_rows'][0]} "
                   f"detector rows x {g.Nang} projections {bt['filter_seconds']:.2f} s; "
                   f"weight+filter of all {g.Nv} detector rows {t_filter_full:.1f} s"),
        "bp_seconds": bt["bp_seconds"], "filter_seconds": bt["filter_seconds"],
        "fp64_oracle_seconds": bt["fp64_oracle_seconds"],
        "full_filter_seconds": t_filter_full,
        # the backprojection extrapolated from the block, the filter stage measured
        "extrapolated_full_fdk_seconds": bt["extrapolated_full_bp_seconds"] + t_filter_full,
        "extrapolated_full_bp_seconds": bt["extrapolated_full_bp_seconds"],
        "extrapolated_full_filter_seconds": bt["extrapolated_full_filter_seconds"],
        "cfg1_full_fdk_seconds": t5 - t4,
        "cfg1_gups": g1.voxel_updates / (t5 - t4) / 1e9,
    }
    return out, errs, block


def fdk_steps(fbp, L, hipm, steps, warmup):
    """Mean stage times [filter, pack, BP] (ms) of `steps` FDK passes, and the
    mean wall time per pass."""
    nullcb = _lib.GET_BLOCK_FN()
    for _ in range(warmup):
        _lib.check(L.fdk_calculate_reconstruction(fbp.handle, nullcb, None), "warmup")
    hipm.hipDeviceSynchronize()
    walls, stages = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        _lib.check(L.fdk_calculate_reconstruction(fbp.handle, nullcb, None), "step")
        hipm.hipDeviceSynchronize()
        walls.append(time.perf_counter() - t0)
        stages.append(fbp.stage_times_ms())
    s = np.array(stages)
    return [float(np.mean(s[:, i])) for i in range(3)], 1e3 * float(np.mean(walls))


def config_leg(name, hipm, storages, steps, cpu_full=False, threads=None, cpu_block=False):
    """GUPS and FDK seconds of one workload per storage type, with each
    low-precision volume's relative RMSE against the FP32 volume (on the GPU).
    cpu_block: also every storage type's error against the FP64 oracle and the
    CPU FP32 mirror's time on a seeded voxel block (block_check, one run), and
    for the 1024^3 workloads the mirror's full filter stage (one run);
    cpu_full: also reconstruct the workload in full with the CPU FP32 mirror
    (one timed run) and compare the GPU FP32 volume with it."""
    L = _lib.load()
    w = geometry.workloads()[name]
    g = w.geom
    fbp = FilteredBackProjection(0)
    g.apply(fbp, None, None)
    _lib.check(fbp.SetFilter(w.filter), "SetFilter")
    _lib.check(fbp.SetStorage(_lib.F32), "SetStorage")
    _lib.check(fbp.Initialize(), "Initialize")
    ells = np.ascontiguousarray(w.ellipsoids())
    _lib.check(L.fdk_phantom_project(fbp.handle, ells.ctypes.data, ells.shape[0], 0, g.Nang,
                                     fbp.device_sinogram_ptr(), None), "phantom")
    if w.I0:
        _lib.check(L.fdk_poisson_log(fbp.device_sinogram_ptr(), g.Nu * g.Nv * g.Nang, w.I0,
                                     w.seed, None), "poisson")
    nvol = g.Nx * g.Ny * g.Nz
    ref = C.c_void_p()
    assert hipm.hipMalloc(C.byref(ref), nvol * 4) == 0
    out = {"workload": w.name, "voxel_updates": g.voxel_updates}
    block = geometry.sub_cube(g, w.seed, n=min(64, g.Nx, g.Ny, g.Nz))
    block_vols = {}
    for st in storages:
        _lib.check(fbp.SetStorage(st), "SetStorage")
        (f, pk, bp), wall = fdk_steps(fbp, L, hipm, steps, 1)
        dvol = C.c_void_p(fbp.device_volume_ptr())
        e = None
        if st == _lib.F32:
            assert hipm.hipMemcpy(ref, dvol, nvol * 4, 3) == 0
        elif _lib.F32 in storages:       # FP32 ran first: compare with its volume
            o3 = (C.c_double * 3)()
            _lib.check(L.fdk_compare(dvol, ref, nvol, o3, None), "compare")
            e = math.sqrt(o3[0] / o3[1])
        out[NAMES[st]] = {"gups": g.voxel_updates / (bp * 1e-3) / 1e9, "fdk_s": (f + pk + bp) / 1e3,
                          "filter_ms": f, "pack_ms": pk, "bp_ms": bp,
                          "rel_rmse_vs_f32_volume": e}
        if cpu_block:
            z0, n = block[2], block[3]
            slab = np.empty((n, g.Ny, g.Nx), np.float32)
            assert hipm.hipMemcpy(slab.ctypes.data, C.c_void_p(dvol.value + z0 * g.Nx * g.Ny * 4),
                                  slab.nbytes, 2) == 0
            block_vols[NAMES[st]] = slab
    if cpu_block:
        bt, errs = block_check(fbp, w, g, block, block_vols, hipm, threads, 1)
        if g.Nz >= 1024:
            bt["full_filter_seconds"] = full_filter_stage(fbp, w, g, hipm)
            bt["extrapolated_full_fdk_seconds"] = (bt["extrapolated_full_bp_seconds"]
                                                   + bt["full_filter_seconds"])
        out["cpu_block"] = bt
        out["rel_rmse_block"] = errs
        log(f"{name} block {block}: CPU mirror {bt['gups']:.2f} GUPS, errors " +
            ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    if cpu_full and _lib.F32 in storages:
        import oracle as O
        raw = np.empty((g.Nv, g.Nang, g.Nu), np.float32)
        assert hipm.hipMemcpy(raw.ctypes.data, C.c_void_p(fbp.device_sinogram_ptr()),
                              raw.nbytes, 2) == 0
        gpu = np.empty((g.Nz, g.Ny, g.Nx), np.float32)
        assert hipm.hipMemcpy(gpu.ctypes.data, ref, nvol * 4, 2) == 0
        t0 = time.perf_counter()
        mirror = O.fdk(O.Geometry(**g.oracle_dict()), raw, kind=w.filter)
        t = time.perf_counter() - t0
        out["cpu_full"] = {"fdk_seconds": t, "gups": g.voxel_updates / t / 1e9,
                           "threads": threads, "kind": "port",
                           "gpu_f32_vs_cpu_mirror_rel_rmse": O.rel_rmse(gpu, mirror)}
        log(f"{name} in full on the CPU ({threads} threads): {t:.1f} s "
            f"({out['cpu_full']['gups']:.2f} GUPS), GPU FP32 vs mirror "
            f"{out['cpu_full']['gpu_f32_vs_cpu_mirror_rel_rmse']:.1e}")
        del raw, gpu, mirror
    hipm.hipFree(ref)
    fbp.close()
    log(f"{name}: " + ", ".join(f"{k} {v['fdk_s'] * 1e3:.2f} ms ({v['gups']:.0f} GUPS)"
                                for k, v in out.items() if isinstance(v, dict) and "fdk_s" in v))
    return out


def toeplitz_leg():
    """The row filter as a Toeplitz GEMM on the matrix cores against the FFT, at
    1024^3 from 1800 x 2048^2: the FFT filter's time over the whole stack (one
    call, weight fused); the time of the whole GEMM (1800 x 2048 rows times the
    2048 x 2048 matrix) through torch.matmul; and, on the detector rows of the
    seeded central 64^3 block (the cfg3 data, weighted in FP32 by the engine),
    the relative error of the filtered rows against the engine's FP64 filter
    with FP32, FP16 and BF16 operands (scaled into FP16 range, products summed
    in FP32 by the GEMM; also with each row's mean taken out first), and the
    error each filter gives the block's FDK (packed in the matching storage
    type, backprojected by the engine) against the engine's FP64 FDK of the
    block.  The matrix holds the filter's own response: H[m][n] = k[n - m],
    read from the FP64 filter's output for unit impulses."""
    import torch
    from tomokvas_amd.stages import Stages
    w = geometry.workloads()["cfg3"]
    g = w.geom
    fbp = FilteredBackProjection(0)
    g.apply(fbp, None, None)
    _lib.check(fbp.SetFilter(w.filter), "SetFilter")
    st = Stages(fbp)
    Nu = st.Nu
    x0, y0, z0, n = geometry.sub_cube(g, w.seed)
    r0, r1 = geometry.detector_rows_of_block(g, x0, y0, z0, n)
    sino = st.phantom(w.ellipsoids())
    st.poisson_log(sino, w.I0, w.seed)
    raw = sino[r0:r1].clone()
    # the product's row filter (weight fused, FP32 FFT) over the whole stack
    st.filter(sino, weight=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st.filter(sino, weight=True)
    torch.cuda.synchronize()
    fft_ms = 1e3 * (time.perf_counter() - t0)
    del sino
    torch.cuda.empty_cache()
    # an engine for detector rows [r0, r1) (the float cZ less r0, an exact
    # difference) and a 64^3 volume whose voxels are the block's
    gb = geometry.workloads()["cfg3"].geom
    gb.Nv, gb.cZ = r1 - r0, float(np.float32(g.cZ)) - r0
    gb.Nx = gb.Ny = gb.Nz = n
    gb.cx, gb.cy = g.cx - x0, g.cy - y0
    gb.cz = (g.Nz - 1) / 2 - (n - 1) / 2 - z0
    fbp2 = FilteredBackProjection(0)
    gb.apply(fbp2, None, None)
    _lib.check(fbp2.SetFilter(w.filter), "SetFilter")
    st2 = Stages(fbp2)
    # the block's FDK in FP64 throughout (the engine's FP64 instantiation)
    f64 = st2.filter(st2.weight(raw.double()))
    s64, c64 = st2.pack(f64, _lib.F64)
    v_ref = st2.backproject(s64, _lib.F64, c64)
    del f64, s64, c64

    def block_err(rows32, storage):
        m = None
        if storage != _lib.F32:
            m = rows32.abs().amax(dim=(0, 2)).contiguous().view(torch.int32)
        stk, sc = st2.pack(rows32.contiguous(), storage, m)
        v = st2.backproject(stk, storage, sc).double()
        return float((v - v_ref).norm() / v_ref.norm())

    x = st2.weight(raw.clone())                          # the engine's FP32 weight
    del raw
    x64 = x.double()
    ref = st2.filter(x64.clone())                        # FP64 FFT filter
    fft32 = st2.filter(x.clone())                        # the product's FP32 FFT
    fft_rel = float((fft32.double() - ref).norm() / ref.norm())
    fdk = {"fft_f32": block_err(fft32, _lib.F32), "fft_f16": block_err(fft32, _lib.F16),
           "fft_bf16": block_err(fft32, _lib.BF16)}
    del fft32
    # impulse responses: row 0 from an impulse at u = 0 gives k[n], n >= 0;
    # from u = Nu-1, k[n - Nu + 1]
    imp = torch.zeros((r1 - r0, 1, Nu), dtype=torch.float64, device="cuda")
    imp[0, 0, 0] = 1.0
    imp[1, 0, Nu - 1] = 1.0
    resp = st2.filter(imp)
    lags = torch.cat([resp[1, 0, :Nu - 1], resp[0, 0, :]])   # k[-(Nu-1)] .. k[Nu-1]
    ni = torch.arange(Nu, device="cuda")
    H = lags[(ni[None, :] - ni[:, None]) + Nu - 1]           # H[m][n] = k[n - m]
    X2 = x64.reshape(-1, Nu)
    del x, x64
    out = {"workload": w.name, "rows": [r0, r1], "block": [x0, y0, z0, n],
           "fft_filter_ms": fft_ms, "fft_f32_rel": fft_rel,
           "fp64_gemm_rel": float(((X2 @ H).reshape(ref.shape) - ref).norm() / ref.norm())}
    sx, sh = float(X2.abs().max()), float(H.abs().max())
    mu = X2.mean(dim=1, keepdim=True)
    d = X2 - mu
    sd = float(d.abs().max())
    for name, dt, storage in (("fp32", torch.float32, _lib.F32),
                              ("fp16", torch.float16, _lib.F16),
                              ("bf16", torch.bfloat16, _lib.BF16)):
        Hs = (H / sh).to(dt)
        y = ((X2 / sx).to(dt) @ Hs).double().reshape(ref.shape) * (sx * sh)
        out[f"{name}_rel"] = float((y - ref).norm() / ref.norm())
        fdk[f"toeplitz_{name}"] = block_err(y.float(), storage)
        del y
        yd = (((d / sd).to(dt) @ Hs).double() * (sd * sh) + mu * H.sum(dim=0)[None, :])
        yd = yd.reshape(ref.shape)
        out[f"{name}_rel_mean_removed"] = float((yd - ref).norm() / ref.norm())
        fdk[f"toeplitz_{name}_mean_removed"] = block_err(yd.float(), storage)
        del yd, Hs
    out["fdk_block_rel_vs_fp64"] = fdk
    del X2, d, ref, v_ref
    torch.cuda.empty_cache()
    M = g.Nv * g.Nang
    for name, dt in (("fp16", torch.float16), ("bf16", torch.bfloat16), ("fp32", torch.float32)):
        chunk = 1 << 20
        A = torch.randn(chunk, Nu, device="cuda").to(dt)
        B = (H / sh).to(dt)
        C = torch.empty(chunk, Nu, device="cuda", dtype=dt)
        torch.matmul(A, B, out=C)
        torch.cuda.synchronize()
        nrep = (M + chunk - 1) // chunk
        t0 = time.perf_counter()
        for _ in range(nrep):
            torch.matmul(A, B, out=C)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) * M / (nrep * chunk)
        out[f"{name}_gemm_ms"] = 1e3 * t
        out[f"{name}_gemm_tflops"] = 2.0 * M * Nu * Nu / t / 1e12
        del A, B, C
    fbp2.close()
    fbp.close()
    return out


def diff_leg(name, steps_chunked):
    """One training step of the differentiable form (tomokvas_amd.diff): loss =
    sum of squares over the central 256^3 crop of the volume (all of it below
    256^3), gradients w.r.t. the projections and the p learnable taps.  Reports
    step time, peak device memory the step allocates above its inputs, and the
    gradient error of each variant against the FP32 step without chunks."""
    import torch
    from tomokvas_amd import diff
    from tomokvas_amd.stages import Stages
    w = geometry.workloads()[name]
    g = w.geom
    fbp = FilteredBackProjection(0)
    g.apply(fbp, None, None)
    st = Stages(fbp)
    sino = st.phantom(w.ellipsoids())
    taps = diff.ramlak_taps(st)
    diff.set_taps(st, taps)
    Nz, Ny, Nx = st.vol_shape
    c = [min(128, n // 2) for n in (Nz, Ny, Nx)]
    crop = tuple(slice(n // 2 - h, n // 2 + h) for n, h in zip((Nz, Ny, Nx), c))

    def step(storage, chunk):
        vol = diff.forward(st, sino, storage=storage, chunk=chunk)
        gvol = torch.zeros_like(vol)
        gvol[crop] = 2 * vol[crop]
        loss = float((vol[crop].double() ** 2).sum())
        del vol
        gp, dh = diff.backward(st, sino, gvol, chunk=chunk)
        return loss, gp, dh

    out = {"workload": w.name, "taps": st.p, "crop": [2 * h for h in c]}
    grads = {}
    for label, storage, chunk in steps_chunked:
        step(storage, chunk)                     # warm-up (code objects, allocator)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        t0 = time.perf_counter()
        loss, gp, dh = step(storage, chunk)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        peak = torch.cuda.max_memory_allocated() - base
        grads[label] = (gp, dh)
        out[label] = {"storage": NAMES[storage], "chunk": chunk, "step_s": dt,
                      "peak_alloc_gib": peak / 2**30,
                      "inputs_gib": sino.numel() * 4 / 2**30, "loss": loss}
    ref_gp, ref_dh = grads["f32_whole"]
    for label, (gp, dh) in grads.items():
        if label == "f32_whole":
            continue
        out[label]["grad_proj_rel_err_vs_f32_whole"] = float(
            torch.linalg.vector_norm((gp - ref_gp).double()) /
            torch.linalg.vector_norm(ref_gp.double()))
        out[label]["grad_taps_rel_err_vs_f32_whole"] = float(
            np.linalg.norm(dh - ref_dh) / np.linalg.norm(ref_dh))
    del grads, sino
    fbp.close()
    torch.cuda.empty_cache()
    log(f"{name} differentiation: " + ", ".join(
        f"{k} {v['step_s']:.3f} s / {v['peak_alloc_gib']:.1f} GiB"
        for k, v in out.items() if isinstance(v, dict)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--storages", default="f32,f16,bf16,f16-packed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="leave out the other workloads' GUPS/FDK legs")
    ap.add_argument("--no-diff", action="store_true",
                    help="leave out the differentiation-form training steps")
    ap.add_argument("--rsq", action="store_true",
                    help="also time the RSQ preprocessing (248 sinograms of 2000 x 1024)")
    ap.add_argument("--tiff", metavar="DIR",
                    help="also export the FP16-store volume as TIFF slices into DIR")
    ap.add_argument("--no-end-to-end", action="store_true",
                    help="leave out the runs that start from projections in host memory")
    ap.add_argument("--no-slab-pipeline", action="store_true",
                    help="leave out the run with every slab handed to host memory (its chunked "
                         "backprojection launches would mix into a kernel trace's averages)")
    # the host CPU share of one GPU on the MI355X machines this runs on is 16
    # CPUs (nproc may report the whole machine)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--calibrate", action="store_true",
                    help="after the timed steps, run accumulate_kernel once over 2^28 floats "
                         "(reads 2 GiB, writes 1 GiB): the reference launch for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE readings to bytes")
    ap.add_argument("--diff-leg", metavar="WORKLOAD",
                    help="only run the differentiation-form training steps of WORKLOAD "
                         "(cfg4 or cfg5) and print their JSON")
    ap.add_argument("--toeplitz-leg", action="store_true",
                    help="only run the Toeplitz-GEMM row filter comparison and print its JSON")
    ap.add_argument("--no-toeplitz", action="store_true",
                    help="leave out the Toeplitz-GEMM row filter comparison")
    args = ap.parse_args()
    if args.toeplitz_leg:
        import torch
        assert torch.cuda.is_available(), "the Toeplitz leg needs a HIP device"
        print(json.dumps(toeplitz_leg()), flush=True)
        return
    if args.diff_leg:
        import torch
        assert torch.cuda.is_available(), "the differentiation leg needs a HIP device"
        print(json.dumps(diff_leg(args.diff_leg, DIFF_STEPS[args.diff_leg])), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the FDK path "
              "is not sharded across GPUs in this build", file=sys.stderr)
        sys.exit(2)

    w = geometry.workloads()[args.workload]
    g = w.geom
    hipm = hip()
    fbp = FilteredBackProjection(0)
    g.apply(fbp, None, None)
    _lib.check(fbp.SetFilter(w.filter), "SetFilter")
    _lib.check(fbp.SetStorage(_lib.F32), "SetStorage")
    t0 = time.perf_counter()
    _lib.check(fbp.Initialize(), "Initialize")
    log(f"initialised {w.name}: {time.perf_counter() - t0:.2f} s")

    L = _lib.load()
    ells = np.ascontiguousarray(w.ellipsoids())
    t0 = time.perf_counter()
    _lib.check(L.fdk_phantom_project(fbp.handle, ells.ctypes.data, ells.shape[0], 0, g.Nang,
                                     fbp.device_sinogram_ptr(), None), "phantom")
    if w.I0:
        _lib.check(L.fdk_poisson_log(fbp.device_sinogram_ptr(), g.Nu * g.Nv * g.Nang, w.I0,
                                     w.seed, None), "poisson")
    hipm.hipDeviceSynchronize()
    log(f"synthetic projections on the GPU: {time.perf_counter() - t0:.2f} s")

    nvol = g.Nx * g.Ny * g.Nz
    ref = C.c_void_p()
    assert hipm.hipMalloc(C.byref(ref), nvol * 4) == 0
    order = [CODES[s] for s in args.storages.split(",")]
    if _lib.F32 in order:
        order.remove(_lib.F32)
        order.insert(0, _lib.F32)
    res = {}
    block = geometry.sub_cube(g, w.seed)
    block_vols = {}
    for st in order:
        _lib.check(fbp.SetStorage(st), "SetStorage")
        (filt_ms, pack_ms, bp_ms), wall_ms = fdk_steps(fbp, L, hipm, args.steps, args.warmup)
        fdk_ms = filt_ms + pack_ms + bp_ms
        dvol = fbp.device_volume_ptr()
        cmp = None
        if st == _lib.F32:
            assert hipm.hipMemcpy(ref, C.c_void_p(dvol), nvol * 4, 3) == 0
        elif _lib.F32 in order:           # FP32 ran first: compare with its volume
            out3 = (C.c_double * 3)()
            _lib.check(L.fdk_compare(C.c_void_p(dvol), ref, nvol, out3, None), "compare")
            cmp = math.sqrt(out3[0] / out3[1])
        x0, y0, z0, n = block
        slab = np.empty((n, g.Ny, g.Nx), np.float32)
        assert hipm.hipMemcpy(slab.ctypes.data, C.c_void_p(dvol + z0 * g.Nx * g.Ny * 4),
                              slab.nbytes, 2) == 0
        block_vols[NAMES[st]] = slab
        res[NAMES[st]] = dict(filter_ms=filt_ms, pack_ms=pack_ms, bp_ms=bp_ms, fdk_ms=fdk_ms,
                              wall_ms=wall_ms, gups=g.voxel_updates / (bp_ms * 1e-3) / 1e9,
                              rel_rmse_vs_f32=cmp)
        log(f"{NAMES[st]}: filter {filt_ms:.2f} ms, pack {pack_ms:.2f} ms, BP {bp_ms:.2f} ms "
            f"({res[NAMES[st]]['gups']:.1f} GUPS), FDK {fdk_ms:.2f} ms, vs f32 {cmp}")
    hipm.hipFree(ref)
    pipe = None
    if not args.no_slab_pipeline:
        pipe_storage = _lib.F16 if "f16" in res else order[0]
        pipe = slab_pipeline(fbp, g, hipm, pipe_storage, args.steps)
        r0 = res[NAMES[pipe_storage]]
        pipe["fdk_without_handout_ms"] = r0["fdk_ms"]
        pipe["fdk_with_handout_ms"] = pipe["filter_ms"] + pipe["pack_ms"] + pipe["bp_ms"] + \
            pipe["exposed_transfer_ms"]
        log(f"slab hand-out ({pipe['storage']}, chunks of {pipe['chunk_slices']} slices): BP "
            f"{pipe['bp_ms']:.2f} ms, exposed transfers {pipe['exposed_transfer_ms']:.2f} ms, "
            f"wall {pipe['wall_ms']:.1f} ms, slabs equal device volume: "
            f"{pipe['slabs_equal_device_volume']}")

    # the headline: FP16 storage with FP32 interpolation weights and accumulation
    head = "f16" if "f16" in res else NAMES[order[0]]
    r = res[head]
    upd = g.voxel_updates
    s_bytes = CELL_BYTES[CODES[head]]
    gather = 4 * s_bytes * upd                      # 2x2 taps per voxel-update
    gather_gbs = gather / (r["bp_ms"] * 1e-3) / 1e9
    traffic, valu_per_update = read_profile(args.workload, head)
    t_bp = r["bp_ms"] * 1e-3
    nlaunch = C.c_int()
    _lib.check(L.fdk_get_angle_chunks(fbp.handle, C.byref(nlaunch)), "angle chunks")
    # one step backprojects the angle loop in K launches of Nang/K angles each:
    # per-launch figures are the step's divided by K; the traffic profile holds
    # the step's bytes (its K dispatches summed)
    K = nlaunch.value
    roofline = {
        "kernel": f"bp_kernel<{head}>", "launches_per_step": K, "avg_launch_ms": r["bp_ms"] / K,
        "traffic": traffic / K if traffic else None,
        "algorithmic_bytes": gather / K,
        "compulsory_hbm_bytes_per_step": g.Nang * g.Nv * g.Nu * s_bytes + 4 * g.Nx * g.Ny * g.Nz,
    }
    if valu_per_update:
        # the gathers are served from LDS and the kernel is bound by VALU issue:
        # achieved = lane-instructions per second (counter-measured instructions
        # per voxel-update x updates / BP time), peak = 256 CUs x 64 lanes x 2.4 GHz
        ach = upd * valu_per_update / t_bp
        roofline.update({"bound": "valu", "achieved": ach / 1e12, "peak": VALU_LANE_OPS / 1e12,
                         "unit": "T lane-instr/s", "frac": ach / VALU_LANE_OPS,
                         "valu_instr_per_update": valu_per_update})
    else:
        roofline.update({"bound": "valu", "achieved": None, "peak": VALU_LANE_OPS / 1e12,
                         "unit": "T lane-instr/s", "frac": None})
    # HBM: the counter-measured bytes against the datasheet peak; L2: the
    # algorithmic 2x2-gather bytes (4 samples per update) against the L2 figure,
    # datasheet estimate and measured on this device
    roofline["hbm_gbs"] = traffic / t_bp / 1e9 if traffic else None
    roofline["hbm_frac"] = traffic / t_bp / 1e9 / HBM_PEAK_GBS if traffic else None
    roofline["gather_gbs"] = gather_gbs
    roofline["l2_peak_est_gbs"] = L2_PEAK_GBS
    roofline["frac_l2_est"] = gather_gbs / L2_PEAK_GBS
    l2_gbs = C.c_double(0.0)
    if L.fdk_probe_l2_read(2 << 20, C.byref(l2_gbs)) == 0 and l2_gbs.value > 0:
        roofline["l2_measured_gbs"] = l2_gbs.value
        roofline["frac_l2_measured"] = gather_gbs / l2_gbs.value
    else:
        log(f"L2 probe failed: {_lib.last_error()}")
    f32ms = res["f32"]["fdk_ms"] if "f32" in res else None
    out = {
        "metric": METRIC, "value": r["gups"], "unit": "GUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["fdk_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": f"{head}-store/f32-interp/f32-accumulate" if head != "f16-packed"
                 else "f16-store/f16-interp/f32-accumulate",
        "data": "synthetic: 64 seeded random ellipsoids, exact line integrals + Poisson(I0=1e5) "
                "counts, generated on the GPU",
        "config": w.describe(),
        "fdk_seconds": {k: v["fdk_ms"] / 1e3 for k, v in res.items()},
        "bp_gups": {k: v["gups"] for k, v in res.items()},
        "stage_ms": res,
        "speedup_fdk_f16_vs_f32": (f32ms / res["f16"]["fdk_ms"]
                                   if "f16" in res and f32ms else None),
        "speedup_fdk_bf16_vs_f32": (f32ms / res["bf16"]["fdk_ms"]
                                    if "bf16" in res and f32ms else None),
        # opt-in, outside the FP32-interpolation precision form
        "speedup_fdk_f16_packed_vs_f32": (f32ms / res["f16-packed"]["fdk_ms"]
                                          if "f16-packed" in res and f32ms else None),
        "rel_rmse": {f"{k}_vs_f32_full_volume": v["rel_rmse_vs_f32"] for k, v in res.items()
                     if v["rel_rmse_vs_f32"] is not None},
        "roofline": roofline,
        "slab_pipeline": pipe,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        cpu, errs, blk = cpu_leg(fbp, w, g, block_vols, hipm, args.threads)
        out["cpu_baseline"] = cpu
        out["rel_rmse"].update(errs)
        log(f"cpu leg {time.perf_counter() - t0:.1f} s: {cpu['value']:.3f} GUPS, errors {errs}")
    if args.tiff:
        _lib.check(fbp.SetStorage(_lib.F16 if "f16" in res else order[0]), "SetStorage")
        _lib.check(L.fdk_calculate_reconstruction(fbp.handle, _lib.GET_BLOCK_FN(), None), "tiff fdk")
        out["tiff"] = tl = tiff_leg(fbp, g, hipm, args.tiff)
        log(f"tiff: minmax {tl['minmax_ms']:.2f} ms ({tl['minmax_gbs']:.0f} GB/s), 16-bit "
            f"scaling {tl['to_u16_ms']:.2f} ms ({tl['to_u16_gbs']:.0f} GB/s), {tl['slices']} "
            f"files in {tl['export_s']:.2f} s")
    if args.rsq:
        out["rsq"] = rq = rsq_leg(hipm, args.steps)
        log(f"rsq: {rq['sinograms']} sinograms in {rq['ms']:.2f} ms "
            f"({rq['algorithmic_gbs']:.0f} GB/s, {100 * rq['frac_hbm']:.0f} % of HBM)")
    if not args.no_end_to_end:
        e2e_storage = _lib.F16 if "f16" in res else order[0]
        t0 = time.perf_counter()
        out["end_to_end"] = e2e = end_to_end(fbp, g, hipm, e2e_storage, args.steps)
        log(f"end to end from host memory ({e2e['storage']}): upload then FDK "
            f"{e2e['upload_then_fdk_s']:.3f} s, overlapped {e2e['overlapped_s']:.3f} s "
            f"({e2e['end_to_end_gups']:.0f} GUPS), rel RMSE {e2e['rel_rmse_overlapped_vs_serial']:.1e}"
            f"; leg {time.perf_counter() - t0:.1f} s")
    if args.calibrate:
        n = 1 << 28
        dst = fbp.device_sinogram_ptr()
        assert g.Nu * g.Nv * g.Nang >= 3 * n, "calibration needs 3 GiB of projection buffer"
        _lib.check(L.fdk_accumulate(C.c_void_p(dst), C.c_void_p(dst + 4 * n), n, None),
                   "calibration")
        hipm.hipDeviceSynchronize()
        out["calibration"] = {"kernel": "accumulate_kernel", "floats": n,
                              "read_bytes": 8 * n, "write_bytes": 4 * n}
    fbp.close()
    if not args.no_configs:
        cpu2 = out["cpu_baseline"] is not None
        out["configs"] = {name: config_leg(name, hipm, order, min(args.steps, 3),
                                           cpu_full=cpu2 and name == "cfg2", threads=args.threads,
                                           cpu_block=cpu2)
                          for name in ("cfg1", "cfg2", "cfg4", "cfg5") if name != args.workload}
        c2 = out["configs"].get("cfg2", {}).get("cpu_full")
        if c2:
            # the CPU baseline's second full reconstruction: 512^3 from 900 x 1024^2
            out["cpu_baseline"]["cfg2_full_fdk_seconds"] = c2["fdk_seconds"]
            out["cpu_baseline"]["cfg2_gups"] = c2["gups"]
    if not args.no_diff:
        # the training steps run on torch tensors: in a child process that
        # initialises torch before anything else touches the GPU
        out["differentiation"] = {}
        for name in ("cfg4", "cfg5"):
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--diff-leg", name],
                               capture_output=True, text=True, timeout=300)
            sys.stderr.write(r.stderr)
            if r.returncode != 0:
                raise RuntimeError(f"differentiation leg {name} failed ({r.returncode})")
            out["differentiation"][name] = json.loads(r.stdout.strip().splitlines()[-1])
    if not args.no_toeplitz:
        # the matrix-core (Toeplitz GEMM) alternative to the FFT row filter, measured (torch
        # in a child process, as above)
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--toeplitz-leg"],
                           capture_output=True, text=True, timeout=300)
        sys.stderr.write(r.stderr)
        if r.returncode != 0:
            raise RuntimeError(f"Toeplitz leg failed ({r.returncode})")
        tz = json.loads(r.stdout.strip().splitlines()[-1])
        out["filter_alternatives"] = tz
        log(f"Toeplitz GEMM row filter: FP16 {tz['fp16_gemm_ms']:.1f} ms "
            f"({tz['fp16_gemm_tflops']:.0f} TFLOP/s), rel error {tz['fp16_rel']:.1e}; BF16 "
            f"{tz['bf16_gemm_ms']:.1f} ms, {tz['bf16_rel']:.1e}; FP32 {tz['fp32_gemm_ms']:.1f} ms, "
            f"{tz['fp32_rel']:.1e}; FFT filter {tz['fft_filter_ms']:.1f} ms")
        log("  64^3 block FDK vs FP64: " + ", ".join(
            f"{k} {v:.1e}" for k, v in tz["fdk_block_rel_vs_fp64"].items()))
    print(json.dumps(out), flush=True)


DIFF_STEPS = {
    "cfg4": [("f32_whole", _lib.F32, None), ("f32_chunk50", _lib.F32, 50)],
    "cfg5": [("f32_whole", _lib.F32, None), ("f16_chunk50", _lib.F16, 50)],
}


if __name__ == "__main__":
    main()
