"""Stage-level calls of the C ABI on torch device tensors.

torch provides device memory and the stream; every computation is a kernel of
libtomokvas_fdk.so, launched on torch's current stream so it orders with the
caller's torch work.  Sinogram tensors are [Nv][nang][Nu], volumes [Nz][Ny][Nx].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import F16, F32, BF16, F64, F16_PACKED  # noqa: F401


def _torch():
    import torch
    return torch


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _dtype_code(t) -> int:
    torch = _torch()
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.float64:
        return F64
    raise TypeError(f"unsupported dtype {t.dtype}")


def _check_dev(t, shape=None):
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")


class Stages:
    """Stage-level operations bound to a prepared engine context."""

    def __init__(self, fbp):
        self.fbp = fbp
        self.L = _lib.load()
        fbp.prepare()
        g = fbp.geometry()
        self.g = g
        self.Nu, self.Nv, self.Nang, self.p = g["Nu"], g["Nv"], g["Nang"], g["p"]
        self.vol_shape = (g["Nz"], g["Ny"], g["Nx"])

    def sino_shape(self, nang=None):
        return (self.Nv, self.Nang if nang is None else nang, self.Nu)

    def weight(self, sino):
        _check_dev(sino)
        nang = sino.shape[1]
        _lib.check(self.L.fdk_weight(self.fbp.handle, _dtype_code(sino), sino.data_ptr(), nang,
                                     _stream()), "fdk_weight")
        return sino

    def filter(self, sino, weight=False, adjoint=False, maxabs=None):
        """In place; returns the tensor.  maxabs: optional int32 tensor [nang]."""
        _check_dev(sino)
        nang = sino.shape[1]
        mp = None
        if maxabs is not None:
            _check_dev(maxabs, (nang,))
            mp = maxabs.data_ptr()
        _lib.check(self.L.fdk_filter(self.fbp.handle, _dtype_code(sino), sino.data_ptr(), nang,
                                     int(weight), int(adjoint), mp, _stream()), "fdk_filter")
        return sino

    def stack_bytes(self, store_dtype, nang=None):
        return self.L.fdk_stack_bytes(self.fbp.handle, store_dtype,
                                      self.Nang if nang is None else nang)

    def pack(self, filtered, store_dtype, maxabs=None):
        """Returns (stored stack as a uint8 tensor, scales tensor)."""
        torch = _torch()
        _check_dev(filtered)
        nang = filtered.shape[1]
        nbytes = self.stack_bytes(store_dtype, nang)
        stack = torch.empty(nbytes, dtype=torch.uint8, device=filtered.device)
        sdt = torch.float64 if store_dtype == F64 else torch.float32
        scales = torch.empty(nang, dtype=sdt, device=filtered.device)
        mp = maxabs.data_ptr() if maxabs is not None else None
        _lib.check(self.L.fdk_pack(self.fbp.handle, _dtype_code(filtered), filtered.data_ptr(),
                                   nang, mp, store_dtype, stack.data_ptr(), scales.data_ptr(),
                                   _stream()), "fdk_pack")
        return stack, scales

    def pack_into(self, filtered, store_dtype, stack, scales, ang0, maxabs=None):
        """Packs filtered [Nv][n][Nu] into projections [ang0, ang0 + n) of a
        stored stack of stack_bytes(store_dtype) bytes and its scales."""
        _check_dev(filtered)
        n = filtered.shape[1]
        per = self.stack_bytes(store_dtype, 1) - self.stack_bytes(store_dtype, 0)
        if ang0 < 0 or ang0 + n > scales.numel() or \
                stack.numel() < self.stack_bytes(store_dtype, ang0 + n):
            raise ValueError("pack_into: projections outside the stack")
        mp = maxabs.data_ptr() if maxabs is not None else None
        _lib.check(self.L.fdk_pack(self.fbp.handle, _dtype_code(filtered), filtered.data_ptr(),
                                   n, mp, store_dtype, stack.data_ptr() + ang0 * per,
                                   scales.data_ptr() + ang0 * scales.element_size(),
                                   _stream()), "fdk_pack")

    def backproject(self, stack, store_dtype, scales=None, ang0=0, nang=None, vol=None,
                    accumulate=False):
        torch = _torch()
        nang = self.Nang - ang0 if nang is None else nang
        vdt = torch.float64 if store_dtype == F64 else torch.float32
        if vol is None:
            vol = torch.zeros(self.vol_shape, dtype=vdt, device=stack.device)
        _check_dev(vol, self.vol_shape)
        sp = scales.data_ptr() if scales is not None else None
        _lib.check(self.L.fdk_backproject(self.fbp.handle, store_dtype, stack.data_ptr(), sp, ang0,
                                          nang, vol.data_ptr(), int(accumulate), _stream()),
                   "fdk_backproject")
        return vol

    def forward_project(self, vol, ang0=0, nang=None, out=None):
        torch = _torch()
        nang = self.Nang - ang0 if nang is None else nang
        _check_dev(vol, self.vol_shape)
        if out is None:
            out = torch.zeros(self.sino_shape(nang), dtype=vol.dtype, device=vol.device)
        _check_dev(out, self.sino_shape(nang))
        _lib.check(self.L.fdk_forward_project(self.fbp.handle, _dtype_code(vol), vol.data_ptr(),
                                              ang0, nang, out.data_ptr(), _stream()),
                   "fdk_forward_project")
        return out

    def tap_gradient(self, q, G, dh=None):
        """Accumulates d<G, filter_h(q)>/dh into dh (numpy float64 [p])."""
        _check_dev(q)
        _check_dev(G, tuple(q.shape))
        if dh is None:
            dh = np.zeros(self.p, np.float64)
        _lib.check(self.L.fdk_tap_gradient(self.fbp.handle, _dtype_code(q), q.data_ptr(),
                                           G.data_ptr(), q.shape[1], dh.ctypes.data, _stream()),
                   "fdk_tap_gradient")
        return dh

    def phantom(self, ells: np.ndarray, ang0=0, nang=None, out=None, device="cuda"):
        torch = _torch()
        nang = self.Nang - ang0 if nang is None else nang
        e = np.ascontiguousarray(ells, np.float64).reshape(-1, 19)
        if out is None:
            out = torch.empty(self.sino_shape(nang), dtype=torch.float32, device=device)
        _lib.check(self.L.fdk_phantom_project(self.fbp.handle, e.ctypes.data, e.shape[0], ang0,
                                              nang, out.data_ptr(), _stream()),
                   "fdk_phantom_project")
        return out

    def poisson_log(self, sino, I0, seed):
        _check_dev(sino)
        _lib.check(self.L.fdk_poisson_log(sino.data_ptr(), sino.numel(), float(I0), int(seed),
                                          _stream()), "fdk_poisson_log")
        return sino


def accumulate(dst, src):
    _check_dev(dst)
    _check_dev(src, tuple(dst.shape))
    _lib.check(_lib.load().fdk_accumulate(dst.data_ptr(), src.data_ptr(), dst.numel(),
                                          _stream()), "fdk_accumulate")
    return dst
