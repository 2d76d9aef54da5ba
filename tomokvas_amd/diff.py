"""The differentiable FDK form on torch device tensors: forward and backward
passes built from the stage-level calls of the C ABI (no autograd here; a
torch.autograd.Function wraps them in the tests).

    x = BP( F_h( W p ) )

W: cosine weight (diagonal, FBPG/kernel.cu:179-201); F_h: row filter with p
learnable spatial taps h, spectrum H = DFT_p(h) (the reference's fConvolution,
kernel.cu:557-781, with the Ram-Lak taps of kernel.cu:783-802 replaced by h);
BP: the backprojector (kernel.cu:68-177).  The adjoints are hand-written:

    dL/dp = W F_h^T BP^T g        fdk_forward_project, fdk_filter(adjoint=1), fdk_weight
    dL/dh = d<BP^T g, F_h q>/dh   fdk_tap_gradient, q = W p

With ``chunk`` set, neither pass holds a filtered or weighted copy of the
whole stack: the forward pass filters and packs ``chunk`` projections at a time
into the stored stack, and the backward pass recomputes q = W p per chunk
(per-angle-chunk recomputation).  The backward pass forward-projects all angles
at once straight into the gradient buffer (one read of the volume gradient).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .stages import Stages


def _torch():
    import torch
    return torch


def set_taps(st: Stages, taps) -> None:
    """Makes the context's filter the spectrum of the p spatial taps."""
    t = np.ascontiguousarray(taps, np.float64)
    _lib.check(st.fbp.SetFilter(_lib.FILTER_USER, t), "SetFilter")
    st.fbp.prepare()


def ramlak_taps(st: Stages) -> np.ndarray:
    """Spatial taps whose DFT is the reference's Ram-Lak spectrum tau*|DFT(t)|
    (kernel.cu:570-655, 783-802): t[i] = RamLak(i - floor(Nu/2))/2 for
    i < Nu, zero to length p; the real, even inverse DFT of the spectrum."""
    g = st.g
    Nu, tau, p = g["Nu"], float(g["tau"]), st.p
    n = np.arange(Nu) - Nu // 2
    t = np.where(n == 0, 1.0 / (4 * tau * tau),
                 np.where(n % 2 == 0, 0.0, -1.0 / (np.maximum(n * n, 1) * np.pi ** 2 * tau * tau)))
    taps = np.zeros(p)
    taps[:Nu] = t / 2
    H = tau * np.abs(np.fft.fft(taps))
    return np.real(np.fft.ifft(H))


def forward(st: Stages, proj, storage=_lib.F32, chunk=None):
    """FDK of proj [Nv][Nang][Nu] with the context's filter; returns the volume.
    FP64 input uses FP64 throughout; FP32 input is stored as `storage`."""
    torch = _torch()
    nang = proj.shape[1]
    chunk = chunk or nang
    store = _lib.F64 if proj.dtype == torch.float64 else storage
    stack = torch.empty(st.stack_bytes(store), dtype=torch.uint8, device=proj.device)
    scales = torch.empty(nang, dtype=torch.float64 if store == _lib.F64 else torch.float32,
                         device=proj.device)
    for a0 in range(0, nang, chunk):
        n = min(chunk, nang - a0)
        y = proj[:, a0:a0 + n, :].contiguous() if n < nang else proj.detach().clone()
        maxabs = torch.zeros(n, dtype=torch.int32, device=proj.device)
        st.filter(y, weight=True, maxabs=maxabs)      # W then F_h, per-projection max |y|
        st.pack_into(y, store, stack, scales, a0, maxabs)
        del y
    return st.backproject(stack, store, scales)


def backward(st: Stages, proj, gvol, chunk=None, need_proj=True, need_taps=True):
    """Gradients of <gvol, x> w.r.t. proj and the taps: (dL/dp tensor or None,
    dL/dh numpy float64 [p] or None)."""
    torch = _torch()
    nang = proj.shape[1]
    chunk = chunk or nang
    gvol = gvol.contiguous().to(proj.dtype)
    gproj = st.forward_project(gvol)                   # BP^T g for every angle
    dh = None
    if need_taps:
        dh = np.zeros(st.p, np.float64)
        for a0 in range(0, nang, chunk):
            n = min(chunk, nang - a0)
            if n == nang:
                q = st.weight(proj.detach().clone())
                gy = gproj
            else:
                q = st.weight(proj[:, a0:a0 + n, :].contiguous())
                gy = gproj[:, a0:a0 + n, :].contiguous()
            st.tap_gradient(q, gy, dh)
            del q, gy
    if not need_proj:
        return None, dh
    st.filter(gproj, adjoint=True)                     # F_h^T, in place
    st.weight(gproj)                                   # W^T = W
    return gproj, dh
