"""torch.autograd wrapper of the differentiable FDK form -- test harness only.

x = BP( F_h( W p ) ): cosine weight W (diagonal), row filter F_h with learnable
spatial taps h (H = DFT_p(h)), backprojection BP.  Gradients use the
hand-written adjoints of the C ABI:
    dL/dp = W F_h^T BP^T g        (fdk_forward_project, fdk_filter adjoint=1, fdk_weight)
    dL/dh = d<BP^T g, F_h q>/dh   (fdk_tap_gradient, q = W p)
With ``chunk`` set, the backward pass recomputes q per block of angles and
never holds the whole filtered or weighted stack (per-angle-chunk
recomputation).
"""
from __future__ import annotations

import numpy as np
import torch

from tomokvas_amd import _lib


class FDKFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj, taps, st, storage, chunk):
        f64 = proj.dtype == torch.float64
        torch.cuda.synchronize()
        _lib.check(st.fbp.SetFilter(_lib.FILTER_USER, taps.detach().cpu().numpy()), "SetFilter")
        st.fbp.prepare()
        y = proj.detach().clone()
        st.filter(y, weight=True)
        if f64:
            stack, scales = st.pack(y, _lib.F64)
            vol = st.backproject(stack, _lib.F64, scales)
        else:
            maxabs = None
            if storage == _lib.F16:
                maxabs = torch.from_numpy(
                    y.abs().amax(dim=(0, 2)).cpu().numpy().astype(np.float32).view(np.int32)).cuda()
            stack, scales = st.pack(y, storage, maxabs)
            vol = st.backproject(stack, storage, scales)
        del y, stack
        ctx.st = st
        ctx.chunk = chunk
        ctx.taps = taps.detach().cpu().numpy().copy()
        ctx.save_for_backward(proj)
        return vol

    @staticmethod
    def backward(ctx, gvol):
        (proj,) = ctx.saved_tensors
        st = ctx.st
        # the context spectrum may have been replaced by other forward calls
        torch.cuda.synchronize()
        _lib.check(st.fbp.SetFilter(_lib.FILTER_USER, ctx.taps), "SetFilter")
        st.fbp.prepare()
        gvol = gvol.contiguous().to(proj.dtype)
        nang = proj.shape[1]
        chunk = ctx.chunk or nang
        gproj = torch.empty_like(proj) if ctx.needs_input_grad[0] else None
        dh = np.zeros(st.p, np.float64)
        for a0 in range(0, nang, chunk):
            n = min(chunk, nang - a0)
            gy = st.forward_project(gvol, ang0=a0, nang=n)              # BP^T
            if ctx.needs_input_grad[1]:
                q = st.weight(proj[:, a0:a0 + n, :].clone(memory_format=torch.contiguous_format))
                st.tap_gradient(q, gy, dh)
                del q
            if gproj is not None:
                gq = st.filter(gy, adjoint=True)                        # F_h^T
                gproj[:, a0:a0 + n, :] = st.weight(gq)                  # W^T = W
            del gy
        gtaps = torch.from_numpy(dh).to(proj.device, proj.dtype) if ctx.needs_input_grad[1] else None
        return gproj, gtaps, None, None, None


def fdk(proj, taps, st, storage=_lib.F32, chunk=None):
    return FDKFunction.apply(proj, taps, st, storage, chunk)


def ramlak_taps(st) -> np.ndarray:
    """Spatial taps whose DFT is the reference Ram-Lak spectrum tau*|DFT(t)|:
    the real, even inverse DFT of the evenly extended spectrum."""
    import oracle as O
    g = st.g
    Hf = O.filter_spectrum(0, g["Nu"], g["tau"])
    p = st.p
    full = np.concatenate([Hf, Hf[1:p // 2][::-1]])
    return np.real(np.fft.ifft(full))
