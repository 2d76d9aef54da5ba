"""torch.autograd wrapper of the differentiable FDK form -- test harness only.

The passes themselves are tomokvas_amd.diff.forward / backward (stage-level C
ABI calls with the hand-written adjoints); this Function only connects them to
autograd.  With ``chunk`` set, neither pass holds a filtered or weighted copy
of the whole stack (per-angle-chunk recomputation).
"""
from __future__ import annotations

import numpy as np
import torch

from tomokvas_amd import _lib, diff


class FDKFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj, taps, st, storage, chunk):
        torch.cuda.synchronize()
        diff.set_taps(st, taps.detach().cpu().numpy())
        vol = diff.forward(st, proj.detach(), storage=storage, chunk=chunk)
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
        diff.set_taps(st, ctx.taps)
        gproj, dh = diff.backward(st, proj, gvol, chunk=ctx.chunk,
                                  need_proj=ctx.needs_input_grad[0],
                                  need_taps=ctx.needs_input_grad[1])
        gtaps = torch.from_numpy(dh).to(proj.device, proj.dtype) if dh is not None else None
        return gproj, gtaps, None, None, None


def fdk(proj, taps, st, storage=_lib.F32, chunk=None):
    return FDKFunction.apply(proj, taps, st, storage, chunk)


def ramlak_taps(st) -> np.ndarray:
    return diff.ramlak_taps(st)
