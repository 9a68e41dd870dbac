"""torch.autograd.Function around the FP32-split forward and the hand-written adjoint kernel.

    I = intensity(pos, phase, amp=None, Ndet=..., pixelsize=..., detz=..., k=..., settings="single nf")
    loss = ((I - I_target) ** 2).sum(); loss.backward()

    loss = intensity_l2_loss(pos, phase, amp=None, target=I_target, Ndet=..., ...)
    loss.backward()          # the same loss and gradients with no torch kernels in between

pos (N, 3) and phase (N,) (and optionally amp (N,)) are CUDA tensors; the
returned intensity is float32 (N0, N1).  The backward pass calls
idisim_adjoint, which computes dL/dpos, dL/dphase and dL/damp from dL/dI and
the saved complex64 field (no reference counterpart).  intensity_l2_loss fuses the
squared-intensity loss: one call (idisim_loss_grad_l2: the forward kernel and the adjoint pass,
which forms dL/dI = 2 (I - T) itself from the forward's partial sums) gives the loss and the
gradients; the target is a constant.
"""
from __future__ import annotations

import torch

from .. import _lib
from .hip import get_handle


class _WavefieldIntensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, phase, amp, handle, extent):
        emit = torch.cat([pos.detach().to(torch.float64), phase.detach().to(torch.float64).reshape(-1, 1)],
                         1).contiguous()
        a = None if amp is None else amp.detach().to(torch.float64).contiguous()
        field = handle.forward(emit, a, None, _lib.OUT_FIELD, extent)
        ctx.handle = handle
        ctx.extent = extent
        ctx.has_amp = amp is not None
        ctx.dtypes = (pos.dtype, phase.dtype, None if amp is None else amp.dtype)
        ctx.save_for_backward(emit, field, a if a is not None else torch.empty(0, device=pos.device))
        return field.real * field.real + field.imag * field.imag

    @staticmethod
    def backward(ctx, gI):
        emit, field, a = ctx.saved_tensors
        amp = a if ctx.has_amp else None
        gpos, gph, gamp = ctx.handle.adjoint(emit, field, gI.to(torch.float32).contiguous(), amp, ctx.extent,
                                             want_amp=ctx.has_amp)
        pd, phd, ad = ctx.dtypes
        return (gpos.to(pd), gph.to(phd).reshape(-1), None if gamp is None else gamp.to(ad), None, None)


def intensity(pos, phase, amp=None, *, Ndet, pixelsize, detz, k, settings="single nf", detangles=(0, 0),
              detoffset=(0, 0), extent=None):
    """|E|^2 on the detector, differentiable w.r.t. pos, phase and amp (FP32-split)."""
    if "single" not in settings:
        settings = settings + " single"
    h = get_handle(Ndet, pixelsize, detz, k, settings, detangles, detoffset, device=pos.device.index)
    if extent is None:
        extent = float(torch.linalg.vector_norm(pos.detach().to(torch.float64), dim=1).max()) if len(pos) else 0.0
    return _WavefieldIntensity.apply(pos, phase, amp, h, float(extent))


class _IntensityL2Loss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, phase, amp, target, handle, extent):
        emit = torch.cat([pos.detach().to(torch.float64), phase.detach().to(torch.float64).reshape(-1, 1)],
                         1).contiguous()
        a = None if amp is None else amp.detach().to(torch.float64).contiguous()
        loss, gpos, gph, gamp = handle.loss_grad_l2(emit, target.detach().to(torch.float32).contiguous(), a,
                                                    extent, want_amp=amp is not None)
        ctx.dtypes = (pos.dtype, phase.dtype, None if amp is None else amp.dtype, loss.dtype)
        ctx.save_for_backward(gpos, gph, gamp if gamp is not None else torch.empty(0, device=pos.device))
        ctx.has_amp = amp is not None
        return loss

    @staticmethod
    def backward(ctx, gL):
        gpos, gph, gamp = ctx.saved_tensors
        s = gL.to(torch.float32)
        pd, phd, ad, _ = ctx.dtypes
        return ((gpos * s).to(pd), (gph * s).to(phd).reshape(-1), (gamp * s).to(ad) if ctx.has_amp else None,
                None, None, None)


def intensity_l2_loss(pos, phase, amp=None, *, target, Ndet, pixelsize, detz, k, settings="single nf",
                      detangles=(0, 0), detoffset=(0, 0), extent=None):
    """sum_p (|E_p|^2 - target_p)^2 (a 0-d float64 tensor), differentiable w.r.t. pos, phase and amp
    (FP32-split); the gradients are computed with the loss, in the same call."""
    if "single" not in settings:
        settings = settings + " single"
    h = get_handle(Ndet, pixelsize, detz, k, settings, detangles, detoffset, device=pos.device.index)
    if extent is None:
        extent = float(torch.linalg.vector_norm(pos.detach().to(torch.float64), dim=1).max()) if len(pos) else 0.0
    return _IntensityL2Loss.apply(pos, phase, amp, target, h, float(extent))
