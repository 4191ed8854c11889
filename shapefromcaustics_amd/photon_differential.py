"""``PDS``: the reference's ``torch.autograd.Function`` for photon-differential splatting.

Same signature and gradient convention as ``model/photon_differential.py:7-19``
of the reference: ``PDS.apply(Ep, xp, Mp, cp, radius, output_size,
max_pixel_radius)`` returns the ``[C, H, W]`` splat; the backward returns
gradients for ``(Ep, xp, Mp)`` and ``None`` for the channel index, radius,
output size and cap (the radius gets no gradient, as in the reference).

An optional eighth argument selects the kernel mode (``"auto"``, ``"plain"``,
``"binned"``, ``"deterministic"``); callers that pass seven arguments get the
default, exactly like the reference.
"""
from __future__ import annotations

import typing as tp

import torch

from .PhotonDifferentialSplatting import _backward, _forward


class PDS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Ep: torch.Tensor, xp: torch.Tensor, Mp: torch.Tensor, cp: torch.Tensor, radius: torch.Tensor,
                output_size: tp.Tuple, max_pixel_radius: int, mode: str = "auto"):
        ctx.save_for_backward(Ep, xp, Mp, cp, radius)
        ctx.max_pixel_radius = max_pixel_radius
        ctx.mode = mode
        # when a gradient will be needed, the forward leaves its tile binning in
        # a buffer that the backward reuses instead of binning again
        keep = any(ctx.needs_input_grad[:3])
        out, bins = _forward(Ep, xp, Mp, cp, radius, output_size, max_pixel_radius, mode, None, keep_bins=keep)
        ctx.bins = bins
        return out

    @staticmethod
    def backward(ctx, grad_pds: torch.Tensor):
        Ep, xp, Mp, cp, radius = ctx.saved_tensors
        grad_Ep, grad_xp, grad_Mp = _backward(grad_pds, Ep, xp, Mp, cp, radius, ctx.max_pixel_radius, ctx.mode, None,
                                              bins=ctx.bins)
        ctx.bins = None
        return grad_Ep, grad_xp, grad_Mp, None, None, None, None, None


pds = PDS  # the reference imports it as ``PDS as pds`` (model/caustics.py:10)
