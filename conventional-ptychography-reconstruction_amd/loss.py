"""Regularisers of the reconstruction objective on the gfx950 kernels.

Reference: utils/loss.py:60-83 (haar_wavelet_sparsity_loss), added to the
step's objective as ``sparsity_weight * haar_wavelet_sparsity_loss(object)``
at Recovery.py:309-313.
"""
from __future__ import annotations

import torch

from . import _lib


class HaarSparsityFn(torch.autograd.Function):
    """L1 of the LH / HL / HH 2x2 Haar sub-bands of the real and imaginary
    parts over the last two axes; the kernel forms the loss and its gradient
    in one pass, the gradient is kept and scaled in backward."""

    @staticmethod
    def forward(ctx, img, reduction: str):
        _lib.require_cuda(img)
        if reduction not in ("sum", "mean"):
            raise ValueError(f"Unknown reduction: {reduction}")
        if img.dim() not in (2, 3, 4):
            raise ValueError(f"haar_wavelet_sparsity_loss takes a 2-D, 3-D or 4-D image, got {img.dim()}-D")
        ctx.real_input = not img.is_complex()
        ctx.in_dtype = img.dtype
        x = img.detach()
        x = torch.complex(x.float(), torch.zeros_like(x, dtype=torch.float32)) if ctx.real_input else x
        x = x.resolve_conj().contiguous().to(torch.complex64)
        W = x.shape[-1]
        rows = x.numel() // W if W else 0
        if x.shape[-2] % 2 or W % 2:
            raise RuntimeError(f"haar_wavelet_sparsity_loss: image {tuple(x.shape[-2:])} has an odd side "
                               "(the 2x2 sub-bands need even sizes)")
        scale = 1.0 if reduction == "sum" else 1.0 / ((rows // 2) * (W // 2))
        loss = torch.empty(1, dtype=torch.float32, device=x.device)
        grad = torch.zeros_like(x)
        ws = _lib.workspace(_lib.query("pty_haar_workspace", rows, W), x.device)
        _lib.call("pty_haar_sparsity", x.data_ptr(), rows, W, scale, 1.0, loss.data_ptr(), grad.data_ptr(),
                  ws.data_ptr(), ws.numel(), _lib.stream_ptr(x.device))
        ctx.save_for_backward(grad)
        return loss[0]

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        out = grad * g
        if ctx.real_input:
            out = out.real.to(ctx.in_dtype)
        return out, None


def haar_wavelet_sparsity_loss(img: torch.Tensor, reduction: str = "sum") -> torch.Tensor:
    """utils/loss.py:60-83: sum (or mean) of |LH| + |HL| + |HH| over the 2x2
    Haar blocks of the real and imaginary parts of ``img`` ([H, W],
    [B, H, W] or [B, C, H, W], even H and W), on the device."""
    return HaarSparsityFn.apply(img, reduction)
