"""Probe normalisation and the far-field loss block on the gfx950 kernels.

Reference: the inline step of Recovery.py:262-307 (same forward model as
utils/optics.py:102-112):
  * probe centring by centre-of-mass roll and energy normalisation to
    sum|p|^2 = M*N (Recovery.py:262-285)  -> ``normalize_probe``
  * per position: I = |fftshift(fft2(probe * object[window]))|^2 and
    smooth_L1 / FD / Poisson loss against the measured pattern, summed over
    the batch (Recovery.py:287-307)        -> ``farfield_loss``
The loss kernel computes the forward and the adjoint in one pass; the
gradients are kept and scaled by the incoming gradient in backward.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

VALID_LOSS_TYPES = tuple(_lib.LOSS_CODES)


class NormalizeProbeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, probe):
        _lib.require_cuda(probe)
        p = probe.resolve_conj().contiguous().to(torch.complex64)
        M, N = p.shape
        out = torch.empty_like(p)
        stats = torch.empty(8, dtype=torch.float32, device=p.device)
        ws = _lib.workspace(_lib.query("pty_probe_workspace", M, N), p.device)
        _lib.call("pty_probe_normalize_fwd", p.data_ptr(), M, N, out.data_ptr(), stats.data_ptr(), ws.data_ptr(),
                  ws.numel(), _lib.stream_ptr(p.device))
        ctx.save_for_backward(p, stats)
        ctx.mark_non_differentiable(stats)
        return out, stats

    @staticmethod
    def backward(ctx, gout, gstats):
        p, stats = ctx.saved_tensors
        M, N = p.shape
        g = gout.resolve_conj().contiguous()
        gp = torch.empty_like(p)
        ws = _lib.workspace(_lib.query("pty_probe_workspace", M, N), p.device)
        _lib.call("pty_probe_normalize_bwd", p.data_ptr(), g.data_ptr(), M, N, stats.data_ptr(), gp.data_ptr(),
                  ws.data_ptr(), ws.numel(), _lib.stream_ptr(p.device))
        return gp


def normalize_probe(probe: torch.Tensor, return_stats: bool = False):
    """COM roll + energy normalisation of Recovery.py:262-285, without host syncs.

    stats (device, 8 floats) = [cp0, cp1, scale, energy, rolled, scaled, 0, 0].
    """
    out, stats = NormalizeProbeFn.apply(probe)
    return (out, stats) if return_stats else out


def prepare_stack(stack_mnf: torch.Tensor) -> torch.Tensor:
    """Measured stack [M, N, frames] (reference layout, Recovery.py:103) ->
    pattern-major [frames, M, N] in natural FFT order (ifftshift applied)."""
    _lib.require_cuda(stack_mnf)
    s = stack_mnf.contiguous().float()
    M, N, F = s.shape
    out = torch.empty(F, M, N, dtype=torch.float32, device=s.device)
    _lib.call("pty_stack_prepare", s.data_ptr(), M, N, F, out.data_ptr(), _lib.stream_ptr(s.device))
    return out


def far_descriptor(M, N, canvas_h, canvas_w, n_batch, n_frames, loss_type) -> _lib.PtyFar:
    if loss_type not in _lib.LOSS_CODES:
        raise ValueError(f"loss_type must be one of {VALID_LOSS_TYPES}, got {loss_type!r}")
    d = _lib.PtyFar()
    d.M, d.N, d.canvas_h, d.canvas_w = M, N, canvas_h, canvas_w
    d.n_batch, d.n_frames, d.loss_type = n_batch, n_frames, _lib.LOSS_CODES[loss_type]
    return d


def positions_tensor(tlY, tlX, device) -> torch.Tensor:
    """[frames, 2] int32 (tlY, tlX) on the device."""
    t = np.stack([np.asarray(tlY, np.int64), np.asarray(tlX, np.int64)], axis=1).astype(np.int32)
    return torch.from_numpy(t).to(device)


def check_windows(tlY, tlX, M, N, canvas_h, canvas_w):
    tlY = np.asarray(tlY)
    tlX = np.asarray(tlX)
    if tlY.min() < 0 or tlX.min() < 0 or tlY.max() + M > canvas_h or tlX.max() + N > canvas_w:
        raise ValueError("scan windows fall outside the canvas")


class FarFieldLossFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, obj, probe_n, stack, batch_idx, tl, loss_type: str):
        _lib.require_cuda(obj, probe_n, stack, batch_idx, tl)
        o = obj.resolve_conj().contiguous().to(torch.complex64)
        p = probe_n.resolve_conj().contiguous().to(torch.complex64)
        bidx = batch_idx.to(torch.int32).contiguous()
        M, N = p.shape
        H, W = o.shape
        d = far_descriptor(M, N, H, W, bidx.numel(), stack.shape[0], loss_type)
        loss = torch.empty(bidx.numel(), dtype=torch.float32, device=o.device)
        dobj = torch.empty_like(o)
        dprobe = torch.empty_like(p)
        ws = _lib.workspace(_lib.query("pty_farfield_workspace", d), o.device)
        _lib.call("pty_farfield_fwd_bwd", d, o.data_ptr(), p.data_ptr(), stack.data_ptr(), bidx.data_ptr(),
                  tl.data_ptr(), loss.data_ptr(), dobj.data_ptr(), dprobe.data_ptr(), ws.data_ptr(), ws.numel(),
                  _lib.stream_ptr(o.device))
        ctx.save_for_backward(dobj, dprobe)
        ctx.mark_non_differentiable(loss)
        return loss.sum(), loss

    @staticmethod
    def backward(ctx, gsum, gper):
        dobj, dprobe = ctx.saved_tensors
        # the per-pattern output is informational; only the sum is differentiated
        return dobj * gsum, dprobe * gsum, None, None, None, None


def farfield_loss(obj, probe_n, stack, batch_idx, tl, loss_type: str = "smooth_L1_loss",
                  return_per_pattern: bool = False):
    """Sum over the batch of the per-pattern loss (Recovery.py:287-307).

    ``stack`` is the output of ``prepare_stack``; ``tl`` is [frames, 2] int32
    (tlY, tlX) on the device; ``batch_idx`` holds frame indices (device).
    """
    total, per = FarFieldLossFn.apply(obj, probe_n, stack, batch_idx, tl, loss_type)
    return (total, per) if return_per_pattern else total


def farfield_intensity(obj, probe_n, batch_idx, tl, n_frames: int) -> torch.Tensor:
    """|fftshift(fft2(probe_n * window))|^2 for each batch entry, [B, M, N] (forward only)."""
    _lib.require_cuda(obj, probe_n, batch_idx, tl)
    o = obj.detach().resolve_conj().contiguous().to(torch.complex64)
    p = probe_n.detach().resolve_conj().contiguous().to(torch.complex64)
    bidx = batch_idx.to(torch.int32).contiguous()
    M, N = p.shape
    H, W = o.shape
    d = far_descriptor(M, N, H, W, bidx.numel(), n_frames, "smooth_L1_loss")
    out = torch.empty(bidx.numel(), M, N, dtype=torch.float32, device=o.device)
    ws = _lib.workspace(_lib.query("pty_farfield_workspace", d), o.device)
    _lib.call("pty_farfield_intensity", d, o.data_ptr(), p.data_ptr(), bidx.data_ptr(), tl.data_ptr(),
              out.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_ptr(o.device))
    return out
