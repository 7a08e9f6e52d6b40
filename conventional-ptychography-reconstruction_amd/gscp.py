"""Drop-in for the reference's 2-D Gaussian-field rasterizer (``gscp``).

Same names, arguments and return values as the reference's extension module
(``gscp/csrc/rasterize_gaussians.cpp:26-254``, bound in ``ext.cpp:69-74``):

  rasterize_forward(xy, scaling, rotation, weights, H, W, max_patch_radius, min_scale, debug)
      -> (num_rendered, out_color [2,H,W], radii [N] int32, geomBuffer, binningBuffer, imgBuffer)
  rasterize_backward(xy, scaling, rotation, weights, radii, H, W, max_patch_radius, min_scale,
                     num_rendered, geomBuffer, binningBuffer, imgBuffer, dL_dout_color, debug)
      -> (dL_dxy, dL_dscaling, dL_drotation, dL_dweight)
  rasterize_forward_cholesky / rasterize_backward_cholesky  (log_L_diag, L_offdiag)

and the autograd functions of ``gscp/cuda/_wrapper.py:18-166``:
``GaussianRasterize2D.apply(xy, scaling, rotation, weight, H, W, max_patch_radius,
min_scale=0.5)`` and ``GaussianRasterize2DCholesky.apply(...)``.

The work runs in the hand-written HIP kernels of ``libptycho_hip.so``
(``pty_gs_*``).  The per-tile state that the reference keeps in ``imgBuffer``
lives in ``geomBuffer`` here, so ``imgBuffer`` is an empty byte tensor kept
for the tuple's shape.  One difference in behaviour: the gradients are summed
in a fixed order (the reference adds them with float atomics), so repeated
backward passes give the same bits.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _descriptor(n, H, W, max_patch_radius, min_scale, param):
    d = _lib.PtyGs()
    d.n, d.width, d.height = int(n), int(W), int(H)
    d.max_patch_radius = int(max_patch_radius)
    d.min_scale = float(min_scale)
    d.param = _lib.GS_PARAM_CODES[param]
    return d


def _check(name, t, n, cols):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name}: expected scalar type Float but found {t.dtype}")
    if not t.is_cuda:
        raise RuntimeError(f"{name}: the HIP rasterizer needs a device tensor")
    if t.numel() != n * cols:
        raise RuntimeError(f"{name} must hold {n} x {cols} values, got shape {tuple(t.shape)}")
    t = t.contiguous()
    if t.data_ptr() % 8:
        t = t.clone()
    return t


def _inputs(xy, p0, p1, weights, names):
    if xy.ndimension() != 2 or xy.size(1) != 2:
        raise RuntimeError("xy must have dimensions (num_points, 2)")
    n = xy.size(0)
    return (n, _check("xy", xy, n, 2), _check(names[0], p0, n, 2), _check(names[1], p1, n, 1),
            _check("weights", weights, n, 2))


def _forward(param, xy, p0, p1, weights, image_height, image_width, max_patch_radius, min_scale, debug,
             names):
    n, xy, p0, p1, weights = _inputs(xy, p0, p1, weights, names)
    H, W = int(image_height), int(image_width)
    dev = xy.device
    d = _descriptor(n, H, W, max_patch_radius, min_scale, param)
    out = torch.empty((2, H, W), dtype=torch.float32, device=dev)
    radii = torch.zeros((n,), dtype=torch.int32, device=dev)
    geom = torch.empty((max(1, _lib.query("pty_gs_geom_bytes", ctypes.byref(d))),), dtype=torch.uint8, device=dev)
    img = torch.empty((0,), dtype=torch.uint8, device=dev)
    s = _lib.stream_ptr(dev)
    with torch.cuda.device(dev):
        _lib.call("pty_gs_preprocess", ctypes.byref(d), xy.data_ptr(), p0.data_ptr(), p1.data_ptr(),
                  radii.data_ptr(), geom.data_ptr(), s)
        nr = ctypes.c_int64(0)
        _lib.call("pty_gs_num_rendered", ctypes.byref(d), geom.data_ptr(), ctypes.byref(nr), s)
        R = int(nr.value)
        # sized for the backward too, so that the saved buffers serve both passes
        binning = torch.empty((max(1, _lib.query("pty_gs_bin_bytes", ctypes.byref(d), R, 1)),), dtype=torch.uint8,
                              device=dev)
        _lib.call("pty_gs_render_fwd", ctypes.byref(d), geom.data_ptr(), R, weights.data_ptr(),
                  binning.data_ptr(), out.data_ptr(), s)
    if debug:
        torch.cuda.synchronize(dev)
    return R, out, radii, geom, binning, img


def _backward(param, xy, p0, p1, weights, radii, image_height, image_width, max_patch_radius, min_scale,
              num_rendered, geomBuffer, binningBuffer, imageBuffer, dL_dout_color, debug, names):
    n, xy, p0, p1, weights = _inputs(xy, p0, p1, weights, names)
    H, W = int(image_height), int(image_width)
    dev = xy.device
    d = _descriptor(n, H, W, max_patch_radius, min_scale, param)
    dxy = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    dp0 = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    dp1 = torch.zeros((n,) if names[1] == "L_offdiag" else (n, 1), dtype=torch.float32, device=dev)
    dw = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    if n == 0:
        return dxy, dp0, dp1, dw
    dout = _check("dL_dout_color", dL_dout_color, 2 * H, W)
    if radii.dtype != torch.int32 or radii.numel() != n or not radii.is_cuda:
        raise RuntimeError("radii must be the int32 [N] device tensor returned by the forward")
    R = int(num_rendered)
    if geomBuffer.numel() < _lib.query("pty_gs_geom_bytes", ctypes.byref(d)) or \
            (R > 0 and binningBuffer.numel() < _lib.query("pty_gs_bin_bytes", ctypes.byref(d), R, 1)):
        raise RuntimeError("geomBuffer / binningBuffer do not come from the forward of these shapes")
    with torch.cuda.device(dev):
        _lib.call("pty_gs_render_bwd", ctypes.byref(d), xy.data_ptr(), p0.data_ptr(), p1.data_ptr(),
                  weights.data_ptr(), radii.contiguous().data_ptr(), geomBuffer.data_ptr(), R,
                  binningBuffer.data_ptr(), dout.data_ptr(), dxy.data_ptr(), dp0.data_ptr(), dp1.data_ptr(),
                  dw.data_ptr(), _lib.stream_ptr(dev))
    if debug:
        torch.cuda.synchronize(dev)
    return dxy, dp0, dp1, dw


_RS = ("scaling", "rotation")
_CHOL = ("log_L_diag", "L_offdiag")


def rasterize_forward(xy, scaling, rotation, weights, image_height, image_width, max_patch_radius,
                      min_scale=0.5, debug=False):
    """``_C.rasterize_forward`` (rasterize_gaussians.cpp:26-78)."""
    return _forward("rs", xy, scaling, rotation, weights, image_height, image_width, max_patch_radius,
                    min_scale, debug, _RS)


def rasterize_backward(xy, scaling, rotation, weights, radii, image_height, image_width, max_patch_radius,
                       min_scale, num_rendered, geomBuffer, binningBuffer, imageBuffer, dL_dout_color,
                       debug=False):
    """``_C.rasterize_backward`` (rasterize_gaussians.cpp:80-137)."""
    return _backward("rs", xy, scaling, rotation, weights, radii, image_height, image_width, max_patch_radius,
                     min_scale, num_rendered, geomBuffer, binningBuffer, imageBuffer, dL_dout_color, debug, _RS)


def rasterize_forward_cholesky(xy, log_L_diag, L_offdiag, weights, image_height, image_width, max_patch_radius,
                               min_scale=0.5, debug=False):
    """``_C.rasterize_forward_cholesky`` (rasterize_gaussians.cpp:143-195)."""
    return _forward("cholesky", xy, log_L_diag, L_offdiag, weights, image_height, image_width, max_patch_radius,
                    min_scale, debug, _CHOL)


def rasterize_backward_cholesky(xy, log_L_diag, L_offdiag, weights, radii, image_height, image_width,
                                max_patch_radius, min_scale, num_rendered, geomBuffer, binningBuffer, imageBuffer,
                                dL_dout_color, debug=False):
    """``_C.rasterize_backward_cholesky`` (rasterize_gaussians.cpp:197-254)."""
    return _backward("cholesky", xy, log_L_diag, L_offdiag, weights, radii, image_height, image_width,
                     max_patch_radius, min_scale, num_rendered, geomBuffer, binningBuffer, imageBuffer,
                     dL_dout_color, debug, _CHOL)


class GaussianRasterize2D(torch.autograd.Function):
    """Rotation + scaling parameterisation (gscp/cuda/_wrapper.py:18-87):
    (xy [N,2], scaling [N,2] log-space, rotation [N,1], weight [N,2], H, W,
    max_patch_radius, min_scale) -> [2, H, W]."""

    @staticmethod
    def forward(ctx, xy, scaling, rotation, weight, H, W, max_patch_radius, min_scale=0.5):
        num_rendered, out, radii, geom, binning, img = rasterize_forward(
            xy.contiguous(), scaling.contiguous(), rotation.contiguous(), weight.contiguous(), H, W,
            max_patch_radius, min_scale, False)
        ctx.save_for_backward(xy, scaling, rotation, weight, radii, geom, binning, img)
        ctx.num_rendered, ctx.H, ctx.W = num_rendered, H, W
        ctx.max_patch_radius, ctx.min_scale = max_patch_radius, min_scale
        return out

    @staticmethod
    def backward(ctx, dL_dout):
        xy, scaling, rotation, weight, radii, geom, binning, img = ctx.saved_tensors
        dxy, dscal, drot, dw = rasterize_backward(
            xy, scaling, rotation, weight, radii, ctx.H, ctx.W, ctx.max_patch_radius, ctx.min_scale,
            ctx.num_rendered, geom, binning, img, dL_dout.contiguous(), False)
        return dxy, dscal, drot.view_as(rotation), dw, None, None, None, None


class GaussianRasterize2DCholesky(torch.autograd.Function):
    """Cholesky parameterisation Sigma = L L^T (gscp/cuda/_wrapper.py:90-166):
    (xy [N,2], log_L_diag [N,2], L_offdiag [N], weight [N,2], H, W,
    max_patch_radius, min_scale) -> [2, H, W]."""

    @staticmethod
    def forward(ctx, xy, log_L_diag, L_offdiag, weight, H, W, max_patch_radius, min_scale=0.5):
        num_rendered, out, radii, geom, binning, img = rasterize_forward_cholesky(
            xy.contiguous(), log_L_diag.contiguous(), L_offdiag.contiguous(), weight.contiguous(), H, W,
            max_patch_radius, min_scale, False)
        ctx.save_for_backward(xy, log_L_diag, L_offdiag, weight, radii, geom, binning, img)
        ctx.num_rendered, ctx.H, ctx.W = num_rendered, H, W
        ctx.max_patch_radius, ctx.min_scale = max_patch_radius, min_scale
        return out

    @staticmethod
    def backward(ctx, dL_dout):
        xy, log_L_diag, L_offdiag, weight, radii, geom, binning, img = ctx.saved_tensors
        dxy, dld, doff, dw = rasterize_backward_cholesky(
            xy, log_L_diag, L_offdiag, weight, radii, ctx.H, ctx.W, ctx.max_patch_radius, ctx.min_scale,
            ctx.num_rendered, geom, binning, img, dL_dout.contiguous(), False)
        return dxy, dld, doff.view_as(L_offdiag), dw, None, None, None, None
