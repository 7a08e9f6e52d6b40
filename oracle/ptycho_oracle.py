"""CPU restatement of the reference INR-ptychography hot path (the oracle).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; it is
used by ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` as the checker that the HIP path is compared against.

Every function restates one piece of the reference, cited by file:line
(paths relative to the reference repository root).  Floating-point work runs
in the dtype passed in (float64 / complex128 for parity checks, float32 for
the CPU timing baseline).  Integer index math of the hash grid is exact and
shares its per-level table construction rules with the product, but the
table is rebuilt here from the rules, not imported.

Pinning (what checks each function, see tests/test_oracle.py):
  level_table          Appendix-A values (integer scales, dense/hash split),
                       next-multiple-of-8 sizes, totals per log2 size.
  grid_coords          torch.linspace (library) + meshgrid transpose identity.
  fma_f32              exact rational arithmetic (fractions) with explicit
                       round-to-nearest-even, on random inputs and on
                       positions within a few ulps of cell boundaries.
  hash_corners         brute force on tiny grids, dense identity x+y*res,
                       uint32 wrap of the prime hash, edge-corner wrap.
  hashgrid_encode      vertex interpolation identity, weights sum to 1,
                       adjoint identity <enc(t),y> = <t,enc^T(y)>.
  complex_mlp          closed form per layer (numpy, written independently),
                       |z| = exp(-w Im v), autograd vs central differences.
  fourier_upsample     numpy.fft reference, exact even-sample identity,
                       adjoint identity <U x, y> = <x, U^H y>.
  probe_center_normalize  closed-form centre of mass, roll invariance.
  farfield / losses    numpy.fft, Parseval, F.smooth_l1_loss, closed-form
                       gradients vs autograd, central differences.
  scan_positions_to_pixels  numpy FP64 round-half-even on constructed inputs.
  haar_wavelet_sparsity_loss  explicit per-block loops, closed forms (impulse,
                       constant, checkerboard), the orthonormal-Haar energy
                       identity of the four bands, autograd vs central
                       differences.
  wsnet_forward        the torch module of the same architecture (nn.Conv2d,
                       nn.ConvTranspose2d) with the same weights, an explicit
                       loop convolution and transposed convolution on a
                       tiny input, central differences of a loss.
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F

# ----------------------------------------------------------------------------
# Hash grid (tiny-cuda-nn HashGrid, 2-D input) -- call sites
# models/complex_inr.py:130-145 (config), :279 and :294 (forward).
# Restated from the public Instant-NGP description (Mueller et al. 2022,
# Eq. 4 prime hash; Appendix A 0.5 offset).  tcnn itself is not available,
# so parity with tcnn's exact bits is unpinned; this restatement is the oracle.
# ----------------------------------------------------------------------------

HASH_PRIME_Y = 2654435761


def _next_multiple(v: int, m: int) -> int:
    return (v + m - 1) // m * m


def level_table(n_levels: int, log2_hashmap_size: int, base_resolution: int,
                per_level_scale: float) -> dict:
    """Per-level scale (f32), resolution, table size and entry offset.

    scale_l = exp2(l * log2(per_level_scale)) * base - 1, every step in f32;
    res_l = ceil(scale_l) + 1; size_l = min(next_mult8(res_l^2), 2^log2T);
    a level is dense when res_l^2 <= size_l, hashed otherwise.
    """
    f32 = np.float32
    log2_pls = f32(np.log2(f32(per_level_scale)))
    T = 1 << int(log2_hashmap_size)
    scale = np.zeros(n_levels, np.float32)
    res = np.zeros(n_levels, np.int64)
    size = np.zeros(n_levels, np.int64)
    offset = np.zeros(n_levels, np.int64)
    dense = np.zeros(n_levels, bool)
    off = 0
    for l in range(n_levels):
        s = f32(np.exp2(f32(l) * log2_pls)) * f32(base_resolution) - f32(1.0)
        s = f32(s)
        r = int(math.ceil(float(s))) + 1
        n = min(_next_multiple(r * r, 8), T)
        scale[l], res[l], size[l], offset[l] = s, r, n, off
        dense[l] = r * r <= n
        off += n
    return dict(scale=scale, res=res, size=size, offset=offset, dense=dense,
                n_entries=off, n_levels=n_levels)


def grid_coords(width: int, height: int) -> torch.Tensor:
    """Cell-centred coordinates, models/complex_inr.py:184-189 (f32).

    Row n = i*height + j holds (ys[j], xs[i]).
    """
    half_dx = 0.5 / width
    half_dy = 0.5 / height
    xs = torch.linspace(half_dx, 1 - half_dx, width)
    ys = torch.linspace(half_dy, 1 - half_dy, height)
    xv, yv = torch.meshgrid([xs, ys], indexing="ij")
    return torch.stack((yv.flatten(), xv.flatten())).t()


def fma_f32(x: np.ndarray, s, c) -> np.ndarray:
    """f32(x*s + c) with ONE rounding (fused multiply-add) for f32 x, s, c.

    The product of two f32 values is exact in f64; TwoSum gives the sum and its
    exact error; the f64 sum rounded to f32 is the correctly rounded result
    except when it sits exactly on a midpoint between two f32 values with a
    nonzero error (double rounding), where the error decides the side.
    """
    f64 = np.float64
    x = np.asarray(x, np.float32)
    p = x.astype(f64) * f64(np.float32(s))
    cc = f64(np.float32(c))
    t = p + cc
    bp = t - cc
    e = (p - bp) + (cc - (t - bp))
    r = t.astype(np.float32)
    r64 = r.astype(f64)
    other = np.nextafter(r, np.where(t > r64, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32))
    mid = (r64 + other.astype(f64)) * 0.5
    flip = (t == mid) & (t != r64) & (e != 0) & (np.sign(e) == np.sign(other.astype(f64) - r64))
    return np.where(flip, other, r).astype(np.float32)


def hash_corners(coords_f32: np.ndarray, table: dict):
    """Corner entry indices [N, L, 4] and fractional positions [N, L, 2].

    pos = fma_f32(x, scale, 0.5) (one rounding, the form of tiny-cuda-nn's
    pos_fract), g = floor(pos), frac = pos - g (exact in f32).  Corner c has
    offsets (c&1, c>>1).
    Corner coordinates are uint32 (tiny-cuda-nn's pos_grid is the uint32 of
    (int)floor(pos)): dense index ((cx + cy*res) mod 2^32) % size, hashed
    index ((cx * 1) ^ (cy * 2654435761)) mod 2^32 % size.  For inputs in
    [0, 1] cx, cy >= 0 and the dense index is plainly (cx + cy*res) % size.
    """
    x = np.asarray(coords_f32, np.float32)
    N = x.shape[0]
    L = table["n_levels"]
    idx = np.zeros((N, L, 4), np.int64)
    frac = np.zeros((N, L, 2), np.float32)
    for l in range(L):
        s = table["scale"][l]
        pos = fma_f32(x, s, 0.5)
        g = np.floor(pos)
        frac[:, l, :] = (pos - g).astype(np.float32)
        gi = g.astype(np.int64)
        res, size, off = int(table["res"][l]), int(table["size"][l]), int(table["offset"][l])
        for c in range(4):
            cx = gi[:, 0] + (c & 1)
            cy = gi[:, 1] + ((c >> 1) & 1)
            if table["dense"][l]:
                local = (((cx & 0xFFFFFFFF) + (cy & 0xFFFFFFFF) * res) & 0xFFFFFFFF) % size
            else:
                h = ((cx & 0xFFFFFFFF) ^ ((cy * HASH_PRIME_Y) & 0xFFFFFFFF)) & 0xFFFFFFFF
                local = h % size
            idx[:, l, c] = off + local
    return idx, frac


def corner_weights(frac: np.ndarray, dtype=torch.float64) -> torch.Tensor:
    """Bilinear weights [N, L, 4]: w_c = (c&1 ? fx : 1-fx) * (c&2 ? fy : 1-fy)."""
    f = torch.from_numpy(frac).to(dtype)
    fx, fy = f[..., 0], f[..., 1]
    w = torch.stack([(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy], -1)
    return w


def hashgrid_encode(coords, params, table: dict, n_features: int = 2,
                    dtype=torch.float64) -> torch.Tensor:
    """Features [N, L*F], column l*F+f = sum_c w_c * params[(off+idx_c)*F + f].

    ``params`` is the tcnn-layout 1-D master table [n_entries*F]; autograd
    through this function gives the table gradient (scatter-add of w_c*dy).
    """
    c = coords.detach().cpu().numpy() if torch.is_tensor(coords) else np.asarray(coords)
    idx, frac = hash_corners(c, table)
    w = corner_weights(frac, dtype)                      # [N, L, 4]
    tab = params.to(dtype).reshape(-1, n_features)       # [entries, F]
    g = tab[torch.from_numpy(idx)]                       # [N, L, 4, F]
    out = (w.unsqueeze(-1) * g).sum(2)                   # [N, L, F]
    return out.reshape(out.shape[0], -1)


# ----------------------------------------------------------------------------
# Complex-exponential MLP -- models/complex_inr.py:6-82, 308-314
# ----------------------------------------------------------------------------

def gabor_layer(x, weight, bias, omega):
    """z = exp(1j * omega * (x W^T + b)), models/complex_inr.py:38-40."""
    lin = F.linear(x, weight, bias)
    return torch.exp(1j * omega * lin)


def complex_mlp(h, layers, out_weight, out_bias):
    """ComplexMLP.net applied to features (models/complex_inr.py:57-78, 313).

    ``layers`` is a list of (weight, bias, omega) for the real first layer and
    the complex hidden layers; the final complex Linear has no activation and
    its complex output is returned (the ``.real`` of ComplexMLP.forward is
    bypassed by ComplexINRModel2D._get_complex_output).
    """
    z = h
    for (w, b, om) in layers:
        z = gabor_layer(z, w, b, om)
    return F.linear(z, out_weight, out_bias)


# ----------------------------------------------------------------------------
# Fourier upsample -- models/complex_inr.py:252-271
# ----------------------------------------------------------------------------

def fourier_upsample(x, scale_factor):
    """ds^2 * ifft2(ifftshift(pad(fftshift(fft2(x))))) over the last two dims."""
    H, W = x.shape[-2], x.shape[-1]
    X = torch.fft.fftshift(torch.fft.fft2(x, dim=(-2, -1)), dim=(-2, -1))
    new_H, new_W = int(H * scale_factor), int(W * scale_factor)
    pad_h, pad_w = (new_H - H) // 2, (new_W - W) // 2
    Xp = F.pad(X, (pad_w, new_W - W - pad_w, pad_h, new_H - H - pad_h), mode="constant", value=0)
    y = torch.fft.ifft2(torch.fft.ifftshift(Xp, dim=(-2, -1)), dim=(-2, -1))
    return y * (scale_factor ** 2)


# ----------------------------------------------------------------------------
# Whole model forward -- models/complex_inr.py:273-306, 316-334
# ----------------------------------------------------------------------------

def inr_field(coords, table_params, table, mlp, width, height, n_features=2,
              dtype=torch.float64):
    """HashGrid -> ComplexMLP over a coordinate set, reshaped (width, height)."""
    cdtype = torch.complex128 if dtype == torch.float64 else torch.complex64
    h = hashgrid_encode(coords, table_params, table, n_features, dtype)
    layers = [(w.to(dtype if i == 0 else cdtype), b.to(dtype if i == 0 else cdtype), om.to(dtype))
              for i, (w, b, om) in enumerate(mlp["layers"])]
    o = complex_mlp(h, layers, mlp["out_weight"].to(cdtype), mlp["out_bias"].to(cdtype))
    return o.reshape(width, height)


def model_forward(obj_coords, obj_table_params, obj_mlp, prb_coords, prb_table_params,
                  prb_mlp, table, ds_w, ds_h, probe_w, probe_h, downsample_factor,
                  k_object=None, object_initial=None, k_probe=None, probe_initial=None,
                  n_features=2, dtype=torch.float64):
    """(object [W,H], probe [Wp,Hp]) as ComplexINRModel2D.forward computes them."""
    o = inr_field(obj_coords, obj_table_params, table, obj_mlp, ds_w, ds_h, n_features, dtype)
    o = fourier_upsample(o.unsqueeze(0).unsqueeze(0), downsample_factor).squeeze(0).squeeze(0)
    p = inr_field(prb_coords, prb_table_params, table, prb_mlp, probe_w, probe_h, n_features, dtype)
    if object_initial is not None:
        o = k_object.to(dtype) * o + object_initial.to(o.dtype)
    if probe_initial is not None:
        p = k_probe.to(dtype) * p + probe_initial.to(p.dtype)
    return o, p


# ----------------------------------------------------------------------------
# Probe centring and energy normalisation -- Recovery.py:262-285
# ----------------------------------------------------------------------------

def probe_com_shift(probe):
    """cp as Recovery.py:262-275 computes it (None when total power is 0)."""
    with torch.no_grad():
        absP2 = probe.abs() ** 2
        Mp, Np = absP2.shape
        tot = absP2.sum()
        if tot == 0:
            return None
        mn = torch.tensor([Mp, Np], dtype=absP2.dtype)
        cp = torch.trunc(mn / 2 - mn * torch.stack((absP2.sum(1).cumsum(0).mean(),
                                                     absP2.sum(0).cumsum(0).mean())) / tot + 1).long()
    return cp


def probe_center_normalize(probe, cp="auto"):
    """Roll by -cp (if any nonzero) then scale so sum|p|^2 = M*N (Recovery.py:276-285)."""
    if isinstance(cp, str):
        cp = probe_com_shift(probe)
    if cp is not None and bool(cp.any()):
        probe = torch.roll(probe, (-int(cp[0]), -int(cp[1])), (0, 1))
    energy = (probe.abs() ** 2).sum()
    if energy > 0:
        target = float(probe.shape[-2] * probe.shape[-1])
        probe = probe * (target / energy).sqrt()
    return probe


# ----------------------------------------------------------------------------
# Far field and losses -- Recovery.py:287-307 (same math utils/optics.py:102-112)
# ----------------------------------------------------------------------------

def farfield_intensity(obj, probe, tlY, tlX):
    """|fftshift(fft2(probe * obj[tlY:tlY+M, tlX:tlX+N]))|^2."""
    M, N = probe.shape
    ew = probe * obj[tlY:tlY + M, tlX:tlX + N]
    return torch.abs(torch.fft.fftshift(torch.fft.fft2(ew), dim=(-2, -1))) ** 2


def pattern_loss(I, meas, loss_type):
    """Per-pattern loss, Recovery.py:294-305 (mean over pixels)."""
    if loss_type == "smooth_L1_loss":
        return F.smooth_l1_loss(I, meas)
    if loss_type == "FD_loss":
        eps = 1e-8
        dW = torch.sqrt(meas + eps) - torch.sqrt(I + eps)
        Rx = torch.cat((dW[:, 1:] - dW[:, :-1], dW[:, :1] - dW[:, -1:]), dim=1)
        Ry = torch.cat((dW[1:, :] - dW[:-1, :], dW[:1, :] - dW[-1:, :]), dim=0)
        return torch.sqrt(Rx ** 2 + Ry ** 2 + eps).mean()
    if loss_type == "Poisson_likelihood_loss":
        eps = 1e-8
        return (I - meas * torch.log(I + eps)).mean()
    raise ValueError(f"unknown loss_type {loss_type!r}")


def batch_loss(obj, probe_n, stack_mnp, batch_idx, tlY, tlX, loss_type):
    """Sum over the batch of per-pattern losses (Recovery.py:287-307).

    ``stack_mnp`` is the measured stack in the reference layout [M, N, frames].
    """
    total = 0.0
    for i in batch_idx:
        i = int(i)
        I = farfield_intensity(obj, probe_n, int(tlY[i]), int(tlX[i]))
        total = total + pattern_loss(I, stack_mnp[:, :, i].to(I.dtype), loss_type)
    return total


def haar_wavelet_sparsity_loss(img, reduction="sum"):
    """utils/loss.py:60-83: L1 of the LH, HL, HH 2x2 Haar sub-bands of the
    real and imaginary parts over the last two axes ('sum' or 'mean' per
    band, as torch.sum / torch.mean there)."""
    if torch.is_complex(img):
        return (haar_wavelet_sparsity_loss(img.real, reduction)
                + haar_wavelet_sparsity_loss(img.imag, reduction))
    a, b = img[..., 0::2, 0::2], img[..., 0::2, 1::2]
    c, d = img[..., 1::2, 0::2], img[..., 1::2, 1::2]
    red = torch.sum if reduction == "sum" else torch.mean
    return red((a - b + c - d).abs()) + red((a + b - c - d).abs()) + red((a - b - c + d).abs())


# ----------------------------------------------------------------------------
# Geometry -- utils/optics.py:68-99 (FP64 host arithmetic)
# ----------------------------------------------------------------------------

def scan_positions_to_pixels(xlocation, ylocation, probe_shape, wavelength,
                             camera_length, camera_pixel_pitch):
    M, N = probe_shape
    x_pos = np.asarray(xlocation, np.float64).squeeze() - np.min(xlocation)
    y_pos = np.asarray(ylocation, np.float64).squeeze() - np.min(ylocation)
    dx = wavelength * camera_length / (np.array([M, N], np.float64) * camera_pixel_pitch)
    tlY = np.round(y_pos / dx[0]).astype(np.int64)
    tlX = np.round(x_pos / dx[1]).astype(np.int64)
    brY, brX = tlY + M, tlX + N
    canvas = math.ceil(max(int(brY.max()), int(brX.max())) / 6) * 6
    return tlX, tlY, brX, brY, dx, canvas


# ----------------------------------------------------------------------------
# One optimizer step of the reference loop -- Recovery.py:247-321
# ----------------------------------------------------------------------------

def step_loss(params: dict, table: dict, geom: dict, stack_mnp, batch_idx, loss_type,
              dtype=torch.float64, sparsity_weight: float = 0.0):
    """Loss of one step for parameter dict ``params`` (autograd-ready).

    ``params`` keys: obj_table, prb_table, obj_mlp, prb_mlp (dicts with
    'layers' [(w,b,omega)...], 'out_weight', 'out_bias'), optional k_probe /
    probe_initial / k_object / object_initial.  ``geom`` keys: obj_coords,
    prb_coords, ds_w, ds_h, probe_w, probe_h, ds, tlY, tlX.
    Returns (loss, object, normalised probe).
    """
    o, p = model_forward(geom["obj_coords"], params["obj_table"], params["obj_mlp"],
                         geom["prb_coords"], params["prb_table"], params["prb_mlp"], table,
                         geom["ds_w"], geom["ds_h"], geom["probe_w"], geom["probe_h"], geom["ds"],
                         params.get("k_object"), params.get("object_initial"),
                         params.get("k_probe"), params.get("probe_initial"), dtype=dtype)
    pn = probe_center_normalize(p)
    loss = batch_loss(o, pn, stack_mnp, batch_idx, geom["tlY"], geom["tlX"], loss_type)
    if sparsity_weight:
        # Recovery.py:309-313: the weighted Haar term joins the objective
        loss = loss + sparsity_weight * haar_wavelet_sparsity_loss(o)
    return loss, o, pn


# ----------------------------------------------------------------------------
# Adam -- torch.optim.Adam semantics used at Recovery.py:238-241,321
# ----------------------------------------------------------------------------

def adam_update(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update on real views (complex parameters as (re, im) pairs)."""
    m.mul_(beta1).add_(g, alpha=1 - beta1)
    v.mul_(beta2).addcmul_(g, g, value=1 - beta2)
    bc1 = 1 - beta1 ** step
    bc2 = 1 - beta2 ** step
    denom = (v.sqrt() / math.sqrt(bc2)).add_(eps)
    p.addcdiv_(m, denom, value=-lr / bc1)
    return p


# ----------------------------------------------------------------------------
# Learned warm start: the conv encoder-decoder of the warm-start row (no
# counterpart in the reference; it seeds object_initial of
# models/complex_inr.py:209-250).  Functional restatement from the flat
# parameter vector the HIP kernels use (torch parameter order; conv weights
# [Co][kh][kw][Ci], transposed-conv weights [kh][kw][Co][Ci]), with the
# library routines F.conv2d / F.conv_transpose2d / F.gelu / F.avg_pool2d.
# ----------------------------------------------------------------------------

def wsnet_layers(base: int):
    """(kind, ci, co) of the 15 layers: 'c3', 'c1' convolutions, 'up' 2x2/2."""
    c = base
    return [("c3", 3, c), ("c3", c, c), ("c3", c, 2 * c), ("c3", 2 * c, 2 * c), ("c3", 2 * c, 4 * c),
            ("c3", 4 * c, 4 * c), ("c3", 4 * c, 8 * c), ("c3", 8 * c, 8 * c), ("up", 8 * c, 4 * c),
            ("c3", 8 * c, 4 * c), ("c3", 4 * c, 4 * c), ("up", 4 * c, 2 * c), ("c3", 4 * c, 2 * c),
            ("c3", 2 * c, 2 * c), ("c1", 2 * c, 2)]


def wsnet_unpack(flat, base: int):
    """[(weight in torch's layout, bias)] per layer, views of ``flat``."""
    out, off = [], 0
    for kind, ci, co in wsnet_layers(base):
        if kind == "up":
            w = flat[off:off + 4 * co * ci].reshape(2, 2, co, ci).permute(3, 2, 0, 1)   # [Ci, Co, kh, kw]
            off += 4 * co * ci
        else:
            k = 3 if kind == "c3" else 1
            w = flat[off:off + k * k * ci * co].reshape(co, k, k, ci).permute(0, 3, 1, 2)  # [Co, Ci, kh, kw]
            off += k * k * ci * co
        b = flat[off:off + co]
        off += co
        out.append((w, b))
    assert off == flat.numel(), (off, flat.numel())
    return out


def wsnet_forward(flat, x_nhwc, base: int):
    """Net output [B, 2, 32, 32] for inputs [B, 64, 64, 3] (NHWC)."""
    L = wsnet_unpack(flat, base)
    x = x_nhwc.permute(0, 3, 1, 2)

    def conv(i, h):
        w, b = L[i]
        return F.gelu(F.conv2d(h, w, b, padding=1))

    def up(i, h):
        w, b = L[i]
        return F.conv_transpose2d(h, w, b, stride=2)

    x1 = conv(1, conv(0, x))
    x2 = conv(3, conv(2, F.avg_pool2d(x1, 2)))
    x3 = conv(5, conv(4, F.avg_pool2d(x2, 2)))
    x4 = conv(7, conv(6, F.avg_pool2d(x3, 2)))
    y3 = conv(10, conv(9, torch.cat([up(8, x4), x3], 1)))
    y2 = conv(13, conv(12, torch.cat([up(11, y3), x2], 1)))
    w, b = L[14]
    return F.conv2d(y2, w, b)
