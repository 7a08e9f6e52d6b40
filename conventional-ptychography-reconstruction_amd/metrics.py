"""Object / probe error after the usual ptychographic ambiguities (host numpy).

Follows the recipe of scripts/compare_object_ssim.py:42-103 in the reference
(coverage mask from the scan footprints, integer alignment by magnitude
cross-correlation, global phase), with a complex scale c = <r, g> / <r, r>
(phase and amplitude) because the energy normalisation fixes the probe
scale, not the object scale:  NRMSE = ||c r - g|| / ||g|| over the mask.
"""
from __future__ import annotations

import numpy as np


def coverage_mask(tlY, tlX, M, N, shape, min_count: int = 1) -> np.ndarray:
    cov = np.zeros(shape, np.int32)
    for y, x in zip(np.asarray(tlY), np.asarray(tlX)):
        cov[y:y + M, x:x + N] += 1
    return cov >= min_count


def xcorr_shift(recon_mag: np.ndarray, gt_mag: np.ndarray, max_shift: int) -> tuple[int, int]:
    """Integer (dy, dx) maximising the circular cross-correlation, |d| <= max_shift."""
    a = recon_mag - recon_mag.mean()
    b = gt_mag - gt_mag.mean()
    xc = np.real(np.fft.ifft2(np.fft.fft2(b) * np.conj(np.fft.fft2(a))))
    best, arg = -np.inf, (0, 0)
    H, W = xc.shape
    for dy in range(-max_shift, max_shift + 1):
        for dx in range(-max_shift, max_shift + 1):
            v = xc[dy % H, dx % W]
            if v > best:
                best, arg = v, (dy, dx)
    return arg


def aligned_nrmse(recon: np.ndarray, gt: np.ndarray, mask: np.ndarray | None = None,
                  max_shift: int = 0) -> float:
    """min over integer shifts (|d| <= max_shift) and complex scale of ||c r - g|| / ||g||."""
    r = np.asarray(recon, np.complex128)
    g = np.asarray(gt, np.complex128)
    if mask is None:
        mask = np.ones(g.shape, bool)
    if max_shift > 0:
        dy, dx = xcorr_shift(np.abs(r) * mask, np.abs(g) * mask, max_shift)
        r = np.roll(r, (dy, dx), axis=(0, 1))
    rm, gm = r[mask], g[mask]
    den = np.vdot(rm, rm)
    c = np.vdot(rm, gm) / den if abs(den) > 0 else 0.0
    return float(np.linalg.norm(c * rm - gm) / np.linalg.norm(gm))
