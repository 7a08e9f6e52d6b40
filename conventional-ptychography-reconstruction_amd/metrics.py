"""Object / probe error after the usual ptychographic ambiguities (host numpy):
NRMSE (below) and the reference's SSIM evaluation (end of the file).

Follows the recipe of scripts/compare_object_ssim.py:42-103 in the reference
(evaluation region from the scan footprints, integer alignment by magnitude
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


def evaluation_mask(tlY, tlX, M, N, shape, frac: float = 0.5) -> np.ndarray:
    """The region the reference evaluates a reconstruction on: the bounding
    box of the pixels covered by at least ceil(frac * max) scan footprints
    (compare_object_ssim.py:64-72, frac 0.5).  Unlike every pixel under some
    footprint, it leaves out the footprint corners a probe smaller than its
    window never illuminates, where the measurements do not determine the
    object (an INR puts arbitrary values there)."""
    cov = np.zeros(shape, np.int32)
    for y, x in zip(np.asarray(tlY), np.asarray(tlX)):
        cov[y:y + M, x:x + N] += 1
    y0, y1, x0, x1 = high_coverage_bbox(cov, frac)
    m = np.zeros(shape, bool)
    m[y0:y1, x0:x1] = True
    return m


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


# ----------------------------------------------------------------------------
# SSIM evaluation (scripts/compare_object_ssim.py:42-103, 132-158 and
# scripts/simulate_and_reconstruct.py:276-290 of the reference)
# ----------------------------------------------------------------------------
def structural_similarity(a: np.ndarray, b: np.ndarray, data_range: float, win_size: int = 7,
                          K1: float = 0.01, K2: float = 0.03) -> float:
    """Mean SSIM (Wang et al. 2004) with the defaults of the
    skimage.metrics.structural_similarity call the reference makes: a
    win_size x win_size uniform window (reflect padding), sample covariances
    (N / (N - 1)), C1 = (K1 R)^2, C2 = (K2 R)^2, and the mean taken over the
    map with (win_size - 1) / 2 pixels cropped on every side."""
    from scipy.ndimage import uniform_filter

    x = np.asarray(a, np.float64)
    y = np.asarray(b, np.float64)
    if x.shape != y.shape:
        raise ValueError(f"structural_similarity: shapes {x.shape} and {y.shape} differ")
    if win_size % 2 == 0 or win_size > min(x.shape):
        raise ValueError("win_size must be odd and not larger than the image")
    npix = win_size ** x.ndim
    cov_norm = npix / (npix - 1.0)
    f = lambda t: uniform_filter(t, size=win_size, mode="reflect")  # noqa: E731
    ux, uy = f(x), f(y)
    vx = cov_norm * (f(x * x) - ux * ux)
    vy = cov_norm * (f(y * y) - uy * uy)
    vxy = cov_norm * (f(x * y) - ux * uy)
    c1, c2 = (K1 * data_range) ** 2, (K2 * data_range) ** 2
    s = ((2 * ux * uy + c1) * (2 * vxy + c2)) / ((ux * ux + uy * uy + c1) * (vx + vy + c2))
    pad = (win_size - 1) // 2
    return float(s[pad:s.shape[0] - pad, pad:s.shape[1] - pad].mean())


def align_by_xcorr(recon_mag: np.ndarray, gt_mag: np.ndarray) -> tuple[int, int]:
    """(row, col) offset of the recon-sized window of the (larger or equal)
    ground truth that maximises the cross-correlation of the mean-removed
    magnitudes (compare_object_ssim.py:75-89)."""
    H, W = recon_mag.shape
    Hg, Wg = gt_mag.shape
    if Hg < H or Wg < W:
        raise ValueError(f"GT ({Hg}x{Wg}) smaller than recon ({H}x{W})")
    rf = np.fft.fft2(recon_mag - recon_mag.mean(), s=(Hg, Wg))
    gf = np.fft.fft2(gt_mag - gt_mag.mean(), s=(Hg, Wg))
    xc = np.fft.ifft2(gf * np.conj(rf)).real[:Hg - H + 1, :Wg - W + 1]
    r, c = np.unravel_index(int(np.argmax(xc)), xc.shape)
    return int(r), int(c)


def resolve_global_phase(recon: np.ndarray, gt: np.ndarray) -> np.ndarray:
    """recon * e^{i alpha} minimising ||recon e^{i alpha} - gt|| (compare_object_ssim.py:92-98)."""
    inner = np.vdot(recon.ravel(), gt.ravel())
    if abs(inner) < 1e-20:
        return recon
    return recon * (inner / abs(inner))


def normalize_amp(x: np.ndarray) -> np.ndarray:
    """x / mean|x| (compare_object_ssim.py:101-103)."""
    m = np.abs(x).mean()
    return x / m if m > 0 else x


def high_coverage_bbox(cov: np.ndarray, frac: float = 0.5) -> tuple[int, int, int, int]:
    """(y0, y1, x0, x1) bounding box of the pixels covered by at least
    ceil(frac * max) footprints (compare_object_ssim.py:64-72)."""
    thr = max(1, int(np.ceil(frac * cov.max())))
    ys, xs = np.where(cov >= thr)
    if ys.size == 0:
        return 0, cov.shape[0], 0, cov.shape[1]
    return int(ys.min()), int(ys.max()) + 1, int(xs.min()), int(xs.max()) + 1


def object_ssim(recon: np.ndarray, gt: np.ndarray, coverage: np.ndarray | None = None,
                coverage_frac: float = 0.5) -> dict:
    """Amplitude and phase SSIM of a reconstructed object against the ground
    truth as compare_object_ssim.py computes them: locate the recon in the GT
    by magnitude cross-correlation, crop, remove the global phase, amplitudes
    normalised to mean 1 with a shared data range, phases mapped by
    (phi + pi) / (2 pi) with data range 1.  With a coverage count map
    (footprints per recon pixel), also the SSIMs on the bounding box of the
    pixels covered by >= coverage_frac * max."""
    recon = np.asarray(recon, np.complex128)
    gt = np.asarray(gt, np.complex128)
    r0, c0 = align_by_xcorr(np.abs(recon), np.abs(gt))
    H, W = recon.shape
    g = gt[r0:r0 + H, c0:c0 + W]
    ra = resolve_global_phase(recon, g)
    a_r = np.abs(normalize_amp(ra)).astype(np.float32)
    a_g = np.abs(normalize_amp(g)).astype(np.float32)
    p_r = (np.angle(ra).astype(np.float32) + np.pi) / (2 * np.pi)
    p_g = (np.angle(g).astype(np.float32) + np.pi) / (2 * np.pi)

    def amp_phase(sl):
        ar, ag, pr, pg = a_r[sl], a_g[sl], p_r[sl], p_g[sl]
        dr = float(max(ar.max(), ag.max()) - min(ar.min(), ag.min()))
        return structural_similarity(ar, ag, dr), structural_similarity(pr, pg, 1.0)

    out = {"offset": (r0, c0)}
    out["ssim_amp"], out["ssim_phase"] = amp_phase((slice(None), slice(None)))
    if coverage is not None and coverage_frac > 0:
        y0, y1, x0, x1 = high_coverage_bbox(np.asarray(coverage), coverage_frac)
        out["coverage_bbox"] = (y0, y1, x0, x1)
        out["ssim_amp_cov"], out["ssim_phase_cov"] = amp_phase((slice(y0, y1), slice(x0, x1)))
    return out
