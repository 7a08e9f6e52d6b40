"""Host evaluation tools (NRMSE / SSIM of scripts/compare_object_ssim.py):
pinned by brute force and known answers (skimage is not installed here, so
the SSIM is restated from its published definition and checked window by
window)."""
import numpy as np
import pytest


def _ssim_loops(a, b, R, w=7, K1=0.01, K2=0.03):
    """Every interior window explicitly, sample (co)variances over w*w pixels."""
    pad = (w - 1) // 2
    n = w * w
    vals = []
    for r in range(pad, a.shape[0] - pad):
        for c in range(pad, a.shape[1] - pad):
            x = a[r - pad:r + pad + 1, c - pad:c + pad + 1].ravel().astype(np.float64)
            y = b[r - pad:r + pad + 1, c - pad:c + pad + 1].ravel().astype(np.float64)
            mx, my = x.mean(), y.mean()
            vx = ((x - mx) ** 2).sum() / (n - 1)
            vy = ((y - my) ** 2).sum() / (n - 1)
            cxy = ((x - mx) * (y - my)).sum() / (n - 1)
            c1, c2 = (K1 * R) ** 2, (K2 * R) ** 2
            vals.append((2 * mx * my + c1) * (2 * cxy + c2) / ((mx * mx + my * my + c1) * (vx + vy + c2)))
    return float(np.mean(vals))


def test_ssim_matches_window_loops(pkg):
    rng = np.random.default_rng(0)
    a = rng.random((19, 23))
    b = 0.7 * a + 0.3 * rng.random((19, 23))
    got = pkg.metrics.structural_similarity(a, b, 1.0)
    assert got == pytest.approx(_ssim_loops(a, b, 1.0), rel=1e-9)
    assert got == pytest.approx(pkg.metrics.structural_similarity(b, a, 1.0), rel=1e-12)
    assert pkg.metrics.structural_similarity(a, a, 1.0) == pytest.approx(1.0, abs=1e-12)
    # anti-correlated structure gives a negative index
    assert pkg.metrics.structural_similarity(a, 1.0 - a, 1.0) < 0


def test_object_ssim_recovers_crop_phase_and_scale(pkg):
    rng = np.random.default_rng(1)
    gt = (0.5 + rng.random((64, 72))) * np.exp(1j * rng.uniform(-1.2, 1.2, (64, 72)))
    recon = 3.0 * np.exp(0.9j) * gt[5:45, 9:57]
    out = pkg.metrics.object_ssim(recon, gt)
    assert out["offset"] == (5, 9)
    assert out["ssim_amp"] == pytest.approx(1.0, abs=1e-5)
    assert out["ssim_phase"] == pytest.approx(1.0, abs=1e-5)
    noisy = recon + 0.5 * (rng.standard_normal(recon.shape) + 1j * rng.standard_normal(recon.shape))
    worse = pkg.metrics.object_ssim(noisy, gt)
    assert worse["offset"] == (5, 9) and worse["ssim_amp"] < 0.9


def test_high_coverage_bbox_and_coverage_ssim(pkg):
    cov = np.zeros((20, 30), np.int32)
    cov[4:9, 6:12] = 3
    cov[10:12, 20:25] = 1
    assert pkg.metrics.high_coverage_bbox(cov, 0.5) == (4, 9, 6, 12)
    assert pkg.metrics.high_coverage_bbox(cov, 0.2) == (4, 12, 6, 25)
    rng = np.random.default_rng(2)
    gt = (0.5 + rng.random((30, 30))) * np.exp(1j * rng.uniform(-1, 1, (30, 30)))
    cov = np.zeros((30, 30), np.int32)
    cov[5:25, 3:28] = 2
    out = pkg.metrics.object_ssim(gt, gt, coverage=cov, coverage_frac=0.5)
    assert out["coverage_bbox"] == (5, 25, 3, 28)
    assert out["ssim_amp_cov"] == pytest.approx(1.0, abs=1e-6)


def test_evaluation_mask_is_the_high_coverage_box(pkg):
    """The evaluation region: bbox of pixels under >= ceil(0.5 * max) footprints
    (compare_object_ssim.py:64-72), here checked against explicit counting."""
    tlY = np.array([0, 0, 10, 10, 20])
    tlX = np.array([0, 10, 0, 10, 20])
    m = pkg.metrics.evaluation_mask(tlY, tlX, 16, 16, (40, 40))
    cov = np.zeros((40, 40), int)
    for y, x in zip(tlY, tlX):
        cov[y:y + 16, x:x + 16] += 1
    assert cov.max() == 4
    ys, xs = np.where(cov >= 2)
    want = np.zeros((40, 40), bool)
    want[ys.min():ys.max() + 1, xs.min():xs.max() + 1] = True
    assert np.array_equal(m, want)
    # single footprint: everything under it
    m1 = pkg.metrics.evaluation_mask([3], [5], 8, 8, (20, 20))
    assert m1.sum() == 64 and m1[3:11, 5:13].all()
