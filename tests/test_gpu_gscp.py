"""GPU parity of the Gaussian-field rasterizer (pty_gs_*) against the oracle.

The oracle (oracle/gscp_oracle.py) restates the reference's CUDA rasterizer:
float32 geometry (radii and tile rects are integers and must match exactly),
FP64 render and adjoints.  Bars:
  radii, pair counts            exact (a radius may differ only where the
                                float 3 sqrt(lambda) lies within 1e-4 of an
                                integer, where one ulp of exp/cos/sin decides)
  canvas                        |gpu - ref| <= 1e-5 * S + 1e-6 per pixel, S the
                                same render with |w| (the sum of magnitudes)
  gradients                     |gpu - ref| <= 1e-4 * max|ref| + 1e-4 * |ref|
                                per tensor (float32 sums over tiles and pixels)
  determinism                   bitwise, forward and backward
"""
import numpy as np
import pytest
import torch

from oracle import gscp_oracle as go

pytestmark = pytest.mark.gpu


def scene(n, W, H, seed, param, smin=0.5, smax=8, spread=(0.6, 0.45), rot_std=1.0, outside=0):
    """Parameters the way the reference model feeds the kernel
    (gscp/models/gaussian_field.py:303-340): sigma in (min, max) through a
    sigmoid, passed as log(sigma); Cholesky: log of the diagonal, free
    off-diagonal."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform([0, 0], [W - 1, H - 1], size=(n, 2))
    if outside:
        xy[:outside] += rng.choice([-1, 1], size=(outside, 2)) * rng.uniform(W / 2, W, size=(outside, 2))
    raw = rng.normal(spread[0], spread[1], size=(n, 2))
    sig = smin + (smax - smin) / (1 + np.exp(-raw)) * 0.5
    p0 = np.log(sig)
    if param == "rs":
        p1 = rng.normal(0, rot_std, size=(n, 1))
    else:
        p1 = rng.normal(0, 1.0, size=(n,)) * sig[:, 0] * 0.5
    w = rng.normal(size=(n, 2))
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return f(xy), f(p0), f(p1), f(w)


def run_gpu(pkg, dev, param, xy, p0, p1, w, H, W, smax, smin, dout=None):
    t = lambda a: torch.from_numpy(a).to(dev)
    fwd = pkg.gscp.rasterize_forward if param == "rs" else pkg.gscp.rasterize_forward_cholesky
    bwd = pkg.gscp.rasterize_backward if param == "rs" else pkg.gscp.rasterize_backward_cholesky
    R, out, radii, geom, binning, img = fwd(t(xy), t(p0), t(p1), t(w), H, W, smax, smin)
    res = dict(R=R, out=out.cpu().numpy(), radii=radii.cpu().numpy())
    if dout is not None:
        g = bwd(t(xy), t(p0), t(p1), t(w), radii, H, W, smax, smin, R, geom, binning, img, t(dout))
        res.update(dxy=g[0].cpu().numpy(), dp0=g[1].cpu().numpy(), dp1=g[2].cpu().numpy(),
                   dweight=g[3].cpu().numpy())
    return res


def radius_margin(geom):
    a, b, c = geom["cov"].astype(np.float64).T
    det = a * c - b * b
    mid = 0.5 * (a + c)
    val = 3 * np.sqrt(mid + np.sqrt(np.maximum(0.1, mid * mid - det)))
    return np.abs(val - np.round(val))


def check_radii(gpu_radii, geom):
    diff = gpu_radii != geom["radii"]
    assert not np.any(diff & (radius_margin(geom) > 1e-4)), np.nonzero(diff)[0][:10]
    assert diff.sum() <= max(1, len(gpu_radii) // 1000)


def check_canvas(out, ref, mag):
    err = np.abs(out - ref)
    bar = 1e-5 * mag + 1e-6
    assert np.all(err <= bar), (float((err / bar).max()), np.unravel_index(np.argmax(err / bar), err.shape))


def check_grad(name, g, ref):
    g = np.asarray(g, np.float64).reshape(ref.shape)
    scale = max(np.abs(ref).max(), 1e-30)
    err = np.abs(g - ref)
    bar = 1e-4 * scale + 1e-4 * np.abs(ref)
    assert np.all(err <= bar), (name, float((err / bar).max()), float(err.max()), scale)


CASES = [
    # param, n, W, H, max_patch_radius, min_scale, seed
    ("rs", 200, 96, 80, 8, 0.5, 0),
    ("cholesky", 200, 96, 80, 8, 0.5, 1),
    ("rs", 1500, 256, 256, 16, 1.0, 2),
    ("cholesky", 1500, 256, 256, 16, 1.0, 3),
    ("rs", 300, 70, 45, 8, 0.5, 4),        # ragged canvas, points outside it (clamped)
    ("cholesky", 64, 16, 16, 16, 1.0, 5),  # one tile, Gaussians larger than the canvas
]


@pytest.mark.parametrize("param,n,W,H,smax,smin,seed", CASES)
def test_forward_backward_match_oracle(pkg, dev, param, n, W, H, smax, smin, seed):
    xy, p0, p1, w = scene(n, W, H, seed, param, smin, smax, outside=n // 10 if W == 70 else 0)
    rng = np.random.default_rng(seed + 100)
    dout = rng.normal(size=(2, H, W)).astype(np.float32)
    gpu = run_gpu(pkg, dev, param, xy, p0, p1, w, H, W, smax, smin, dout)
    geom32 = (go.preprocess_rs if param == "rs" else go.preprocess_cholesky)(xy, p0, p1, W, H, smax, smin)
    check_radii(gpu["radii"], geom32)
    ref = go.rasterize(param, xy, p0, p1, w, H, W, smax, smin, dout=dout.astype(np.float64), radii=gpu["radii"])
    assert gpu["R"] == ref["num_rendered"]
    mag = go.render(ref["geom"], np.abs(w), ref["lists"])
    check_canvas(gpu["out"], ref["out"], np.abs(mag).sum(0, keepdims=True) * np.ones((2, 1, 1)))
    check_grad("dweight", gpu["dweight"], ref["dweight"])
    check_grad("dxy", gpu["dxy"], ref["dxy"])
    check_grad("dp0", gpu["dp0"], ref["dp0"])
    check_grad("dp1", gpu["dp1"], ref["dp1"].reshape(-1))
    print(f"{param} n={n} {W}x{H}: pairs {gpu['R']}, canvas max err {np.abs(gpu['out'] - ref['out']).max():.2e}")


def test_dense_tile_uses_the_long_list_sort(pkg, dev):
    """More than 8192 Gaussians in one tile: the list is ordered by counting
    ranks instead of the LDS sort; the result still follows the ascending
    index order (the canvas matches the oracle's sum in that order)."""
    n, W, H = 9000, 32, 32
    rng = np.random.default_rng(7)
    xy = rng.uniform(14, 18, size=(n, 2)).astype(np.float32)
    p0 = np.log(rng.uniform(0.6, 1.5, size=(n, 2))).astype(np.float32)
    p1 = rng.normal(0, 1, size=(n, 1)).astype(np.float32)
    w = (rng.normal(size=(n, 2)) / 30).astype(np.float32)
    dout = rng.normal(size=(2, H, W)).astype(np.float32)
    gpu = run_gpu(pkg, dev, "rs", xy, p0, p1, w, H, W, 8, 0.5, dout)
    ref = go.rasterize("rs", xy, p0, p1, w, H, W, 8, 0.5, dout=dout.astype(np.float64), radii=gpu["radii"])
    assert max(len(l) for l in ref["lists"]) > 8192
    mag = go.render(ref["geom"], np.abs(w), ref["lists"])
    check_canvas(gpu["out"], ref["out"], np.abs(mag).sum(0, keepdims=True) * np.ones((2, 1, 1)))
    check_grad("dweight", gpu["dweight"], ref["dweight"])
    check_grad("dxy", gpu["dxy"], ref["dxy"])


def test_empty_culled_and_clamped_inputs(pkg, dev):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
    # no Gaussians
    R, out, radii, geom, binning, img = pkg.gscp.rasterize_forward(
        t(np.zeros((0, 2))), t(np.zeros((0, 2))), t(np.zeros((0, 1))), t(np.zeros((0, 2))), 40, 24, 8, 0.5)
    assert R == 0 and radii.numel() == 0 and out.shape == (2, 40, 24) and not out.any()
    g = pkg.gscp.rasterize_backward(t(np.zeros((0, 2))), t(np.zeros((0, 2))), t(np.zeros((0, 1))),
                                    t(np.zeros((0, 2))), radii, 40, 24, 8, 0.5, R, geom, binning, img,
                                    torch.ones(2, 40, 24, device=dev))
    assert all(x.numel() == 0 for x in g)
    # a NaN rotation culls its Gaussian (NaN covariance); a NaN scale is
    # clamped to min_scale by fmaxf, as in the reference; an out-of-range
    # scale is clamped with a zero gradient
    xy = np.array([[10, 12], [30, 20], [50, 33], [20, 40]], np.float32)
    p0 = np.log(np.array([[2, 2], [1, 1], [100, 1.5], [np.nan, 0.5]], np.float32))
    p1 = np.array([[0.3], [np.nan], [0.2], [0.0]], np.float32)
    w = np.array([[1, 0.5], [2, 2], [-1, 0.25], [0.5, 0.5]], np.float32)
    dout = np.ones((2, 48, 64), np.float32)
    gpu = run_gpu(pkg, dev, "rs", xy, p0, p1, w, 48, 64, 8, 0.5, dout)
    assert gpu["radii"][1] == 0 and gpu["radii"][0] > 0 and gpu["radii"][2] > 0
    assert gpu["radii"][3] == int(np.ceil(3 * np.sqrt(0.25 + np.sqrt(np.float32(0.1)))))
    assert np.all(gpu["dxy"][1] == 0) and np.all(gpu["dweight"][1] == 0) and gpu["dp1"][1] == 0
    assert gpu["dp0"][2, 0] == 0 and gpu["dp0"][2, 1] != 0
    assert np.all(np.isfinite(gpu["out"]))
    ref = go.rasterize("rs", xy, p0, p1, w, 48, 64, 8, 0.5, dout=dout.astype(np.float64), radii=gpu["radii"])
    np.testing.assert_array_equal(ref["geom"]["radii"], gpu["radii"])
    np.testing.assert_allclose(gpu["out"], ref["out"], rtol=1e-5, atol=1e-6)
    # wrong shapes and dtypes are refused on the host, as the reference's AT_ERROR does
    with pytest.raises(RuntimeError, match="xy must have dimensions"):
        pkg.gscp.rasterize_forward(torch.zeros(4, 3, device=dev), t(np.zeros((4, 2))), t(np.zeros((4, 1))),
                                   t(np.zeros((4, 2))), 16, 16, 8, 0.5)
    with pytest.raises(RuntimeError, match="expected scalar type Float"):
        pkg.gscp.rasterize_forward(t(np.zeros((4, 2))), torch.zeros(4, 2, device=dev, dtype=torch.float64),
                                   t(np.zeros((4, 1))), t(np.zeros((4, 2))), 16, 16, 8, 0.5)


@pytest.mark.parametrize("param", ["rs", "cholesky"])
def test_autograd_dropin_and_determinism(pkg, dev, param):
    xy, p0, p1, w = scene(800, 200, 150, 11, param, 1.0, 16)
    tensors = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in (xy, p0, p1, w)]
    fn = pkg.gscp.GaussianRasterize2D if param == "rs" else pkg.gscp.GaussianRasterize2DCholesky
    target = torch.randn(2, 150, 200, device=dev, generator=torch.Generator(device=dev).manual_seed(3))
    outs, grads = [], []
    for _ in range(2):
        for x in tensors:
            x.grad = None
        out = fn.apply(*tensors, 150, 200, 16, 1.0)
        loss = ((out - target) ** 2).sum()
        loss.backward()
        outs.append(out.detach().clone())
        grads.append([x.grad.clone() for x in tensors])
    assert torch.equal(outs[0], outs[1])
    for a, b in zip(grads[0], grads[1]):
        assert torch.equal(a, b)
    for x, g in zip(tensors, grads[0]):
        assert g.shape == x.shape and torch.isfinite(g).all()
    # the autograd path equals the tuple API with dL/dout = 2 (out - target)
    dout = (2 * (outs[0] - target)).cpu().numpy()
    direct = run_gpu(pkg, dev, param, xy, p0, p1, w, 150, 200, 16, 1.0, dout)
    np.testing.assert_array_equal(direct["out"], outs[0].cpu().numpy())
    np.testing.assert_array_equal(direct["dweight"], grads[0][3].cpu().numpy())
    np.testing.assert_array_equal(direct["dxy"], grads[0][0].cpu().numpy())


def test_full_size_properties(pkg, dev):
    """50,000 Cholesky Gaussians on a 1024^2 canvas (the object count of the
    reference's simulate_and_reconstruct.py:170; max_patch_radius 16,
    min_scale 1): pair count and sampled tiles / Gaussians against the oracle,
    linearity in the weights, bitwise reruns."""
    n, W, H, smax, smin = 50000, 1024, 1024, 16, 1.0
    xy, p0, p1, w = scene(n, W, H, 21, "cholesky", smin, smax)
    rng = np.random.default_rng(22)
    dout = rng.normal(size=(2, H, W)).astype(np.float32)
    gpu = run_gpu(pkg, dev, "cholesky", xy, p0, p1, w, H, W, smax, smin, dout)
    gpu2 = run_gpu(pkg, dev, "cholesky", xy, p0, p1, w, H, W, smax, smin, dout)
    for k in ("out", "dxy", "dp0", "dp1", "dweight", "radii"):
        np.testing.assert_array_equal(gpu[k], gpu2[k])
    geom32 = go.preprocess_cholesky(xy, p0, p1, W, H, smax, smin)
    check_radii(gpu["radii"], geom32)
    geom = dict(geom32, radii=gpu["radii"].astype(np.int64))
    geom["rect"] = np.stack(go._rects(geom["means"][:, 0], geom["means"][:, 1], geom["radii"], W, H), axis=1)
    geom["touched"] = (geom["rect"][:, 3] - geom["rect"][:, 1]) * (geom["rect"][:, 2] - geom["rect"][:, 0])
    assert gpu["R"] == int(geom["touched"].sum())
    lists = go.tile_lists(geom)
    # sampled tiles: the full canvas of those tiles against the oracle
    gx, _ = go.tile_grid(W, H)
    tiles = rng.choice(len(lists), size=24, replace=False)
    sel = [l if t in set(tiles.tolist()) else [] for t, l in enumerate(lists)]
    ref = go.render(geom, w, sel)
    mag = go.render(geom, np.abs(w), sel).sum(0, keepdims=True) * np.ones((2, 1, 1))
    mask = np.zeros((H, W), bool)
    for t in tiles:
        ty, tx = divmod(int(t), gx)
        mask[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16] = True
    check_canvas(gpu["out"][:, mask], ref[:, mask], mag[:, mask])
    # sampled Gaussians: their gradients need only their own terms
    ids = set(rng.choice(n, size=64, replace=False).tolist())
    only = [[g for g in l if g in ids] for l in lists]
    dw, dm, dc = go.render_backward(geom, w, dout.astype(np.float64), only)
    dxy, dld, doff = go.preprocess_backward_cholesky(p0, p1, geom["radii"], dm, dc, smax, smin)
    idx = np.array(sorted(ids))
    check_grad("dweight", gpu["dweight"][idx], dw[idx])
    check_grad("dxy", gpu["dxy"][idx], dxy[idx])
    check_grad("dlog_L_diag", gpu["dp0"][idx], dld[idx])
    check_grad("dL_offdiag", gpu["dp1"][idx], doff[idx])
    # linearity in the weights (same lists, same order)
    w2 = rng.normal(size=(n, 2)).astype(np.float32)
    a = run_gpu(pkg, dev, "cholesky", xy, p0, p1, w2, H, W, smax, smin)["out"]
    b = run_gpu(pkg, dev, "cholesky", xy, p0, p1, (w + w2).astype(np.float32), H, W, smax, smin)["out"]
    np.testing.assert_allclose(b, gpu["out"] + a, rtol=0, atol=2e-5 * np.abs(b).max())
    print(f"full size: pairs {gpu['R']} ({gpu['R'] / n:.2f} tiles per Gaussian)")


def test_large_canvas_uses_the_global_scatter(pkg, dev):
    """A canvas of more than 16384 tiles (2112^2 = 132^2 tiles) bins through
    global slot counters instead of the per-chunk LDS counters; the canvas
    and the gradients still match the oracle."""
    n, W, H, smax, smin = 3000, 2112, 2112, 16, 1.0
    xy, p0, p1, w = scene(n, W, H, 31, "cholesky", smin, smax)
    rng = np.random.default_rng(32)
    dout = rng.normal(size=(2, H, W)).astype(np.float32)
    gpu = run_gpu(pkg, dev, "cholesky", xy, p0, p1, w, H, W, smax, smin, dout)
    ref = go.rasterize("cholesky", xy, p0, p1, w, H, W, smax, smin, dout=dout.astype(np.float64),
                       radii=gpu["radii"])
    assert go.tile_grid(W, H)[0] * go.tile_grid(W, H)[1] > 16384
    assert gpu["R"] == ref["num_rendered"]
    mag = go.render(ref["geom"], np.abs(w), ref["lists"])
    check_canvas(gpu["out"], ref["out"], np.abs(mag).sum(0, keepdims=True) * np.ones((2, 1, 1)))
    check_grad("dweight", gpu["dweight"], ref["dweight"])
    check_grad("dxy", gpu["dxy"], ref["dxy"])
    check_grad("dlog_L_diag", gpu["dp0"], ref["dp0"])
    check_grad("dL_offdiag", gpu["dp1"], ref["dp1"].reshape(-1))
