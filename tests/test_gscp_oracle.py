"""Pins of the Gaussian-field rasterizer oracle (oracle/gscp_oracle.py).

Each check is written independently of the oracle's own formulas:
covariances and conics from numpy matrices, coverage from pixel intervals,
a dense renderer, the bivariate normal density, and central differences.
"""
import numpy as np
import pytest

from oracle import gscp_oracle as go


def _scene(n, W, H, seed, param="rs", smin=0.5, smax=8, frac64=True):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(2, [W - 3, H - 3], size=(n, 2))
    if frac64:  # multiples of 1/64: every f32 and f64 sum of the rect stays exact
        xy = np.round(xy * 64) / 64
    xy = xy.astype(np.float32)
    if param == "rs":
        p0 = rng.uniform(np.log(smin * 1.3), np.log(smax / 2.0), size=(n, 2)).astype(np.float32)
        p1 = rng.uniform(-np.pi, np.pi, size=(n, 1)).astype(np.float32)
    else:
        p0 = rng.uniform(np.log(smin * 1.3), np.log(smax / 2.0), size=(n, 2)).astype(np.float32)
        p1 = rng.uniform(-1.5, 1.5, size=(n,)).astype(np.float32)
    w = rng.normal(size=(n, 2)).astype(np.float32)
    return xy, p0, p1, w


def test_rs_conic_is_inverse_covariance():
    xy, s, th, _ = _scene(64, 96, 80, 0)
    g = go.preprocess_rs(xy, s, th, 96, 80, 8, 0.5, dt=np.float64)
    for i in range(64):
        t = float(th[i, 0])
        R = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
        S = np.diag(np.clip(np.exp(s[i].astype(np.float64)), 0.5, 8.0) ** 2)
        cov = R @ S @ R.T
        inv = np.linalg.inv(cov)
        # conic = (inv[0,0], off-diagonal, inv[1,1]) with power = -0.5 x inv[0,0] ... - y dx dy
        np.testing.assert_allclose(g["conic"][i], [inv[0, 0], inv[0, 1], inv[1, 1]], rtol=1e-10, atol=1e-12)
        lam = np.linalg.eigvalsh(cov).max()
        assert g["radii"][i] >= 3 * np.sqrt(lam) - 1e-9
        assert g["radii"][i] <= 3 * np.sqrt(lam + 0.1) + 1


def test_cholesky_matches_rs_on_the_same_covariance():
    xy, s, th, _ = _scene(40, 64, 64, 1)
    g_rs = go.preprocess_rs(xy, s, th, 64, 64, 8, 0.5, dt=np.float64)
    a, b, c = g_rs["cov"].T
    l11 = np.sqrt(a)
    l21 = b / l11
    l22 = np.sqrt(c - l21 ** 2)
    g_ch = go.preprocess_cholesky(xy, np.stack([np.log(l11), np.log(l22)], 1), l21, 64, 64, 8, 0.5,
                                  dt=np.float64)
    L = np.zeros((40, 2, 2))
    L[:, 0, 0], L[:, 1, 0], L[:, 1, 1] = l11, l21, l22
    cov = L @ np.transpose(L, (0, 2, 1))
    np.testing.assert_allclose(cov[:, 0, 0], a, rtol=1e-12)
    np.testing.assert_allclose(cov[:, 1, 0], b, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g_ch["conic"], g_rs["conic"], rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(g_ch["radii"], g_rs["radii"])


@pytest.mark.parametrize("W,H", [(64, 64), (70, 45), (16, 16), (33, 17)])
def test_tile_lists_brute_force_and_coverage(W, H):
    xy, s, th, _ = _scene(50, W, H, 2)
    # some points on the canvas border and beyond it (clamped)
    xy[:3] = np.array([[0, 0], [W - 1, H - 1], [W + 5, -4]], np.float32)
    g = go.preprocess_rs(xy, s, th, W, H, 8, 0.5)
    lists = go.tile_lists(g)
    gx, gy = go.tile_grid(W, H)
    px = np.clip(xy[:, 0].astype(np.float64), 0, W - 1)
    py = np.clip(xy[:, 1].astype(np.float64), 0, H - 1)
    for i in range(50):
        r = int(g["radii"][i])
        assert r > 0
        # tile x-range: from the tile holding max(p - r, 0) to the tile holding
        # p + r + 15 (exclusive), written from pixel intervals
        x0 = min(gx, int(max(px[i] - r, 0.0) // 16))
        x1 = min(gx, int((px[i] + r + 15) // 16))
        y0 = min(gy, int(max(py[i] - r, 0.0) // 16))
        y1 = min(gy, int((py[i] + r + 15) // 16))
        want = {ty * gx + tx for ty in range(y0, y1) for tx in range(x0, x1)}
        got = {t for t, ids in enumerate(lists) if i in ids}
        assert got == want
        # every pixel within [p - r, p + r - 1] on both axes is in a listed tile
        for x in range(max(0, int(np.ceil(px[i] - r))), min(W, int(np.floor(px[i] + r - 1)) + 1)):
            for y in (int(py[i]),):
                assert (y // 16) * gx + (x // 16) in got
    for ids in lists:
        assert ids == sorted(ids)


def _dense_render(geom, w, lists):
    """Every Gaussian on every pixel of the canvas, masked by its listed tiles."""
    W, H = geom["W"], geom["H"]
    gx, _ = go.tile_grid(W, H)
    Y, X = np.mgrid[0:H, 0:W].astype(np.float64)
    tile_of = (Y // 16) * gx + (X // 16)
    out = np.zeros((2, H, W))
    for i in range(geom["n"]):
        tiles = [t for t, ids in enumerate(lists) if i in ids]
        mask = np.isin(tile_of, tiles)
        dx = geom["means"][i, 0] - X
        dy = geom["means"][i, 1] - Y
        cx, cy, cz = geom["conic"][i].astype(np.float64)
        q = cx * dx * dx + 2 * cy * dx * dy + cz * dy * dy
        G = np.exp(-0.5 * q)
        G = np.where(mask & (G >= 1e-7), G, 0.0)
        out[0] += w[i, 0] * G
        out[1] += w[i, 1] * G
    return out


@pytest.mark.parametrize("param", ["rs", "cholesky"])
def test_render_equals_dense_render(param):
    xy, p0, p1, w = _scene(30, 56, 40, 3, param)
    res = go.rasterize(param, xy, p0, p1, w, 40, 56, 8, 0.5)
    dense = _dense_render(res["geom"], w, res["lists"])
    np.testing.assert_allclose(res["out"], dense, rtol=1e-12, atol=1e-12)


def test_isolated_gaussian_is_the_bivariate_normal_density():
    xy = np.array([[31.25, 29.5]], np.float32)
    s = np.log(np.array([[2.5, 1.2]], np.float32))
    th = np.array([[0.7]], np.float32)
    w = np.array([[1.5, -0.25]], np.float32)
    res = go.rasterize("rs", xy, s, th, w, 64, 64, 8, 0.5)
    ex = np.exp(s[0].astype(np.float64))
    R = np.array([[np.cos(0.7), -np.sin(0.7)], [np.sin(0.7), np.cos(0.7)]])
    cov = R @ np.diag(ex ** 2) @ R.T
    Y, X = np.mgrid[0:64, 0:64].astype(np.float64)
    d = np.stack([X - 31.25, Y - 29.5], -1)
    q = np.einsum("...i,ij,...j->...", d, np.linalg.inv(cov), d)
    pdf = np.exp(-0.5 * q) / (2 * np.pi * np.sqrt(np.linalg.det(cov)))
    G = pdf * 2 * np.pi * np.sqrt(np.linalg.det(cov))
    covered = res["out"][0] != 0
    # the footprint reaches at least 3 sigma along both principal axes
    r = int(res["geom"]["radii"][0])
    assert covered[int(29.5), max(0, 31 - r + 1):31 + r - 1].all()
    np.testing.assert_allclose(res["out"][0][covered], 1.5 * G[covered], rtol=1e-5)
    np.testing.assert_allclose(res["out"][1][covered], -0.25 * G[covered], rtol=1e-5)
    assert G[~covered].max() < 0.05


@pytest.mark.parametrize("param", ["rs", "cholesky"])
def test_backward_matches_central_differences(param):
    W, H = 48, 40
    xy, p0, p1, w = _scene(12, W, H, 4, param, frac64=False)
    rng = np.random.default_rng(5)
    dout = rng.normal(size=(2, H, W))
    base = go.rasterize(param, xy, p0, p1, w, H, W, 8, 0.5, dout=dout, geom_dtype=np.float64)
    radii = base["geom"]["radii"]

    def loss(xy_, p0_, p1_, w_):
        r = go.rasterize(param, xy_, p0_, p1_, w_, H, W, 8, 0.5, radii=radii, geom_dtype=np.float64)
        return float(np.sum(r["out"] * dout))

    args = [xy.astype(np.float64), p0.astype(np.float64), p1.astype(np.float64), w.astype(np.float64)]
    grads = [base["dxy"], base["dp0"], base["dp1"].reshape(p1.shape), base["dweight"]]
    h = 1e-4
    for k in range(4):
        for idx in [(0, 0), (3, 1), (7, 0), (11, 1)]:
            idx = idx if args[k].ndim == 2 else (idx[0],)
            if args[k].ndim == 2 and args[k].shape[1] == 1:
                idx = (idx[0], 0)
            a_p = [a.copy() for a in args]
            a_m = [a.copy() for a in args]
            a_p[k][idx] += h
            a_m[k][idx] -= h
            fd = (loss(*a_p) - loss(*a_m)) / (2 * h)
            g = grads[k][idx]
            assert abs(fd - g) <= 2e-4 * max(1.0, abs(fd)), (param, k, idx, fd, g)


def test_backward_zero_where_scale_clamped_and_radius_zero():
    xy, s, th, w = _scene(6, 48, 48, 6)
    s[0, 0] = np.log(20.0)       # above max_patch_radius 8: clamped
    s[1, 1] = np.log(0.1)        # below min_scale 0.5: clamped
    dout = np.ones((2, 48, 48))
    res = go.rasterize("rs", xy, s, th, w, 48, 48, 8, 0.5, dout=dout, geom_dtype=np.float64)
    assert res["dp0"][0, 0] == 0 and res["dp0"][1, 1] == 0
    assert res["dp0"][0, 1] != 0 and res["dp0"][1, 0] != 0
    r0 = np.array(res["geom"]["radii"])
    r0[2] = 0
    res2 = go.rasterize("rs", xy, s, th, w, 48, 48, 8, 0.5, dout=dout, radii=r0, geom_dtype=np.float64)
    assert np.all(res2["dxy"][2] == 0) and np.all(res2["dp0"][2] == 0) and res2["dp1"][2] == 0
    assert np.all(res2["dweight"][2] == 0)


def test_nan_inputs_follow_fminf_fmaxf():
    """A NaN scale is clamped to min_scale (fmaxf returns the non-NaN operand);
    a NaN rotation gives a NaN covariance and culls the Gaussian."""
    xy = np.array([[8, 8], [20, 20]], np.float32)
    s = np.array([[np.nan, np.log(2.0)], [0.0, 0.0]], np.float32)
    th = np.array([[0.0], [np.nan]], np.float32)
    g = go.preprocess_rs(xy, s, th, 32, 32, 8, 0.5)
    ref = go.preprocess_rs(xy, np.array([[np.log(0.5), np.log(2.0)], [0, 0]], np.float32),
                           np.array([[0.0], [0.0]], np.float32), 32, 32, 8, 0.5)
    assert g["radii"][0] == ref["radii"][0] > 0
    assert g["radii"][1] == 0 and g["touched"][1] == 0
