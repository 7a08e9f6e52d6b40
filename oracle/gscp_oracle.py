"""CPU restatement of the reference's 2-D Gaussian-field rasterizer (the oracle).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; it is
used by ``tests/`` and the ``cpu_baseline`` leg of ``bench.py`` (Gaussian-field
mode) as the checker that the HIP rasterizer is compared against.

The reference rasterizer is CUDA (``gscp/csrc/gscp_rasterizer``); each
function below restates one of its stages, cited by file:line (paths relative
to the reference repository root):

  preprocess_rs / preprocess_cholesky   forward.cu:18-98 / :103-170
      per-Gaussian mean clamp, activated scales, covariance, conic, radius
      (legacy circular AABB, USE_SNUGBOX = 0 in config.h:18-20), tile rect
      (auxiliary.h:17-30) and tile count.  Computed in float32, one IEEE
      operation at a time in the reference's order, because the radius and
      the rect are integers derived from float arithmetic: the kernel must
      reproduce them exactly.
  tile_lists                            rasterizer_impl.cu:45-106, :212-240
      (duplicateWithKeys + stable radix sort on the tile bits +
      identifyTileRanges): for every 16x16 tile, the Gaussians (radius > 0)
      whose rect contains it, in ascending Gaussian index.
  render                                forward.cu:176-258 (renderCUDA):
      per pixel, sum over its tile's list of w * exp(power), skipping
      power > 0 and G < 1e-7.  float64 (or the dtype passed).
  render_backward                       backward.cu:21-133 (renderCUDA_backward):
      per-Gaussian sums of the weight, mean and conic gradients over the
      pixels of every tile in its list.
  preprocess_backward_rs / _cholesky    backward.cu:138-201 / :206-266:
      conic gradient -> covariance -> parameter gradients, including the
      reference's 1/(det^2 + 1e-7), the zero gradient where the activated
      scale is clamped, and the mean gradient passed through the position
      clamp unchanged.

Pinning (tests/test_gscp_oracle.py):
  * conic = inverse of the covariance built as R diag(s^2) R^T with numpy
    matrices (RS) or L L^T (Cholesky); both parameterisations agree on the
    same covariance;
  * radius >= 3 sqrt(largest eigenvalue) (numpy eigvalsh), and every pixel
    within [p - r, p + r - 1] on both axes lies in a listed tile;
  * tile lists against a brute-force rect/tile membership written from
    pixel intervals;
  * render against a dense renderer (every Gaussian on every pixel with an
    explicit coverage mask) and against the closed-form Gaussian integral
    for an isolated Gaussian;
  * render_backward + preprocess_backward against central differences of the
    FP64 render (rects held fixed, so the render is smooth in the parameters).
"""
from __future__ import annotations

import numpy as np

BLOCK = 16          # config.h:13-14 (BLOCK_X = BLOCK_Y = 16)
G_MIN = 1e-7        # forward.cu:243, backward.cu:88


def tile_grid(W: int, H: int):
    """Tiles along x and y (rasterizer_impl.cu:168)."""
    return (W + BLOCK - 1) // BLOCK, (H + BLOCK - 1) // BLOCK


def _rects(px, py, radii, W, H):
    """getRect (auxiliary.h:17-30), one operation at a time in the dtype of
    ``px`` (float32 for kernel parity)."""
    gx, gy = tile_grid(W, H)
    dt = px.dtype.type
    r = radii.astype(px.dtype)
    sixteen = dt(BLOCK)
    lo_x = ((px - r) / sixteen)
    lo_y = ((py - r) / sixteen)
    hi_x = (((px + r) + sixteen) - dt(1.0)) / sixteen
    hi_y = (((py + r) + sixteen) - dt(1.0)) / sixteen
    # (int) truncates toward zero; max(0, .) then min(grid, .)
    def clampi(v, g):
        return np.minimum(g, np.maximum(0, np.trunc(v).astype(np.int64)))
    return (clampi(lo_x, gx), clampi(lo_y, gy), clampi(hi_x, gx), clampi(hi_y, gy))


def _finish_geometry(px, py, a, b, c, det, W, H):
    """Conic, radius, rect and tile count from the covariance (forward.cu:64-97)."""
    n = px.shape[0]
    dt = px.dtype.type
    valid = det > dt(0.0)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        det_inv = dt(1.0) / det
        conic = np.stack([c * det_inv, -b * det_inv, a * det_inv], axis=1)
        mid = dt(0.5) * (a + c)
        lam = mid + np.sqrt(np.fmax(dt(0.1), mid * mid - det))
        rad_f = np.ceil(dt(3.0) * np.sqrt(lam))
    # NaN or non-positive -> culled; beyond the int range -> INT_MAX (the
    # saturating float-to-int conversion of the kernels)
    live = valid & (rad_f > 0)
    radii = np.where(live, np.minimum(np.where(live, rad_f, 0), 2147483520.0), 0).astype(np.int64)
    radii[live & (rad_f >= 2147483520.0)] = 2147483647
    rect = _rects(px, py, radii, W, H)
    touched = (rect[3] - rect[1]) * (rect[2] - rect[0])
    touched = np.where(radii > 0, touched, 0)
    means = np.stack([px, py], axis=1)
    conic = np.where((radii > 0)[:, None], conic, 0).astype(px.dtype)
    return dict(means=means, conic=conic, radii=radii, rect=np.stack(rect, axis=1),
                touched=touched.astype(np.int64), n=n, W=W, H=H,
                cov=np.stack([a, b, c], axis=1))


def _clamped_means(xy, W, H, dt):
    xy = np.asarray(xy, dtype=dt)
    # fminf / fmaxf semantics (a NaN operand yields the other one)
    px = np.fmin(dt(W - 1), np.fmax(dt(0.0), xy[:, 0]))
    py = np.fmin(dt(H - 1), np.fmax(dt(0.0), xy[:, 1]))
    return px, py


def preprocess_rs(xy, scaling, rotation, W, H, max_patch_radius, min_scale, dt=np.float32):
    """forward.cu:18-98 (preprocessCUDA); float32 by default (kernel parity),
    float64 for the central-difference pins."""
    px, py = _clamped_means(xy, W, H, dt)
    s = np.asarray(scaling, dtype=dt)
    th = np.asarray(rotation, dtype=dt).reshape(-1)
    smax, smin = dt(max_patch_radius), dt(min_scale)
    sx = np.fmin(smax, np.fmax(smin, np.exp(s[:, 0])))
    sy = np.fmin(smax, np.fmax(smin, np.exp(s[:, 1])))
    sx2, sy2 = sx * sx, sy * sy
    ct, st = np.cos(th), np.sin(th)
    a = ct * ct * sx2 + st * st * sy2
    b = ct * st * (sx2 - sy2)
    c = st * st * sx2 + ct * ct * sy2
    det = a * c - b * b
    return _finish_geometry(px, py, a, b, c, det, W, H)


def preprocess_cholesky(xy, log_L_diag, L_offdiag, W, H, max_patch_radius, min_scale, dt=np.float32):
    """forward.cu:103-170 (preprocessCUDA_cholesky); float32 by default."""
    px, py = _clamped_means(xy, W, H, dt)
    ld = np.asarray(log_L_diag, dtype=dt)
    l21 = np.asarray(L_offdiag, dtype=dt).reshape(-1)
    smax, smin = dt(max_patch_radius), dt(min_scale)
    l11 = np.fmin(smax, np.fmax(smin, np.exp(ld[:, 0])))
    l22 = np.fmin(smax, np.fmax(smin, np.exp(ld[:, 1])))
    a = l11 * l11
    b = l11 * l21
    c = l21 * l21 + l22 * l22
    det = a * l22 * l22
    return _finish_geometry(px, py, a, b, c, det, W, H)


def tile_lists(geom, radii=None):
    """Per tile, the ascending Gaussian ids whose rect holds it
    (rasterizer_impl.cu:45-106 and the stable sort on the tile bits at :217-225).
    ``radii`` (optional) overrides the geometry's radii, as the reference's
    duplicateWithKeys reads the radii buffer (rasterizer_impl.cu:58)."""
    W, H = geom["W"], geom["H"]
    gx, gy = tile_grid(W, H)
    if radii is None:
        radii = geom["radii"]
        rect = geom["rect"]
    else:
        radii = np.asarray(radii, dtype=np.int64)
        rect = np.stack(_rects(geom["means"][:, 0], geom["means"][:, 1], radii, W, H), axis=1)
    lists = [[] for _ in range(gx * gy)]
    for g in np.nonzero(radii > 0)[0]:
        x0, y0, x1, y1 = rect[g]
        for ty in range(y0, y1):
            for tx in range(x0, x1):
                lists[ty * gx + tx].append(int(g))
    return lists


def render(geom, weights, lists=None, dtype=np.float64):
    """forward.cu:176-258: out[2, H, W]; sums in list order, in ``dtype``."""
    W, H = geom["W"], geom["H"]
    gx, gy = tile_grid(W, H)
    if lists is None:
        lists = tile_lists(geom)
    means = geom["means"].astype(dtype)
    con = geom["conic"].astype(dtype)
    w = np.asarray(weights, dtype=dtype)
    out = np.zeros((2, H, W), dtype=dtype)
    for t, ids in enumerate(lists):
        if not ids:
            continue
        ty, tx = divmod(t, gx)
        ys = np.arange(ty * BLOCK, min(H, ty * BLOCK + BLOCK))
        xs = np.arange(tx * BLOCK, min(W, tx * BLOCK + BLOCK))
        X, Y = np.meshgrid(xs.astype(dtype), ys.astype(dtype))
        acc0 = np.zeros(X.shape, dtype=dtype)
        acc1 = np.zeros(X.shape, dtype=dtype)
        for g in ids:
            dx = means[g, 0] - X
            dy = means[g, 1] - Y
            power = -0.5 * (con[g, 0] * dx * dx + con[g, 2] * dy * dy) - con[g, 1] * dx * dy
            G = np.exp(np.minimum(power, 0.0))
            keep = (power <= 0) & (G >= G_MIN)
            acc0 += np.where(keep, w[g, 0] * G, 0.0)
            acc1 += np.where(keep, w[g, 1] * G, 0.0)
        out[0, ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1] = acc0
        out[1, ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1] = acc1
    return out


def render_backward(geom, weights, dout, lists=None, dtype=np.float64):
    """backward.cu:21-133: per-Gaussian dL/dweight [N,2], dL/dmean [N,2],
    dL/dconic [N,3] (conic order: x = Sigma^-1[0,0], y = off-diagonal,
    z = Sigma^-1[1,1] with power = -0.5(x dx^2 + z dy^2) - y dx dy)."""
    W, H = geom["W"], geom["H"]
    gx, gy = tile_grid(W, H)
    if lists is None:
        lists = tile_lists(geom)
    n = geom["n"]
    means = geom["means"].astype(dtype)
    con = geom["conic"].astype(dtype)
    w = np.asarray(weights, dtype=dtype)
    dout = np.asarray(dout, dtype=dtype)
    dw = np.zeros((n, 2), dtype=dtype)
    dm = np.zeros((n, 2), dtype=dtype)
    dc = np.zeros((n, 3), dtype=dtype)
    for t, ids in enumerate(lists):
        if not ids:
            continue
        ty, tx = divmod(t, gx)
        ys = np.arange(ty * BLOCK, min(H, ty * BLOCK + BLOCK))
        xs = np.arange(tx * BLOCK, min(W, tx * BLOCK + BLOCK))
        X, Y = np.meshgrid(xs.astype(dtype), ys.astype(dtype))
        d0 = dout[0, ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1]
        d1 = dout[1, ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1]
        for g in ids:
            dx = means[g, 0] - X
            dy = means[g, 1] - Y
            power = -0.5 * (con[g, 0] * dx * dx + con[g, 2] * dy * dy) - con[g, 1] * dx * dy
            G = np.exp(np.minimum(power, 0.0))
            G = np.where((power <= 0) & (G >= G_MIN), G, 0.0)
            dLdG = w[g, 0] * d0 + w[g, 1] * d1
            dw[g, 0] += np.sum(G * d0)
            dw[g, 1] += np.sum(G * d1)
            dm[g, 0] += np.sum(dLdG * G * (-con[g, 0] * dx - con[g, 1] * dy))
            dm[g, 1] += np.sum(dLdG * G * (-con[g, 2] * dy - con[g, 1] * dx))
            dc[g, 0] += np.sum(-0.5 * G * dx * dx * dLdG)
            dc[g, 1] += np.sum(-G * dx * dy * dLdG)
            dc[g, 2] += np.sum(-0.5 * G * dy * dy * dLdG)
    return dw, dm, dc


def _dcov_from_dconic(a, b, c, dc):
    """backward.cu:174-181 (conic -> covariance, with the reference's 1e-7)."""
    det = a * c - b * b
    inv = 1.0 / (det * det + 1e-7)
    da = inv * (-c * c * dc[:, 0] + b * c * dc[:, 1] + (det - a * c) * dc[:, 2])
    db = inv * (2.0 * b * c * dc[:, 0] - (det + 2.0 * b * b) * dc[:, 1] + 2.0 * a * b * dc[:, 2])
    dcc = inv * ((det - a * c) * dc[:, 0] + a * b * dc[:, 1] - a * a * dc[:, 2])
    return da, db, dcc


def preprocess_backward_rs(scaling, rotation, radii, dm, dc, max_patch_radius, min_scale):
    """backward.cu:138-201: (dxy, dscaling, drotation); zero where radius <= 0."""
    s = np.asarray(scaling, dtype=np.float64)
    th = np.asarray(rotation, dtype=np.float64).reshape(-1)
    smax, smin = float(max_patch_radius), float(min_scale)
    sxr, syr = np.exp(s[:, 0]), np.exp(s[:, 1])
    sx, sy = np.fmin(smax, np.fmax(smin, sxr)), np.fmin(smax, np.fmax(smin, syr))
    sx2, sy2 = sx * sx, sy * sy
    ct, st = np.cos(th), np.sin(th)
    a = ct * ct * sx2 + st * st * sy2
    b = ct * st * (sx2 - sy2)
    c = st * st * sx2 + ct * ct * sy2
    da, db, dcc = _dcov_from_dconic(a, b, c, dc)
    dsx2 = ct * ct * da + ct * st * db + st * st * dcc
    dsy2 = st * st * da - ct * st * db + ct * ct * dcc
    dth = (2.0 * ct * st * (sy2 - sx2) * da + (ct * ct - st * st) * (sx2 - sy2) * db
           + 2.0 * st * ct * (sx2 - sy2) * dcc)
    inx = ((sxr >= smin) & (sxr <= smax)).astype(np.float64)
    iny = ((syr >= smin) & (syr <= smax)).astype(np.float64)
    live = (np.asarray(radii) > 0).astype(np.float64)
    dscal = np.stack([dsx2 * 2.0 * sx2 * inx, dsy2 * 2.0 * sy2 * iny], axis=1) * live[:, None]
    drot = (dth * live)[:, None]
    dxy = dm * live[:, None]
    return dxy, dscal, drot


def preprocess_backward_cholesky(log_L_diag, L_offdiag, radii, dm, dc, max_patch_radius, min_scale):
    """backward.cu:206-266: (dxy, dlog_L_diag, dL_offdiag); zero where radius <= 0."""
    ld = np.asarray(log_L_diag, dtype=np.float64)
    l21 = np.asarray(L_offdiag, dtype=np.float64).reshape(-1)
    smax, smin = float(max_patch_radius), float(min_scale)
    r11, r22 = np.exp(ld[:, 0]), np.exp(ld[:, 1])
    l11, l22 = np.fmin(smax, np.fmax(smin, r11)), np.fmin(smax, np.fmax(smin, r22))
    a = l11 * l11
    b = l11 * l21
    c = l21 * l21 + l22 * l22
    # det = a * l22^2 in the reference (backward.cu:237), equal to a c - b^2
    det = a * l22 * l22
    inv = 1.0 / (det * det + 1e-7)
    da = inv * (-c * c * dc[:, 0] + b * c * dc[:, 1] + (det - a * c) * dc[:, 2])
    db = inv * (2.0 * b * c * dc[:, 0] - (det + 2.0 * b * b) * dc[:, 1] + 2.0 * a * b * dc[:, 2])
    dcc = inv * ((det - a * c) * dc[:, 0] + a * b * dc[:, 1] - a * a * dc[:, 2])
    dl11 = 2.0 * l11 * da + l21 * db
    dl21 = l11 * db + 2.0 * l21 * dcc
    dl22 = 2.0 * l22 * dcc
    i11 = ((r11 >= smin) & (r11 <= smax)).astype(np.float64)
    i22 = ((r22 >= smin) & (r22 <= smax)).astype(np.float64)
    live = (np.asarray(radii) > 0).astype(np.float64)
    dld = np.stack([dl11 * l11 * i11, dl22 * l22 * i22], axis=1) * live[:, None]
    doff = dl21 * live
    dxy = dm * live[:, None]
    return dxy, dld, doff


def rasterize(param, xy, p0, p1, weights, H, W, max_patch_radius, min_scale, dout=None,
              radii=None, dtype=np.float64, geom_dtype=np.float32):
    """Whole forward (and backward when ``dout`` is given), the reference's
    ``_C.rasterize_forward[_cholesky]`` / ``rasterize_backward[_cholesky]``.
    ``param`` is "rs" or "cholesky"; ``radii`` (optional) replaces the computed
    radii (the reference's duplicateWithKeys and backward read the radii
    buffer); ``geom_dtype`` is float32 for kernel parity, float64 for the
    central-difference pins.  Returns a dict."""
    pre = preprocess_rs if param == "rs" else preprocess_cholesky
    geom = pre(xy, p0, p1, W, H, max_patch_radius, min_scale, dt=geom_dtype)
    if radii is not None:
        radii = np.asarray(radii, dtype=np.int64)
        geom = dict(geom, radii=radii,
                    rect=np.stack(_rects(geom["means"][:, 0], geom["means"][:, 1], radii, W, H), axis=1))
        geom["touched"] = np.where(radii > 0, (geom["rect"][:, 3] - geom["rect"][:, 1])
                                   * (geom["rect"][:, 2] - geom["rect"][:, 0]), 0)
    lists = tile_lists(geom)
    out = render(geom, weights, lists, dtype=dtype)
    res = dict(geom=geom, lists=lists, out=out, num_rendered=int(geom["touched"].sum()))
    if dout is not None:
        dw, dm, dc = render_backward(geom, weights, dout, lists, dtype=dtype)
        bwd = preprocess_backward_rs if param == "rs" else preprocess_backward_cholesky
        dxy, d0, d1 = bwd(p0, p1, geom["radii"], dm, dc, max_patch_radius, min_scale)
        res.update(dxy=dxy, dp0=d0, dp1=d1, dweight=dw, dmean=dm, dconic=dc)
    return res
