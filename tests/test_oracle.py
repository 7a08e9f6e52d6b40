"""CPU checks that pin the oracle (no GPU).

The reference repository ships no tests, fixtures or data, so every function
of the oracle is pinned here by closed forms, library routines (numpy.fft,
torch.nn.functional), brute force on tiny inputs and invariants, each chosen
so that a plausible slip (dropped term, wrong sign, swapped axis) fails.
"""
import math
from fractions import Fraction

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import ptycho_oracle as orc
from tests.helpers import boundary_coords


# ------------------------------------------------------------------ level table
@pytest.mark.parametrize("log2T,total", [(14, 190344), (18, 2139016), (20, 6778280)])
def test_level_table_totals(log2T, total):
    t = orc.level_table(16, log2T, 16, 1.5)
    assert t["n_entries"] == total


def test_level_table_integer_scales_round_down():
    t = orc.level_table(16, 18, 16, 1.5)
    # exact scales 15, 23, 35, 53 are reached; level 4's exact 80 is 79.9999924 in f32
    assert [float(t["scale"][l]) for l in range(4)] == [15.0, 23.0, 35.0, 53.0]
    assert abs(float(t["scale"][4]) - 79.9999924) < 1e-6
    assert list(t["res"][:6]) == [16, 24, 36, 54, 81, 122]
    # dense up to level 8 with T = 2^18, hashed from level 9 on
    assert list(t["dense"]) == [True] * 9 + [False] * 7
    assert list(t["size"][:5]) == [256, 576, 1296, 2920, 6568]
    assert all(int(s) % 8 == 0 for s in t["size"])


def test_level_table_matches_product(pkg):
    for log2T in (12, 14, 18, 20):
        a = orc.level_table(16, log2T, 16, 1.5)
        b = pkg.level_table(16, 2, log2T, 16, 1.5)
        assert tuple(int(x) for x in a["res"]) == b.res
        assert tuple(int(x) for x in a["size"]) == b.size
        assert tuple(int(x) for x in a["offset"]) == b.offset
        assert tuple(bool(x) for x in a["dense"]) == b.dense
        assert tuple(np.float32(x) for x in a["scale"]) == tuple(np.float32(x) for x in b.scale)


# ------------------------------------------------------------------ coordinates
def test_grid_coords_layout():
    W, H = 6, 4
    c = orc.grid_coords(W, H)
    xs = torch.linspace(0.5 / W, 1 - 0.5 / W, W)
    ys = torch.linspace(0.5 / H, 1 - 0.5 / H, H)
    for i in range(W):
        for j in range(H):
            assert c[i * H + j, 0] == ys[j] and c[i * H + j, 1] == xs[i]


def test_grid_coords_match_product_model(pkg):
    m = pkg.ComplexINRModel2D(32, 32, 2, probe_width=16, probe_height=16, log2_hashmap_size=10)
    assert torch.equal(m.xy_coords.data, orc.grid_coords(16, 16))
    assert torch.equal(m.probe_coords.data, orc.grid_coords(16, 16))
    gp = m.grid_points("object")
    n0 = gp.height
    for n in (0, 5, 77, 255):
        i, j = divmod(n, n0)
        assert m.xy_coords[n, 0] == gp.ys[j] and m.xy_coords[n, 1] == gp.xs[i]


# ------------------------------------------------------------------ hash indices
def _f32_round_exact(q: Fraction) -> np.float32:
    """Round a rational to the nearest f32 (ties to even mantissa) by comparing
    the exact distances to the neighbouring f32 values."""
    c = np.float32(float(q))
    cands = [np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))]
    best = min(cands, key=lambda v: (abs(Fraction(float(v)) - q), int(np.array(v).view(np.uint32)) & 1))
    return np.float32(best)


def _fma_exact(x, s, c=0.5) -> np.float32:
    return _f32_round_exact(Fraction(float(np.float32(x))) * Fraction(float(np.float32(s)))
                            + Fraction(float(np.float32(c))))


def _brute_indices(x, table):
    """Pure-Python integer arithmetic per point (independent of the vectorised oracle)."""
    out = []
    for l in range(table["n_levels"]):
        s = np.float32(table["scale"][l])
        row = []
        for c in range(4):
            pos = [_fma_exact(x[d], s) for d in range(2)]
            g = [int(math.floor(float(p))) for p in pos]
            cx, cy = g[0] + (c & 1), g[1] + ((c >> 1) & 1)
            res, size = int(table["res"][l]), int(table["size"][l])
            if res * res <= size:
                idx = ((cx % 2**32 + (cy % 2**32) * res) % 2**32) % size
            else:
                idx = ((cx * 1) % 2**32 ^ (cy * 2654435761) % 2**32) % size
            row.append(int(table["offset"][l]) + idx)
        out.append(row)
    return out


def test_hash_indices_brute_force():
    table = orc.level_table(6, 9, 4, 2.0)   # levels 3.. are hashed (res^2 > 512)
    assert not all(table["dense"]) and any(table["dense"])
    rng = np.random.default_rng(0)
    pts = rng.random((40, 2)).astype(np.float32)
    idx, _ = orc.hash_corners(pts, table)
    for n in range(pts.shape[0]):
        assert idx[n].tolist() == _brute_indices(pts[n], table)


def test_fma_f32_matches_exact_rounding():
    """The oracle's single-rounding x*s + 0.5 against exact rationals, on random
    inputs and next to cell boundaries; and the two-rounding form differs
    from it on some of those inputs (so the choice is visible)."""
    table = orc.level_table(16, 18, 16, 1.5)
    rng = np.random.default_rng(5)
    xb = boundary_coords(table, per_level=6)
    xr = rng.random(300).astype(np.float32)
    n_diff = 0
    for l in range(16):
        s = np.float32(table["scale"][l])
        for x in (xr[:40], xb):
            got = orc.fma_f32(x, s, 0.5)
            want = np.array([_fma_exact(v, s) for v in x], np.float32)
            assert np.array_equal(got, want), l
        two = ((xb * s).astype(np.float32) + np.float32(0.5)).astype(np.float32)
        n_diff += int((two != orc.fma_f32(xb, s, 0.5)).sum())
    assert n_diff > 0
    # a double-rounding case: x*s = 2^-24 + 2^-60 (2^36 + 1 = 4097 * 16773121),
    # so x*s + 1 rounds in f64 onto the f32 midpoint 1 + 2^-24, whose tie goes
    # to 1.0, while the exact value lies above it: one rounding gives 1 + 2^-23
    x = np.float32(4097 * 2.0 ** -30)
    s = np.float32(16773121 * 2.0 ** -30)
    assert np.float32(np.float64(x) * np.float64(s) + 1.0) == np.float32(1.0)
    got = orc.fma_f32(np.array([x]), s, 1.0)[0]
    assert got == np.float32(1.0 + 2.0 ** -23) == _fma_exact(x, s, 1.0)


def test_hash_indices_brute_force_at_cell_boundaries():
    table = orc.level_table(16, 14, 16, 1.5)
    xb = boundary_coords(table, per_level=3, seed=3)
    rng = np.random.default_rng(4)
    pts = np.stack([xb, rng.permutation(xb)], 1).astype(np.float32)[:120]
    idx, _ = orc.hash_corners(pts, table)
    for n in range(pts.shape[0]):
        assert idx[n].tolist() == _brute_indices(pts[n], table)


def test_hash_indices_brute_force_outside_unit_square():
    """Inputs outside [0, 1]: the corners are uint32 (negative cells wrap
    modulo 2^32), which differs from a plain modulo when the level size does
    not divide 2^32."""
    table = orc.level_table(16, 14, 16, 1.5)
    rng = np.random.default_rng(6)
    pts = rng.uniform(-0.3, 1.3, (60, 2)).astype(np.float32)
    idx, _ = orc.hash_corners(pts, table)
    for n in range(pts.shape[0]):
        assert idx[n].tolist() == _brute_indices(pts[n], table)
    # level 1: scale 23, res 24, size 576 (dense); x = (-0.2, -0.2) lies in cell
    # -5 on both axes: (-5 - 5*24) mod 2^32 = 2^32 - 125, and 2^32 = 256 mod 576
    assert (table["scale"][1], table["res"][1], table["size"][1]) == (23.0, 24, 576)
    i, _ = orc.hash_corners(np.array([[-0.2, -0.2]], np.float32), table)
    assert i[0, 1, 0] == table["offset"][1] + 131
    assert (-125) % 576 == 451   # the plain modulo would give another entry


def test_hash_dense_edge_corner_wraps_into_next_row():
    table = orc.level_table(1, 14, 16, 1.5)   # level 0: scale 15, res 16, size 256
    # x0 near 1: pos0 = 15.5 -> g0 = 15 = res-1, the +1 corner is cx = 16 = res
    x = np.array([[0.999, 0.2]], np.float32)
    idx, frac = orc.hash_corners(x, table)
    g1 = int(np.floor(_fma_exact(0.2, 15.0)))
    assert idx[0, 0, 1] == (16 + g1 * 16) % 256        # wraps to column 0 of the next row
    assert idx[0, 0, 1] == (g1 + 1) * 16


def test_hash_prime_uint32_wrap():
    table = orc.level_table(16, 12, 16, 1.5)
    l = 15
    assert not table["dense"][l]
    x = np.array([[0.77, 0.93]], np.float32)
    idx, _ = orc.hash_corners(x, table)
    s = np.float32(table["scale"][l])
    gx = int(np.floor(_fma_exact(0.77, s)))
    gy = int(np.floor(_fma_exact(0.93, s)))
    # 2654435761 * gy overflows 32 bits for gy >= 2: the wrap is part of the hash
    assert gy * 2654435761 >= 2**32
    expect = ((gx ^ ((gy * 2654435761) & 0xFFFFFFFF)) & 0xFFFFFFFF) % table["size"][l]
    assert idx[0, l, 0] == table["offset"][l] + expect


# ------------------------------------------------------------------ hash encode
def test_encode_at_vertices_returns_entries():
    table = orc.level_table(2, 12, 4, 2.0)
    s0 = float(table["scale"][0])  # 3
    # pos = x*3 + 0.5 is an integer when x = (k - 0.5)/3 exactly representable? use k=2 -> x=0.5
    x = np.array([[0.5, 0.5]], np.float32)
    params = torch.randn(table["n_entries"] * 2, dtype=torch.float64)
    out = orc.hashgrid_encode(x, params, table)
    idx, frac = orc.hash_corners(x, table)
    assert frac[0, 0, 0] == 0 and frac[0, 0, 1] == 0 and s0 == 3.0
    e = params.reshape(-1, 2)[idx[0, 0, 0]]
    assert torch.allclose(out[0, 0:2], e, atol=0, rtol=0)


def test_encode_weights_sum_to_one_and_linear():
    table = orc.level_table(4, 10, 8, 1.7)
    rng = np.random.default_rng(1)
    x = rng.random((30, 2)).astype(np.float32)
    ones = torch.ones(table["n_entries"] * 2, dtype=torch.float64)
    out = orc.hashgrid_encode(x, ones, table)
    assert torch.allclose(out, torch.ones_like(out), atol=1e-12)
    t1 = torch.randn(table["n_entries"] * 2, dtype=torch.float64)
    t2 = torch.randn(table["n_entries"] * 2, dtype=torch.float64)
    lhs = orc.hashgrid_encode(x, 2 * t1 - 3 * t2, table)
    rhs = 2 * orc.hashgrid_encode(x, t1, table) - 3 * orc.hashgrid_encode(x, t2, table)
    assert torch.allclose(lhs, rhs, atol=1e-12)


def test_encode_adjoint_matches_explicit_scatter():
    table = orc.level_table(4, 9, 8, 1.7)
    rng = np.random.default_rng(2)
    x = rng.random((25, 2)).astype(np.float32)
    t = torch.randn(table["n_entries"] * 2, dtype=torch.float64, requires_grad=True)
    y = torch.randn(25, 8, dtype=torch.float64)
    (orc.hashgrid_encode(x, t, table) * y).sum().backward()
    idx, frac = orc.hash_corners(x, table)
    fx = frac[..., 0].astype(np.float64)
    fy = frac[..., 1].astype(np.float64)
    w = np.stack([(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy], -1)
    g = np.zeros((table["n_entries"], 2))
    yy = y.numpy().reshape(25, 4, 2)
    for n in range(25):
        for l in range(4):
            for c in range(4):
                g[idx[n, l, c]] += w[n, l, c] * yy[n, l]
    assert np.allclose(t.grad.numpy().reshape(-1, 2), g, atol=1e-12)


# ------------------------------------------------------------------ MLP
def _np_mlp(h, layers, wo, bo):
    z = h.astype(np.complex128)
    for (w, b, om) in layers:
        lin = z @ w.T + b
        z = np.exp(-om * lin.imag) * (np.cos(om * lin.real) + 1j * np.sin(om * lin.real))
    return z @ wo.T + bo


def test_complex_mlp_closed_form():
    rng = np.random.default_rng(3)
    h = rng.standard_normal((7, 6)) * 0.3
    W0 = rng.standard_normal((4, 6)) * 0.3
    b0 = rng.standard_normal(4) * 0.3
    A = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) * 0.3
    c = (rng.standard_normal(4) + 1j * rng.standard_normal(4)) * 0.3
    wo = rng.standard_normal((1, 4)) + 1j * rng.standard_normal((1, 4))
    bo = rng.standard_normal(1) + 1j * rng.standard_normal(1)
    layers_np = [(W0, b0, 10.0), (A, c, 1.3)]
    ref = _np_mlp(h, layers_np, wo, bo)
    t = lambda a: torch.from_numpy(np.asarray(a))
    out = orc.complex_mlp(t(h), [(t(W0), t(b0), torch.tensor([10.0], dtype=torch.float64)),
                                  (t(A), t(c), torch.tensor([1.3], dtype=torch.float64))], t(wo), t(bo))
    assert np.allclose(out.numpy(), ref, rtol=1e-12, atol=1e-12)


def test_gabor_magnitude_is_exp_of_minus_omega_imag():
    v = torch.randn(50, dtype=torch.complex128)
    I = torch.eye(50, dtype=torch.complex128)
    z = orc.gabor_layer(v, I, torch.zeros(50, dtype=torch.complex128), torch.tensor([2.5], dtype=torch.float64))
    assert torch.allclose(z.abs(), torch.exp(-2.5 * v.imag), rtol=1e-12)


def test_mlp_gradients_central_differences():
    torch.manual_seed(0)
    h = torch.randn(5, 6, dtype=torch.float64) * 0.3
    W0 = (torch.randn(4, 6, dtype=torch.float64) * 0.3).requires_grad_()
    om = torch.tensor([3.0], dtype=torch.float64, requires_grad=True)
    A = (torch.randn(4, 4, dtype=torch.complex128) * 0.3).requires_grad_()
    c = (torch.randn(4, dtype=torch.complex128) * 0.3)
    wo = torch.randn(1, 4, dtype=torch.complex128)
    bo = torch.zeros(1, dtype=torch.complex128)

    def f():
        o = orc.complex_mlp(h, [(W0, torch.zeros(4, dtype=torch.float64), om),
                                (A, c, torch.tensor([0.7], dtype=torch.float64))], wo, bo)
        return (o.abs() ** 2).sum()

    f().backward()
    eps = 1e-6
    for (p, idx) in [(W0, (1, 2)), (om, (0,)), (A, (2, 3))]:
        with torch.no_grad():
            if p.is_complex():
                for part, unit in (("re", 1.0), ("im", 1j)):
                    p[idx] += eps * unit
                    fp = f().item()
                    p[idx] -= 2 * eps * unit
                    fm = f().item()
                    p[idx] += eps * unit
                    fd = (fp - fm) / (2 * eps)
                    g = p.grad[idx].real if part == "re" else p.grad[idx].imag
                    assert abs(fd - g.item()) < 1e-6 * max(1.0, abs(fd))
            else:
                p[idx] += eps
                fp = f().item()
                p[idx] -= 2 * eps
                fm = f().item()
                p[idx] += eps
                fd = (fp - fm) / (2 * eps)
                assert abs(fd - p.grad[idx].item()) < 1e-6 * max(1.0, abs(fd))


# ------------------------------------------------------------------ upsample
def _np_upsample(x, ds):
    n = x.shape[0]
    m = n * ds
    X = np.fft.fft2(x)
    Y = np.zeros((m, m), complex)
    k = np.arange(n)
    q = np.where(k < n // 2, k, k - n + m)   # Nyquist n/2 lands on the negative side
    Y[np.ix_(q, q)] = X
    return np.fft.ifft2(Y) * ds * ds


@pytest.mark.parametrize("n", [4, 8, 16])
def test_fourier_upsample_numpy_and_even_samples(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    y = orc.fourier_upsample(torch.from_numpy(x)[None, None], 2)[0, 0].numpy()
    assert np.allclose(y, _np_upsample(x, 2), atol=1e-12)
    assert np.allclose(y[::2, ::2], x, atol=1e-12)


@pytest.mark.parametrize("n,ds", [(5, 2), (7, 3), (9, 2), (8, 3), (57, 2)])
def test_fourier_upsample_general_sizes(n, ds):
    """Odd sizes and ds > 2: fftshift/pad/ifftshift place small frequency k at
    q = ((k + n//2) % n + (m-n)//2 - m//2) mod m (for odd n the DC bin lands on
    m-1, a half-cell shift the kernels reproduce)."""
    rng = np.random.default_rng(n * 10 + ds)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    m = n * ds
    y = orc.fourier_upsample(torch.from_numpy(x)[None, None], ds)[0, 0].numpy()
    q = np.array([((k + n // 2) % n + (m - n) // 2 - m // 2) % m for k in range(n)])
    Y = np.zeros((m, m), complex)
    Y[np.ix_(q, q)] = np.fft.fft2(x)
    assert np.allclose(y, np.fft.ifft2(Y) * ds * ds, atol=1e-12)


def test_fourier_upsample_adjoint_formula():
    n, ds = 8, 2
    m = n * ds
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    g = torch.from_numpy(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
    xr = x.clone().requires_grad_(True)
    y = orc.fourier_upsample(xr[None, None], ds)[0, 0]
    (torch.real(torch.conj(g) * y)).sum().backward()   # Re<g, U x>: autograd gives U^H g
    # closed form used by the kernels: U^H g = ifft2_n(crop_q(fft2_m(g)))
    G = np.fft.fft2(g.numpy())
    k = np.arange(n)
    q = np.where(k < n // 2, k, k - n + m)
    adj = np.fft.ifft2(G[np.ix_(q, q)])
    assert np.allclose(xr.grad.numpy(), adj, atol=1e-10)


# ------------------------------------------------------------------ probe COM / normalisation
def test_probe_com_closed_form_and_energy():
    rng = np.random.default_rng(6)
    for _ in range(5):
        M, N = 16, 32
        p = torch.from_numpy(rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N)))
        p = p * torch.from_numpy(np.exp(-((np.arange(M)[:, None] - rng.uniform(2, 14)) ** 2) / 8.0))
        cp = orc.probe_com_shift(p)
        a = (p.abs() ** 2).numpy()
        E = a.sum()
        com_y = (np.arange(M) * a.sum(1)).sum() / E
        com_x = (np.arange(N) * a.sum(0)).sum() / E
        assert int(cp[0]) == math.trunc(com_y - M / 2 + 1)
        assert int(cp[1]) == math.trunc(com_x - N / 2 + 1)
        pn = orc.probe_center_normalize(p)
        assert abs((pn.abs() ** 2).sum().item() - M * N) < 1e-8 * M * N
        expect = torch.roll(p, (-int(cp[0]), -int(cp[1])), (0, 1))
        ratio = pn / expect
        assert torch.allclose(ratio, ratio[0, 0].expand_as(ratio), rtol=1e-10)


def test_probe_zero_energy_is_left_alone():
    p = torch.zeros(8, 8, dtype=torch.complex128)
    assert orc.probe_com_shift(p) is None
    assert torch.equal(orc.probe_center_normalize(p), p)


# ------------------------------------------------------------------ far field + losses
def test_farfield_parseval_and_numpy():
    rng = np.random.default_rng(7)
    obj = torch.from_numpy(rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40)))
    probe = torch.from_numpy(rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16)))
    I = orc.farfield_intensity(obj, probe, 5, 11)
    ew = probe.numpy() * obj.numpy()[5:21, 11:27]
    assert np.allclose(I.numpy(), np.abs(np.fft.fftshift(np.fft.fft2(ew))) ** 2, rtol=1e-12)
    assert abs(I.sum().item() - 256 * (np.abs(ew) ** 2).sum()) < 1e-9 * I.sum().item()


def _closed_form_dI(I, m, loss_type):
    """dL/dI of the per-pattern mean loss (formulas used by the far-field kernel)."""
    MN = I.size
    if loss_type == "smooth_L1_loss":
        return np.clip(I - m, -1, 1) / MN
    if loss_type == "Poisson_likelihood_loss":
        return (1 - m / (I + 1e-8)) / MN
    eps = 1e-8
    dW = np.sqrt(m + eps) - np.sqrt(I + eps)
    Rx = np.roll(dW, -1, 1) - dW
    Ry = np.roll(dW, -1, 0) - dW
    G = np.sqrt(Rx ** 2 + Ry ** 2 + eps)
    gx, gy = Rx / (G * MN), Ry / (G * MN)
    ddW = -gx - gy + np.roll(gx, 1, 1) + np.roll(gy, 1, 0)
    return -ddW / (2 * np.sqrt(I + eps))


@pytest.mark.parametrize("loss_type", ["smooth_L1_loss", "FD_loss", "Poisson_likelihood_loss"])
def test_loss_gradients_closed_form(loss_type):
    rng = np.random.default_rng(8)
    I = torch.from_numpy(rng.uniform(0.0, 5.0, (12, 10))).requires_grad_(True)
    m = torch.from_numpy(rng.poisson(2.0, (12, 10)).astype(np.float64))
    L = orc.pattern_loss(I, m, loss_type)
    L.backward()
    assert np.allclose(I.grad.numpy(), _closed_form_dI(I.detach().numpy(), m.numpy(), loss_type), atol=1e-12)


def test_smooth_l1_matches_definition():
    rng = np.random.default_rng(9)
    I = torch.from_numpy(rng.uniform(0, 4, 100))
    m = torch.from_numpy(rng.uniform(0, 4, 100))
    d = (I - m).numpy()
    ref = np.where(np.abs(d) < 1, 0.5 * d * d, np.abs(d) - 0.5).mean()
    assert abs(orc.pattern_loss(I, m, "smooth_L1_loss").item() - ref) < 1e-14


def test_fd_loss_is_mean_gradient_magnitude():
    rng = np.random.default_rng(10)
    I = rng.uniform(0, 4, (6, 5))
    m = rng.uniform(0, 4, (6, 5))
    dW = np.sqrt(m + 1e-8) - np.sqrt(I + 1e-8)
    tot = 0.0
    for y in range(6):
        for x in range(5):
            rx = dW[y, (x + 1) % 5] - dW[y, x]
            ry = dW[(y + 1) % 6, x] - dW[y, x]
            tot += math.sqrt(rx * rx + ry * ry + 1e-8)
    assert abs(orc.pattern_loss(torch.from_numpy(I), torch.from_numpy(m), "FD_loss").item() - tot / 30) < 1e-14


def test_batch_loss_central_difference_on_object():
    rng = np.random.default_rng(11)
    obj = torch.from_numpy(rng.standard_normal((24, 24)) + 1j * rng.standard_normal((24, 24))).requires_grad_(True)
    probe = torch.from_numpy(rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8)))
    stack = torch.from_numpy(rng.uniform(0, 30, (8, 8, 3)))
    tlY, tlX = np.array([0, 5, 16]), np.array([3, 9, 14])
    for lt in ("smooth_L1_loss", "FD_loss", "Poisson_likelihood_loss"):
        obj.grad = None
        L = orc.batch_loss(obj, probe, stack, [0, 1, 2], tlY, tlX, lt)
        L.backward()
        eps = 1e-6
        for (y, x) in [(2, 5), (7, 12), (20, 20)]:
            for unit, part in ((1.0, "re"), (1j, "im")):
                with torch.no_grad():
                    obj[y, x] += eps * unit
                    fp = orc.batch_loss(obj, probe, stack, [0, 1, 2], tlY, tlX, lt).item()
                    obj[y, x] -= 2 * eps * unit
                    fm = orc.batch_loss(obj, probe, stack, [0, 1, 2], tlY, tlX, lt).item()
                    obj[y, x] += eps * unit
                fd = (fp - fm) / (2 * eps)
                g = obj.grad[y, x].real if part == "re" else obj.grad[y, x].imag
                assert abs(fd - g.item()) < 1e-5 * max(1.0, abs(fd)), (lt, y, x, part)


# ------------------------------------------------------------------ Haar sparsity
def _haar_loops(img):
    """Per-block loops over a real 2-D array, written from the sub-band
    definitions of utils/loss.py:74-81 (a top-left, b top-right, c
    bottom-left, d bottom-right)."""
    tot = 0.0
    for r in range(img.shape[0] // 2):
        for s in range(img.shape[1] // 2):
            a, b = img[2 * r, 2 * s], img[2 * r, 2 * s + 1]
            c, d = img[2 * r + 1, 2 * s], img[2 * r + 1, 2 * s + 1]
            tot += abs(a - b + c - d) + abs(a + b - c - d) + abs(a - b - c + d)
    return tot


def test_haar_matches_block_loops_and_closed_forms():
    rng = np.random.default_rng(5)
    z = rng.standard_normal((6, 8)) + 1j * rng.standard_normal((6, 8))
    got = orc.haar_wavelet_sparsity_loss(torch.from_numpy(z)).item()
    assert got == pytest.approx(_haar_loops(z.real) + _haar_loops(z.imag), rel=1e-12)
    # constant image: no detail; impulse at a: |lh| = |hl| = |hh| = 1
    assert orc.haar_wavelet_sparsity_loss(torch.ones(4, 4, dtype=torch.float64)).item() == 0.0
    imp = torch.zeros(4, 4, dtype=torch.float64)
    imp[2, 2] = 1.0
    assert orc.haar_wavelet_sparsity_loss(imp).item() == 3.0
    # the impulse at d instead: same magnitudes (signs -,-,+)
    imp = torch.zeros(4, 4, dtype=torch.float64)
    imp[1, 3] = 2.0
    assert orc.haar_wavelet_sparsity_loss(imp).item() == 6.0
    # checkerboard (+1 on even parity): lh = hl = 0, hh = 4 per block
    cb = torch.from_numpy(np.indices((6, 4)).sum(0) % 2 == 0).to(torch.float64) * 2 - 1
    assert orc.haar_wavelet_sparsity_loss(cb).item() == 4.0 * 6
    # imaginary part counts like the real part
    assert orc.haar_wavelet_sparsity_loss(1j * cb.to(torch.complex128)).item() == 4.0 * 6
    # 'mean' divides each band's sum by the number of blocks
    x = torch.from_numpy(z)
    assert orc.haar_wavelet_sparsity_loss(x, "mean").item() == pytest.approx(got / 12, rel=1e-12)
    # a leading batch axis stacks images
    xb = torch.stack([x, 2 * x])
    assert orc.haar_wavelet_sparsity_loss(xb).item() == pytest.approx(3 * got, rel=1e-12)


def test_haar_bands_are_orthogonal():
    """(ll, lh, hl, hh) / 2 is an orthonormal transform of (a, b, c, d): with
    ll = a + b + c + d the energies agree, which fails for a wrong sign in any
    detail band."""
    rng = np.random.default_rng(6)
    img = torch.from_numpy(rng.standard_normal((8, 10)))
    a, b, c, d = img[0::2, 0::2], img[0::2, 1::2], img[1::2, 0::2], img[1::2, 1::2]
    ll = a + b + c + d
    lh, hl, hh = a - b + c - d, a + b - c - d, a - b - c + d
    e_bands = (ll ** 2 + lh ** 2 + hl ** 2 + hh ** 2).sum() / 4
    assert e_bands.item() == pytest.approx((img ** 2).sum().item(), rel=1e-12)
    # and the oracle's loss is the L1 of exactly these three detail bands
    l1 = (lh.abs().sum() + hl.abs().sum() + hh.abs().sum()).item()
    assert orc.haar_wavelet_sparsity_loss(img).item() == pytest.approx(l1, rel=1e-12)


def test_haar_gradient_central_differences():
    rng = np.random.default_rng(7)
    z = torch.from_numpy(rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))).requires_grad_(True)
    orc.haar_wavelet_sparsity_loss(z).backward()
    eps = 1e-6
    for (y, x) in [(0, 0), (1, 4), (5, 5), (3, 2)]:
        for unit in (1.0, 1j):
            zp = z.detach().clone()
            zp[y, x] += eps * unit
            zm = z.detach().clone()
            zm[y, x] -= eps * unit
            fd = (orc.haar_wavelet_sparsity_loss(zp) - orc.haar_wavelet_sparsity_loss(zm)).item() / (2 * eps)
            g = z.grad[y, x].real if unit == 1.0 else z.grad[y, x].imag
            assert abs(fd - g.item()) < 1e-6


# ------------------------------------------------------------------ geometry
def test_scan_positions_round_half_even():
    dx = 0.675e-6 * 0.0295 / (64 * 2.96e-5)
    # positions at 0.5 and 1.5 and 2.5 pixels: numpy rounds half to even -> 0, 2, 2
    x = np.array([0.0, 0.5, 1.5, 2.5, 10.0]) * dx
    y = np.array([0.0, 3.0, 3.5, 4.5, 0.0]) * dx
    tlX, tlY, brX, brY, d, canvas = orc.scan_positions_to_pixels(x, y, (64, 64), 0.675e-6, 0.0295, 2.96e-5)
    assert tlX.tolist() == [0, 0, 2, 2, 10]
    assert tlY.tolist() == [0, 3, 4, 4, 0]
    assert canvas == math.ceil(74 / 6) * 6 and canvas % 6 == 0
    assert np.allclose(d, dx)


def test_scan_positions_match_product(pkg):
    rng = np.random.default_rng(12)
    x, y = rng.uniform(0, 1e-4, 30), rng.uniform(0, 1e-4, 30)
    a = orc.scan_positions_to_pixels(x, y, (64, 64), 0.675e-6, 0.0295, 2.96e-5)
    b = pkg.scan_positions_to_pixels(x, y, (64, 64), 0.675e-6, 0.0295, 2.96e-5)
    for u, v in zip(a[:4], b[:4]):
        assert np.array_equal(u, v)
    assert a[5] == b[5]


# ------------------------------------------------------------------ Adam
def test_adam_update_matches_torch():
    torch.manual_seed(1)
    p0 = torch.randn(20, dtype=torch.float64)
    p_ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([p_ref], lr=1e-2)
    p = p0.clone()
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    for step in range(1, 6):
        g = torch.randn(20, dtype=torch.float64)
        p_ref.grad = g.clone()
        opt.step()
        orc.adam_update(p, g, m, v, step, 1e-2)
    assert torch.allclose(p, p_ref.detach(), rtol=1e-12, atol=1e-14)


# ------------------------------------------------------------------ synthetic data / metric
def test_synthetic_object_and_photons(pkg):
    inst = pkg.synthetic.make_named("small", seed=4)
    a = np.abs(inst.obj)
    assert abs(a.min() - 0.5) < 1e-9 and abs(a.max() - 1.0) < 1e-9
    ph = np.angle(inst.obj)
    assert ph.min() >= -np.pi / 2 - 1e-9 and ph.max() <= np.pi / 2 + 1e-9
    clean = pkg.synthetic.make_named("small", seed=4, noise=False)
    assert abs(clean.stack.sum(axis=(0, 1)).mean() - 1e3) < 1e-2
    # 1/f^2 power spectrum: radially averaged power falls ~ 2 decades per decade
    f = pkg.synthetic.gaussian_field_1f2(np.random.default_rng(0), 256)
    P = np.abs(np.fft.fft2(f)) ** 2
    fy, fx = np.meshgrid(np.fft.fftfreq(256), np.fft.fftfreq(256), indexing="ij")
    r = np.hypot(fx, fy)
    lo = P[(r > 0.02) & (r < 0.03)].mean()
    hi = P[(r > 0.2) & (r < 0.3)].mean()
    slope = np.log10(hi / lo) / np.log10(10)
    assert -2.3 < slope < -1.7


def test_aligned_nrmse_ambiguities(pkg):
    rng = np.random.default_rng(13)
    g = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    r = np.roll(g, (2, -1), axis=(0, 1)) * 0.3 * np.exp(1j * 1.1)
    assert pkg.metrics.aligned_nrmse(r, g, max_shift=3) < 1e-12
    noisy = g + 0.1 * (rng.standard_normal((32, 32)))
    e = pkg.metrics.aligned_nrmse(noisy, g)
    assert 0.05 < e < 0.1


# ------------------------------------------------------------------ warm-start net
def _loop_conv3(flat_w, bias, x, ci, co):
    """out[y, x, o] = b[o] + sum W[o][kh][kw][i] in[y + kh - 1][x + kw - 1][i] (NHWC, zero padding)."""
    H, W = x.shape[:2]
    w = flat_w.reshape(co, 3, 3, ci)
    out = np.zeros((H, W, co))
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    for y in range(H):
        for xx in range(W):
            patch = xp[y:y + 3, xx:xx + 3, :]
            out[y, xx] = np.tensordot(w, patch, axes=([1, 2, 3], [0, 1, 2])) + bias
    return out


def _loop_up(flat_w, bias, x, ci, co):
    """out[2y + a, 2x + c, o] = b[o] + sum_i W[a][c][o][i] in[y, x, i]."""
    H, W = x.shape[:2]
    w = flat_w.reshape(2, 2, co, ci)
    out = np.zeros((2 * H, 2 * W, co))
    for a in range(2):
        for c in range(2):
            out[a::2, c::2] = np.einsum("oi,yxi->yxo", w[a, c], x) + bias
    return out


def test_wsnet_layer_layouts_against_loops():
    """The oracle's unpacking (flat kernel layout -> torch layouts) against
    explicit loops written from the layout definition, one conv and one
    transposed conv."""
    rng = np.random.default_rng(0)
    base = 2
    layers = orc.wsnet_layers(base)
    n = sum((9 if k == "c3" else 4 if k == "up" else 1) * ci * co + co for k, ci, co in layers)
    flat = torch.from_numpy(rng.standard_normal(n))
    L = orc.wsnet_unpack(flat, base)
    # layer 1: conv3 c -> c
    w, b = L[1]
    off = 9 * 3 * base + base
    x = rng.standard_normal((6, 5, base))
    got = F.conv2d(torch.from_numpy(x).permute(2, 0, 1)[None], w, b, padding=1)[0].permute(1, 2, 0).numpy()
    want = _loop_conv3(flat[off:off + 9 * base * base].numpy(), flat[off + 9 * base * base:off + 9 * base * base + base]
                       .numpy(), x, base, base)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # layer 8: up 8c -> 4c
    off = sum((9 if k == "c3" else 4 if k == "up" else 1) * ci * co + co for k, ci, co in layers[:8])
    ci, co = 8 * base, 4 * base
    w, b = L[8]
    x = rng.standard_normal((3, 4, ci))
    got = F.conv_transpose2d(torch.from_numpy(x).permute(2, 0, 1)[None], w, b, stride=2)[0].permute(1, 2, 0).numpy()
    want = _loop_up(flat[off:off + 4 * ci * co].numpy(), flat[off + 4 * ci * co:off + 4 * ci * co + co].numpy(), x,
                    ci, co)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_wsnet_oracle_matches_torch_module(pkg):
    """The functional oracle = the torch module of the same architecture with
    the same weights (FP64), through the product's parameter packing."""
    ws = pkg.warmstart
    torch.manual_seed(1)
    mod = ws.WarmStartNet(base=4).double()
    flat = ws.pack_params(mod).double()
    x = torch.randn(2, 64, 64, 3, dtype=torch.float64)
    want = mod(x.permute(0, 3, 1, 2))
    got = orc.wsnet_forward(flat, x, 4)
    assert got.shape == (2, 2, 32, 32)
    assert torch.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_wsnet_oracle_gradient_central_differences():
    rng = np.random.default_rng(2)
    base = 2
    layers = orc.wsnet_layers(base)
    n = sum((9 if k == "c3" else 4 if k == "up" else 1) * ci * co + co for k, ci, co in layers)
    flat = torch.from_numpy(rng.standard_normal(n) * 0.3).requires_grad_(True)
    x = torch.from_numpy(rng.standard_normal((1, 16, 16, 3)))
    r = torch.from_numpy(rng.standard_normal((1, 2, 8, 8)))
    (orc.wsnet_forward(flat, x, base) * r).sum().backward()
    h = 1e-5
    for i in rng.choice(n, 12, replace=False):
        fp = flat.detach().clone()
        fm = flat.detach().clone()
        fp[i] += h
        fm[i] -= h
        d = ((orc.wsnet_forward(fp, x, base) * r).sum() - (orc.wsnet_forward(fm, x, base) * r).sum()) / (2 * h)
        assert abs(float(d) - float(flat.grad[i])) < 1e-6 * max(1.0, abs(float(d))), i
