"""GPU parity: every HIP entry point against the FP64 CPU oracle, through the C ABI.

Tolerances (FP32 kernels vs FP64 oracle on the same inputs):
  hash indices: bit-exact; features 1e-6 relative (norm);
  far-field intensities: 1e-5 relative per pattern (max-norm);
  fields / gradients: stated per test (FP32 accumulation of long sums).
"""
import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as orc
from tests.helpers import boundary_coords, oracle_params, oracle_table, rel_err, step_reference

pytestmark = pytest.mark.gpu


def _model(pkg, dev, canvas, M, log2T=12, residual=True, probe_init=None, seed=0, hidden_layers=2,
           table_precision="fp32", mlp_precision="fp32", first_omega=10.0, hidden_omega=1.0):
    torch.manual_seed(seed)
    init = probe_init if probe_init is not None else torch.ones(M, M, dtype=torch.complex64)
    return pkg.ComplexINRModel2D(canvas, canvas, 2, probe_width=M, probe_height=M, log2_hashmap_size=log2T,
                                 first_omega_0=first_omega, hidden_omega_0=hidden_omega, trainable_omega0=True,
                                 use_residual=residual,
                                 probe_initial=init if residual else None, hidden_layers=hidden_layers,
                                 table_precision=table_precision, mlp_precision=mlp_precision).to(dev)


def _round_tables_to_fp16(params):
    """Oracle leaves holding the tables as the FP16 kernels see them."""
    for key in ("obj_table", "prb_table"):
        t = params[key].detach().to(torch.float16).to(torch.float64)
        params[key] = t.clone().requires_grad_(True)
    return params


# ---------------------------------------------------------------- hash grid
@pytest.mark.parametrize("log2T", [14, 18, 20])
def test_hash_indices_bit_exact(pkg, dev, log2T):
    """Corner indices bit-exact against the oracle: random points, a regular
    grid, points within a few ulps of cell boundaries on every level (where
    the single rounding of x*scale + 0.5 decides the cell) and, for the
    canvas-3648 table (log2T 20), a row subsample of its 1824^2 INR grid."""
    spec = pkg.level_table(16, 2, log2T, 16, 1.5)
    table = orc.level_table(16, log2T, 16, 1.5)
    rng = np.random.default_rng(log2T)
    xb = boundary_coords(table, per_level=20, seed=log2T)
    parts = [rng.random((4000, 2)), orc.grid_coords(128, 128).numpy()[::7],
             np.stack([xb, rng.permutation(xb)], 1), np.stack([rng.random(xb.size), xb], 1),
             np.array([[0.0, 0.0], [0.999999, 0.999999], [0.5, 0.25]]),
             rng.uniform(-0.3, 1.3, (2000, 2))]   # outside [0, 1]: uint32 corner wrap
    if log2T == 20:
        g = orc.grid_coords(1824, 1824).numpy().reshape(1824, 1824, 2)
        parts.append(g[::97].reshape(-1, 2))
    pts = np.concatenate(parts).astype(np.float32)
    idx = pkg.hashgrid_indices(torch.from_numpy(pts).to(dev), spec).cpu().numpy()
    ref, _ = orc.hash_corners(pts, table)
    assert np.array_equal(idx, ref)


def test_hash_encode_fwd_bwd(pkg, dev):
    torch.manual_seed(0)
    enc = pkg.HashGridEncoding(2, {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2,
                                   "log2_hashmap_size": 14, "base_resolution": 16, "per_level_scale": 1.5}).to(dev)
    with torch.no_grad():
        enc.params.normal_()
    x = torch.rand(3000, 2, device=dev)
    out = enc(x)
    table = orc.level_table(16, 14, 16, 1.5)
    tp = enc.params.detach().cpu().double().requires_grad_(True)
    ref = orc.hashgrid_encode(x.cpu().numpy(), tp, table)
    assert rel_err(out, ref) < 1e-6
    y = torch.randn_like(out)
    (out * y).sum().backward()
    (ref * y.cpu().double()).sum().backward()
    assert rel_err(enc.params.grad, tp.grad) < 1e-5


# ---------------------------------------------------------------- fused INR field
@pytest.mark.parametrize("hidden_layers", [1, 2])
def test_inr_field_fwd_bwd(pkg, dev, hidden_layers):
    model = _model(pkg, dev, 64, 16, log2T=12, residual=False, hidden_layers=hidden_layers)
    with torch.no_grad():   # features of order one so that every layer matters
        model.xy_hash_encoder.params.uniform_(-1, 1)
    pts = model.grid_points("object")
    f = model._field(model.xy_hash_encoder, model.object_complex_mlp, "object")
    params = oracle_params(model)
    table = oracle_table(model)
    ref = orc.inr_field(model.xy_coords.cpu(), params["obj_table"], table, params["obj_mlp"], 32, 32)
    assert rel_err(f.reshape(32, 32), ref) < 2e-6
    g = torch.randn(f.shape, dtype=torch.complex64, device=dev)
    (torch.real(torch.conj(g) * f)).sum().backward()
    (torch.real(torch.conj(g.cpu().to(torch.complex128)) * ref.reshape(-1))).sum().backward()
    assert rel_err(model.xy_hash_encoder.params.grad, params["obj_table"].grad) < 1e-4
    mlp = params["obj_mlp"]
    ours = model.object_complex_mlp.tensors()
    theirs = []
    for (w, b, om) in mlp["layers"]:
        theirs += [w, b, om]
    theirs += [mlp["out_weight"], mlp["out_bias"]]
    for a, b in zip(ours, theirs):
        assert rel_err(a.grad, b.grad) < 1e-4, (a.shape, rel_err(a.grad, b.grad))


def test_inr_field_fp16_table(pkg, dev):
    """FP16 table storage: the kernels gather half entries and interpolate in
    f32, so against the oracle evaluated at the half-rounded table the field
    and every gradient keep the FP32 bars; against the unrounded table the
    difference is of the order of half rounding."""
    model = _model(pkg, dev, 64, 16, log2T=12, residual=False, table_precision="fp16")
    with torch.no_grad():
        model.xy_hash_encoder.params.uniform_(-1, 1)
    f = model._field(model.xy_hash_encoder, model.object_complex_mlp, "object")
    params = oracle_params(model)
    table = oracle_table(model)
    exact = orc.inr_field(model.xy_coords.cpu(), params["obj_table"].detach(), table, params["obj_mlp"], 32, 32)
    params = _round_tables_to_fp16(params)
    ref = orc.inr_field(model.xy_coords.cpu(), params["obj_table"], table, params["obj_mlp"], 32, 32)
    assert rel_err(f.reshape(32, 32), ref) < 2e-6
    assert 1e-5 < rel_err(f.reshape(32, 32), exact) < 1e-2
    g = torch.randn(f.shape, dtype=torch.complex64, device=dev)
    (torch.real(torch.conj(g) * f)).sum().backward()
    (torch.real(torch.conj(g.cpu().to(torch.complex128)) * ref.reshape(-1))).sum().backward()
    assert model.xy_hash_encoder.params.grad.dtype == torch.float32
    assert rel_err(model.xy_hash_encoder.params.grad, params["obj_table"].grad) < 1e-4
    mlp = params["obj_mlp"]
    theirs = []
    for (w, b, om) in mlp["layers"]:
        theirs += [w, b, om]
    theirs += [mlp["out_weight"], mlp["out_bias"]]
    for a, b in zip(model.object_complex_mlp.tensors(), theirs):
        assert rel_err(a.grad, b.grad) < 1e-4, (a.shape, rel_err(a.grad, b.grad))


def test_hash_encode_fp16_table(pkg, dev):
    enc = pkg.HashGridEncoding(2, {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2,
                                   "log2_hashmap_size": 14, "base_resolution": 16, "per_level_scale": 1.5},
                               dtype=torch.float16).to(dev)
    assert enc.table_precision == "fp16" and enc.params.dtype == torch.float32
    with torch.no_grad():
        enc.params.uniform_(-1, 1)
    x = torch.rand(3000, 2, generator=torch.Generator().manual_seed(4))
    out = enc(x.to(dev))
    table = orc.level_table(16, 14, 16, 1.5)
    rounded = enc.params.detach().cpu().to(torch.float16).to(torch.float64)
    ref = orc.hashgrid_encode(x.numpy(), rounded, table)     # f32 coordinates, as the kernel sees them
    assert rel_err(out, ref) < 1e-6


def test_mlp_forward_from_features(pkg, dev):
    model = _model(pkg, dev, 64, 16, residual=False)
    feats = torch.randn(777, 32, device=dev)
    with torch.no_grad():
        out = model.object_complex_mlp.net[0].linear.weight  # noqa: F841 (touch)
        got = pkg.mlp_forward(model.object_complex_mlp, feats)
    params = oracle_params(model)["obj_mlp"]
    ref = orc.complex_mlp(feats.cpu().double(), params["layers"], params["out_weight"], params["out_bias"])
    assert rel_err(got, ref) < 2e-6


# ---------------------------------------------------------------- upsample
@pytest.mark.parametrize("n,ds", [(16, 2), (32, 2), (64, 2), (128, 2), (256, 2), (512, 2), (1024, 2), (64, 3),
                                  (5, 2), (7, 3), (57, 2), (114, 2), (513, 2), (912, 2), (1824, 2)])
def test_upsample_fwd_adj(pkg, dev, n, ds):
    """Power-of-two (64 .. 512 at ds 2 on the wavefront register FFT, the rest on the LDS
    Stockham kernels) and mixed-radix sizes (57 = 3*19, 1824 = 2^5*3*19), odd n (DC on m-1)."""
    torch.manual_seed(n)
    m = n * ds
    x = torch.randn(n, n, dtype=torch.complex64, device=dev).requires_grad_(True)
    y = pkg.fourier_upsample(x, ds)
    ref = orc.fourier_upsample(x.detach().cpu().to(torch.complex128)[None, None], ds)[0, 0]
    assert rel_err(y, ref) < 5e-6
    if n % 2 == 0 and ds == 2:
        assert rel_err(y[::2, ::2], x) < 5e-6          # exact even-sample identity
    g = torch.randn(m, m, dtype=torch.complex64, device=dev)
    (torch.real(torch.conj(g) * y)).sum().backward()
    xr = x.detach().cpu().to(torch.complex128).requires_grad_(True)
    (torch.real(torch.conj(g.cpu().to(torch.complex128)) *
                orc.fourier_upsample(xr[None, None], ds)[0, 0])).sum().backward()
    assert rel_err(x.grad, xr.grad) < 5e-6


# ---------------------------------------------------------------- probe normalisation
def test_probe_normalize_fwd_bwd(pkg, dev):
    rng = np.random.default_rng(0)
    M = 64
    yy, xx = np.mgrid[0:M, 0:M]
    amp = np.exp(-((yy - 40.3) ** 2 + (xx - 22.7) ** 2) / 80.0)
    p0 = (amp * np.exp(1j * rng.uniform(-1, 1, (M, M)))).astype(np.complex64)
    p = torch.from_numpy(p0).to(dev).requires_grad_(True)
    pn, stats = pkg.normalize_probe(p, return_stats=True)
    pr = torch.from_numpy(p0).to(torch.complex128).requires_grad_(True)
    ref = orc.probe_center_normalize(pr)
    cp = orc.probe_com_shift(pr)
    st = stats.cpu().numpy()
    assert int(st[0]) == int(cp[0]) % M and int(st[1]) == int(cp[1]) % M and st[4] == 1.0
    assert rel_err(pn, ref) < 1e-6
    g = torch.randn(M, M, dtype=torch.complex64, device=dev)
    (torch.real(torch.conj(g) * pn)).sum().backward()
    (torch.real(torch.conj(g.cpu().to(torch.complex128)) * ref)).sum().backward()
    assert rel_err(p.grad, pr.grad) < 1e-5


# ---------------------------------------------------------------- far field
@pytest.mark.parametrize("M", [16, 64, 128, 256, 512])
def test_farfield_intensity(pkg, dev, M):
    rng = np.random.default_rng(M)
    H = 2 * M + 24
    obj = (rng.standard_normal((H, H)) + 1j * rng.standard_normal((H, H))).astype(np.complex64)
    probe = (rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))).astype(np.complex64)
    tlY = rng.integers(0, H - M + 1, 7)
    tlX = rng.integers(0, H - M + 1, 7)
    tl = pkg.positions_tensor(tlY, tlX, dev)
    bidx = torch.tensor([3, 0, 6, 6, 1], dtype=torch.int32, device=dev)
    I = pkg.farfield_intensity(torch.from_numpy(obj).to(dev), torch.from_numpy(probe).to(dev), bidx, tl, 7)
    for b, i in enumerate([3, 0, 6, 6, 1]):
        ref = orc.farfield_intensity(torch.from_numpy(obj).to(torch.complex128),
                                     torch.from_numpy(probe).to(torch.complex128), int(tlY[i]), int(tlX[i]))
        err = (I[b].cpu().double() - ref).abs().max() / ref.abs().max()
        assert err < 1e-5, (b, float(err))


@pytest.mark.parametrize("loss_type,M", [(lt, m) for m in (16, 64, 128, 256) for lt in
                                         ("smooth_L1_loss", "FD_loss", "Poisson_likelihood_loss")]
                         + [("FD_loss", 512), ("smooth_L1_loss", 512)])
def test_farfield_loss_and_grads(pkg, dev, loss_type, M):
    rng = np.random.default_rng(M + len(loss_type))
    H = 3 * M
    F = 9
    obj = (0.5 + 0.5 * rng.random((H, H))) * np.exp(1j * rng.uniform(-1, 1, (H, H)))
    probe = (rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))) * 0.3
    tlY = rng.integers(0, H - M + 1, F)
    tlX = rng.integers(0, H - M + 1, F)
    stack = rng.poisson(5.0, (M, M, F)).astype(np.float32)
    batch = [4, 1, 7, 2]
    o = torch.from_numpy(obj.astype(np.complex64)).to(dev).requires_grad_(True)
    p = torch.from_numpy(probe.astype(np.complex64)).to(dev).requires_grad_(True)
    st = pkg.prepare_stack(torch.from_numpy(stack).to(dev))
    tl = pkg.positions_tensor(tlY, tlX, dev)
    L, per = pkg.farfield_loss(o, p, st, torch.tensor(batch, dtype=torch.int32, device=dev), tl, loss_type,
                               return_per_pattern=True)
    L.backward()
    o64 = torch.from_numpy(obj.astype(np.complex64)).to(torch.complex128).requires_grad_(True)
    p64 = torch.from_numpy(probe.astype(np.complex64)).to(torch.complex128).requires_grad_(True)
    Lr = orc.batch_loss(o64, p64, torch.from_numpy(stack).double(), batch, tlY, tlX, loss_type)
    Lr.backward()
    assert abs(L.item() - Lr.item()) < 2e-5 * abs(Lr.item())
    for b, i in enumerate(batch):
        li = orc.batch_loss(o64.detach(), p64.detach(), torch.from_numpy(stack).double(), [i], tlY, tlX, loss_type)
        assert abs(per[b].item() - li.item()) < 2e-5 * abs(li.item())
    assert rel_err(o.grad, o64.grad) < 1e-4
    assert rel_err(p.grad, p64.grad) < 1e-4


# Probe shapes other than powers of two (the reference takes the shape of the
# measured patterns): odd sizes (one-column blocks, fftshift by M // 2), prime
# factors 3, 5, 7, 19, non-square windows, a power-of-two M with a
# non-power-of-two N (register-FFT column passes over N = 200 columns) and
# the reverse; 1000 x 1000 (2^3 5^3) has column passes of 72 KB of LDS.
ANY_SHAPES = [(45, 63), (100, 128), (64, 200), (57, 114), (96, 40), (384, 384), (1000, 1000)]


@pytest.mark.parametrize("M,N", ANY_SHAPES)
def test_farfield_intensity_any_shape(pkg, dev, M, N):
    rng = np.random.default_rng(M * 1000 + N)
    H, W = 2 * M + 11, 2 * N + 5
    obj = (rng.standard_normal((H, W)) + 1j * rng.standard_normal((H, W))).astype(np.complex64)
    probe = (rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))).astype(np.complex64)
    tlY = rng.integers(0, H - M + 1, 5)
    tlX = rng.integers(0, W - N + 1, 5)
    tl = pkg.positions_tensor(tlY, tlX, dev)
    order = [2, 0, 4, 4]
    bidx = torch.tensor(order, dtype=torch.int32, device=dev)
    I = pkg.farfield_intensity(torch.from_numpy(obj).to(dev), torch.from_numpy(probe).to(dev), bidx, tl, 5)
    assert I.shape == (4, M, N)
    for b, i in enumerate(order):
        ref = orc.farfield_intensity(torch.from_numpy(obj).to(torch.complex128),
                                     torch.from_numpy(probe).to(torch.complex128), int(tlY[i]), int(tlX[i]))
        err = (I[b].cpu().double() - ref).abs().max() / ref.abs().max()
        assert err < 1e-5, (b, float(err))


@pytest.mark.parametrize("loss_type", ["smooth_L1_loss", "FD_loss", "Poisson_likelihood_loss"])
@pytest.mark.parametrize("M,N", ANY_SHAPES)
def test_farfield_loss_and_grads_any_shape(pkg, dev, loss_type, M, N):
    rng = np.random.default_rng(M * 7 + N + len(loss_type))
    H, W = 2 * M + 9, 2 * N + 3
    F = 6
    obj = (0.5 + 0.5 * rng.random((H, W))) * np.exp(1j * rng.uniform(-1, 1, (H, W)))
    probe = (rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))) * 0.3
    tlY = rng.integers(0, H - M + 1, F)
    tlX = rng.integers(0, W - N + 1, F)
    stack = rng.poisson(5.0, (M, N, F)).astype(np.float32)
    batch = [3, 0, 5]
    o = torch.from_numpy(obj.astype(np.complex64)).to(dev).requires_grad_(True)
    p = torch.from_numpy(probe.astype(np.complex64)).to(dev).requires_grad_(True)
    st = pkg.prepare_stack(torch.from_numpy(stack).to(dev))
    tl = pkg.positions_tensor(tlY, tlX, dev)
    L, per = pkg.farfield_loss(o, p, st, torch.tensor(batch, dtype=torch.int32, device=dev), tl, loss_type,
                               return_per_pattern=True)
    L.backward()
    o64 = torch.from_numpy(obj.astype(np.complex64)).to(torch.complex128).requires_grad_(True)
    p64 = torch.from_numpy(probe.astype(np.complex64)).to(torch.complex128).requires_grad_(True)
    Lr = orc.batch_loss(o64, p64, torch.from_numpy(stack).double(), batch, tlY, tlX, loss_type)
    Lr.backward()
    assert abs(L.item() - Lr.item()) < 2e-5 * abs(Lr.item())
    for b, i in enumerate(batch):
        li = orc.batch_loss(o64.detach(), p64.detach(), torch.from_numpy(stack).double(), [i], tlY, tlX, loss_type)
        assert abs(per[b].item() - li.item()) < 2e-5 * abs(li.item())
    assert rel_err(o.grad, o64.grad) < 1e-4
    assert rel_err(p.grad, p64.grad) < 1e-4


@pytest.mark.parametrize("M,N", [(45, 63), (100, 100), (57, 114)])
def test_probe_normalize_any_shape(pkg, dev, M, N):
    rng = np.random.default_rng(M + N)
    yy, xx = np.mgrid[0:M, 0:N]
    # the centre of mass far from the middle: rolls of both signs
    amp = np.exp(-((yy - 0.8 * M) ** 2 + (xx - 0.2 * N) ** 2) / (0.05 * M * N))
    p0 = (amp * np.exp(1j * rng.uniform(-1, 1, (M, N)))).astype(np.complex64)
    p = torch.from_numpy(p0).to(dev).requires_grad_(True)
    pn, stats = pkg.normalize_probe(p, return_stats=True)
    pr = torch.from_numpy(p0).to(torch.complex128).requires_grad_(True)
    ref = orc.probe_center_normalize(pr)
    cp = orc.probe_com_shift(pr)
    st = stats.cpu().numpy()
    assert int(st[0]) == int(cp[0]) % M and int(st[1]) == int(cp[1]) % N and st[4] == 1.0
    assert int(cp[0]) != 0 and int(cp[1]) != 0
    assert rel_err(pn, ref) < 1e-6
    g = torch.randn(M, N, dtype=torch.complex64, device=dev)
    (torch.real(torch.conj(g) * pn)).sum().backward()
    (torch.real(torch.conj(g.cpu().to(torch.complex128)) * ref)).sum().backward()
    assert rel_err(p.grad, pr.grad) < 1e-5


@pytest.mark.parametrize("loss_type", ["smooth_L1_loss", "FD_loss"])
def test_fused_step_non_power_of_two_probe(pkg, dev, loss_type):
    """The whole fused step with a 40 x 40 probe (2^3 * 5) on the small canvas."""
    inst = pkg.synthetic.make_named("small", seed=2, M=40)
    model = _model(pkg, dev, inst.canvas, 40, log2T=12, residual=True,
                   probe_init=torch.from_numpy(inst.probe.astype(np.complex64)))
    params = oracle_params(model)
    eng = pkg.FusedReconstruction(model, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX,
                                  loss_type=loss_type, batch_size=4)
    batch = [1, 8, 20, 33]
    loss = eng.step(torch.tensor(batch, dtype=torch.int32, device=dev), update=False).sum().item()
    ref_loss, g = step_reference(model, params, inst, batch, loss_type)
    assert abs(loss - ref_loss) < 1e-4 * abs(ref_loss), (loss, ref_loss)
    assert rel_err(eng.flat.g(model.xy_hash_encoder.params), g["obj_table"]) < 2e-3
    for a, b in zip(model.object_complex_mlp.tensors(), g["obj_mlp"]):
        assert rel_err(eng.flat.g(a), b) < 2e-3, ("obj", a.shape, rel_err(eng.flat.g(a), b))
    _, g32 = step_reference(model, oracle_params(model, torch.float32), inst, batch, loss_type,
                            dtype=torch.float32)
    bar = max(2e-3, 4 * rel_err(g32["prb_table"], g["prb_table"]))
    assert rel_err(eng.flat.g(model.probe_hash_encoder.params), g["prb_table"]) < bar


# ---------------------------------------------------------------- fused step
@pytest.mark.parametrize("loss_type", ["smooth_L1_loss", "FD_loss", "Poisson_likelihood_loss"])
def test_fused_step_matches_oracle(pkg, dev, loss_type):
    inst = pkg.synthetic.make_named("small", seed=1)
    M = inst.probe.shape[0]
    model = _model(pkg, dev, inst.canvas, M, log2T=12, residual=True,
                   probe_init=torch.from_numpy(inst.probe.astype(np.complex64)))
    params = oracle_params(model)
    eng = pkg.FusedReconstruction(model, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX,
                                  loss_type=loss_type, batch_size=5)
    batch = [0, 7, 13, 22, 35]
    loss = eng.step(torch.tensor(batch, dtype=torch.int32, device=dev), update=False).sum().item()
    ref_loss, g = step_reference(model, params, inst, batch, loss_type)
    assert abs(loss - ref_loss) < 1e-4 * abs(ref_loss), (loss, ref_loss)
    assert rel_err(eng.flat.g(model.xy_hash_encoder.params), g["obj_table"]) < 2e-3
    # The probe branch passes through the adjoint of the energy normalisation,
    # g_q = s g_n - (s/E) Re<q, g_n> q, which removes the component of g_n
    # along q.  The Poisson loss gradient is dominated by that (scale)
    # component, so an FP32 computation keeps fewer correct digits there.
    # The bar of each probe tensor is the larger of 2e-3 (6e-3 Poisson) and
    # four times the error of the oracle's own autograd run in FP32: these
    # gradients are cancelling sums (e.g. the residual scale's Re sum dprobe
    # conj(mlp) over the probe), so two correct FP32 evaluation orders differ
    # by several times their rounding.
    tol_p = 6e-3 if loss_type == "Poisson_likelihood_loss" else 2e-3
    _, g32 = step_reference(model, oracle_params(model, torch.float32), inst, batch, loss_type,
                            dtype=torch.float32)

    def bar(ours32, ref64):
        return max(tol_p, 4 * rel_err(ours32, ref64))

    assert rel_err(eng.flat.g(model.probe_hash_encoder.params), g["prb_table"]) < bar(g32["prb_table"],
                                                                                      g["prb_table"])
    for a, b in zip(model.object_complex_mlp.tensors(), g["obj_mlp"]):
        assert rel_err(eng.flat.g(a), b) < 2e-3, ("obj", a.shape, rel_err(eng.flat.g(a), b))
    for a, b, c in zip(model.probe_complex_mlp.tensors(), g["prb_mlp"], g32["prb_mlp"]):
        assert rel_err(eng.flat.g(a), b) < bar(c, b), ("prb", a.shape, rel_err(eng.flat.g(a), b), rel_err(c, b))
    assert rel_err(eng.flat.g(model.k_probe), g["k_probe"]) < bar(g32["k_probe"], g["k_probe"])


@pytest.mark.parametrize("shape", [(2, 2), (6, 10), (64, 96), (1024, 1024), (3, 8, 12)])
def test_haar_sparsity_loss_and_grad(pkg, dev, shape):
    """HIP Haar sparsity (loss and adjoint) against the oracle in FP64: the
    loss to 1e-6 relative, the gradient (sums of band signs) to 1e-6 relative
    in norm (a band within f32 rounding of 0 may take the other sign)."""
    rng = np.random.default_rng(sum(shape))
    z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    for reduction in ("sum", "mean"):
        x = torch.from_numpy(z).to(dev).requires_grad_(True)
        L = pkg.haar_wavelet_sparsity_loss(x, reduction)
        (2.5 * L).backward()
        x64 = torch.from_numpy(z).to(torch.complex128).requires_grad_(True)
        Lr = orc.haar_wavelet_sparsity_loss(x64, reduction)
        (2.5 * Lr).backward()
        assert abs(L.item() - Lr.item()) < 1e-6 * abs(Lr.item()), (reduction, L.item(), Lr.item())
        assert rel_err(x.grad, x64.grad) < 1e-6, reduction
    # a real image: real gradient, imaginary part ignored
    r = torch.from_numpy(z.real.copy()).to(dev).requires_grad_(True)
    pkg.haar_wavelet_sparsity_loss(r).backward()
    r64 = torch.from_numpy(z.real.astype(np.float64)).requires_grad_(True)
    orc.haar_wavelet_sparsity_loss(r64).backward()
    assert r.grad.dtype == torch.float32 and rel_err(r.grad, r64.grad) < 1e-6


def test_haar_sparsity_rejects_odd_sizes(pkg, dev):
    with pytest.raises(RuntimeError, match="odd"):
        pkg.haar_wavelet_sparsity_loss(torch.zeros(6, 7, dtype=torch.complex64, device=dev))


def test_fused_step_with_sparsity_matches_oracle(pkg, dev):
    """sparsity_weight > 0 (Recovery.py:309-313): the step's objective is the
    batch loss plus the weighted Haar term of the object.  The object carries
    a textured object_initial (use_residual) so that its Haar bands are far
    from 0 and the band signs agree between FP32 and FP64; loss and every
    gradient against the oracle at the fused-step bars, and the canvas
    gradient equals the far-field one plus the weighted Haar adjoint."""
    inst = pkg.synthetic.make_named("small", seed=2)
    M = inst.probe.shape[0]
    rng = np.random.default_rng(11)
    tex = (rng.standard_normal((inst.canvas, inst.canvas)) + 1j * rng.standard_normal((inst.canvas, inst.canvas)))
    torch.manual_seed(0)
    model = pkg.ComplexINRModel2D(inst.canvas, inst.canvas, 2, probe_width=M, probe_height=M, log2_hashmap_size=12,
                                  hidden_omega_0=1.0, trainable_omega0=True, use_residual=True,
                                  object_initial=torch.from_numpy(tex.astype(np.complex64)),
                                  probe_initial=torch.from_numpy(inst.probe.astype(np.complex64))).to(dev)
    params = oracle_params(model)
    stack = torch.from_numpy(inst.stack).to(dev)
    bidx = torch.tensor([3, 8, 14, 21, 30], dtype=torch.int32, device=dev)
    plain = pkg.FusedReconstruction(model, stack, inst.tlY, inst.tlX, batch_size=5)
    data0 = plain.step(bidx, update=False).sum().item()
    dobj0 = plain.dobj.clone()
    w = 0.5
    eng = pkg.FusedReconstruction(model, stack, inst.tlY, inst.tlX, batch_size=5, sparsity_weight=w)
    data = eng.step(bidx, update=False).sum().item()
    assert data == data0
    total = data + w * eng.sparsity.item()
    assert abs(total - data) > 0.1 * abs(total), "the Haar term must be visible at this weight"
    # canvas gradient = far-field gradient + w * Haar adjoint (same FP32 object)
    o = eng.obj.detach().clone().requires_grad_(True)
    pkg.haar_wavelet_sparsity_loss(o).backward()
    assert rel_err(eng.dobj - dobj0, w * o.grad) < 1e-6
    ref_total, g = step_reference(model, params, inst, bidx.tolist(), "smooth_L1_loss", sparsity_weight=w)
    assert abs(total - ref_total) < 1e-4 * abs(ref_total), (total, ref_total)
    assert rel_err(eng.flat.g(model.xy_hash_encoder.params), g["obj_table"]) < 2e-3
    for a, b in zip(model.object_complex_mlp.tensors(), g["obj_mlp"]):
        assert rel_err(eng.flat.g(a), b) < 2e-3, ("obj", a.shape, rel_err(eng.flat.g(a), b))
    assert rel_err(eng.flat.g(model.k_object), g["k_object"]) < 2e-3


def test_fused_step_fp16_tables(pkg, dev):
    """Fused step with FP16 table storage against the oracle at the
    half-rounded tables (FP32 bars of the fused-step test)."""
    inst = pkg.synthetic.make_named("small", seed=1)
    M = inst.probe.shape[0]
    model = _model(pkg, dev, inst.canvas, M, log2T=12, residual=True, table_precision="fp16",
                   probe_init=torch.from_numpy(inst.probe.astype(np.complex64)))
    with torch.no_grad():   # table values where half rounding is visible
        model.xy_hash_encoder.params.uniform_(-0.3, 0.3)
        model.probe_hash_encoder.params.uniform_(-0.3, 0.3)
    eng = pkg.FusedReconstruction(model, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX,
                                  loss_type="smooth_L1_loss", batch_size=5)
    batch = [0, 7, 13, 22, 35]
    loss = eng.step(torch.tensor(batch, dtype=torch.int32, device=dev), update=False).sum().item()
    params = _round_tables_to_fp16(oracle_params(model))
    ref_loss, g = step_reference(model, params, inst, batch, "smooth_L1_loss")
    assert abs(loss - ref_loss) < 1e-4 * abs(ref_loss), (loss, ref_loss)
    assert rel_err(eng.flat.g(model.xy_hash_encoder.params), g["obj_table"]) < 2e-3
    assert rel_err(eng.flat.g(model.probe_hash_encoder.params), g["prb_table"]) < 2e-3
    for a, b in zip(model.object_complex_mlp.tensors(), g["obj_mlp"]):
        assert rel_err(eng.flat.g(a), b) < 2e-3, ("obj", a.shape, rel_err(eng.flat.g(a), b))


def test_adjoint_central_differences(pkg, dev):
    """The hand-written adjoint at 32 random parameters against central
    differences (h = 1e-3, FP64) of the oracle's step loss (smooth-L1, the
    plant-like setting).  The oracle's own autograd gradient is held to the
    same differences at 1e-5, which pins both."""
    inst = pkg.synthetic.make_named("small", seed=3)
    M = inst.probe.shape[0]
    model = _model(pkg, dev, inst.canvas, M, log2T=12, residual=True, seed=7,
                   probe_init=torch.from_numpy(inst.probe.astype(np.complex64)))
    eng = pkg.FusedReconstruction(model, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX,
                                  loss_type="smooth_L1_loss", batch_size=5)
    batch = [2, 5, 11, 17, 30]
    eng.step(torch.tensor(batch, dtype=torch.int32, device=dev), update=False)
    params = oracle_params(model)
    _, g_auto = step_reference(model, params, inst, batch, "smooth_L1_loss")
    table = oracle_table(model)
    from tests.helpers import oracle_geom
    geom = oracle_geom(model, inst.tlY, inst.tlX)
    stack = torch.from_numpy(np.asarray(inst.stack)).to(torch.float64)

    def loss_now():
        with torch.no_grad():
            return float(orc.step_loss(params, table, geom, stack, batch, "smooth_L1_loss")[0])

    def real(t):
        t = t.detach()
        return torch.view_as_real(t).reshape(-1) if t.is_complex() else t.reshape(-1)

    # (oracle leaf, its autograd gradient, the product parameter) per tensor
    groups = [[(params["obj_table"], g_auto["obj_table"], model.xy_hash_encoder.params)],
              [(params["prb_table"], g_auto["prb_table"], model.probe_hash_encoder.params)]]
    for key, mlp in (("obj_mlp", model.object_complex_mlp), ("prb_mlp", model.probe_complex_mlp)):
        leaves = []
        for (w, b, om) in params[key]["layers"]:
            leaves += [w, b, om]
        leaves += [params[key]["out_weight"], params[key]["out_bias"]]
        groups.append(list(zip(leaves, g_auto[key], mlp.tensors())))
    groups.append([(params["k_probe"], g_auto["k_probe"], model.k_probe)])
    rng = np.random.default_rng(11)
    picks = []
    for grp, count in zip(groups, [12, 6, 6, 6, 2]):
        for _ in range(count):
            leaf, ga, prod = grp[int(rng.integers(len(grp)))]
            picks.append((leaf, ga, prod, int(rng.integers(real(leaf).numel()))))
    assert len(picks) == 32
    h = 1e-3
    fd, ours, auto = [], [], []
    for leaf, ga, prod, i in picks:
        flat = real(leaf.data)
        x0 = flat[i].item()
        flat[i] = x0 + h
        lp = loss_now()
        flat[i] = x0 - h
        lm = loss_now()
        flat[i] = x0
        fd.append((lp - lm) / (2 * h))
        auto.append(real(ga)[i].item())
        ours.append(real(eng.flat.g(prod)).cpu().double()[i].item())
    fd, ours, auto = map(np.asarray, (fd, ours, auto))
    scale = np.linalg.norm(fd)
    assert scale > 0
    assert np.linalg.norm(auto - fd) < 1e-5 * scale, np.abs(auto - fd).max()
    assert np.linalg.norm(ours - fd) < 2e-3 * scale, (np.linalg.norm(ours - fd) / scale)


def test_fused_step_equals_autograd_dropin(pkg, dev):
    """The fused step and the drop-in autograd path (model.forward + loss) agree."""
    inst = pkg.synthetic.make_named("small", seed=2)
    M = inst.probe.shape[0]
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    m1 = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=5)
    m2 = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=5)
    batch = torch.tensor([1, 4, 9], dtype=torch.int32, device=dev)
    eng = pkg.FusedReconstruction(m1, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX, batch_size=3)
    l1 = eng.step(batch, update=False).sum()
    obj, probe = m2()
    pn = pkg.normalize_probe(probe)
    st = pkg.prepare_stack(torch.from_numpy(inst.stack).to(dev))
    tl = pkg.positions_tensor(inst.tlY, inst.tlX, dev)
    l2 = pkg.farfield_loss(obj, pn, st, batch, tl)
    l2.backward()
    assert abs(l1.item() - l2.item()) < 1e-5 * abs(l2.item())
    assert rel_err(eng.flat.g(m1.xy_hash_encoder.params), m2.xy_hash_encoder.params.grad) < 1e-5
    assert rel_err(eng.flat.g(m1.k_probe), m2.k_probe.grad) < 1e-5


def test_adam_matches_torch(pkg, dev):
    torch.manual_seed(0)
    p = torch.randn(1000, device=dev)
    ref = p.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=1e-3)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    hyper = torch.tensor([1e-3, 0.0], device=dev)
    s = pkg._lib.stream_ptr(dev)
    for _ in range(5):
        g = torch.randn(1000, device=dev)
        ref.grad = g.clone()
        opt.step()
        pkg._lib.call("pty_adam_tick", hyper.data_ptr(), s)
        pkg._lib.call("pty_adam", p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), 1000, hyper.data_ptr(),
                      0.9, 0.999, 1e-8, 1, s)
        assert not g.any()      # consumed gradient cleared
    assert rel_err(p, ref.detach()) < 1e-6


# ---------------------------------------------------------------- full-size properties
def test_full_size_properties(pkg, dev):
    """cfg1 shapes (INR 512^2 -> canvas 1024, probe 256^2, T=2^18): size-independent checks."""
    torch.manual_seed(0)
    model = _model(pkg, dev, 1024, 256, log2T=18, residual=True)
    obj, probe = model()
    assert obj.shape == (1024, 1024) and probe.shape == (256, 256)
    small = model._field(model.xy_hash_encoder, model.object_complex_mlp, "object").reshape(512, 512)
    assert rel_err(obj.detach()[::2, ::2], small.detach()) < 2e-6      # upsample identity at full size
    assert torch.isfinite(torch.view_as_real(obj)).all() and torch.isfinite(torch.view_as_real(probe)).all()
    # Parseval of the far field at 256^2
    pn = pkg.normalize_probe(probe.detach())
    tl = pkg.positions_tensor(np.array([0, 300, 768]), np.array([10, 500, 768]), dev)
    I = pkg.farfield_intensity(obj.detach(), pn, torch.tensor([0, 1, 2], dtype=torch.int32, device=dev), tl, 3)
    for b, (y, x) in enumerate([(0, 10), (300, 500), (768, 768)]):
        ew = (pn * obj.detach()[y:y + 256, x:x + 256]).abs().pow(2).double().sum()
        assert abs(I[b].double().sum() / (256 * 256) - ew) < 1e-5 * ew


@pytest.mark.parametrize("canvas,log2T", [(128, 10), (128, 12), (128, 14), (128, 18), (512, 14), (1024, 14)])
def test_table_gradient_paths_agree(pkg, dev, canvas, log2T):
    """Regular-grid owner-computes (cached plan / plan built per call) == explicit-coords scatter == oracle.

    log2T 10 and 12 have hashed levels with fewer than 2^8 columns per hash
    bucket (the per-entry path of k_table_hashed), 14 and 18 the shared-bucket
    path.  k_table_dense splits each entry of the leading dense levels with at
    least eight rows of points per cell over eight lanes: levels 0-1 of the
    256^2 grid (canvas 512), levels 0-3 of the 512^2 grid (canvas 1024).  The
    grid path runs three times and must give the same bits."""
    torch.manual_seed(3)
    model = _model(pkg, dev, canvas, 16, log2T=log2T, residual=False)
    with torch.no_grad():
        model.xy_hash_encoder.params.uniform_(-1, 1)
    spec = model.xy_hash_encoder.spec
    ts = [t.detach() for t in model.object_complex_mlp.tensors()]
    table = model.xy_hash_encoder.params.detach()
    grid = model.grid_points("object")
    coords = pkg.CoordPoints(model.xy_coords.to(dev))
    g = torch.randn(grid.n, dtype=torch.complex64, device=dev)
    outs = []
    for pts, with_spec in ((grid, True), (grid, True), (grid, False), (coords, False)):
        pdesc = pts.descriptor(spec if with_spec else None)
        grads = [torch.zeros_like(t) for t in ts]
        dtab = torch.zeros_like(table)
        desc = pkg.mlp_descriptor(ts, 32)
        gdesc = pkg.mlp_descriptor(grads, 32)
        ws = pkg._lib.workspace(pkg._lib.query("pty_inr_bwd_workspace", spec.descriptor(), desc, pdesc), dev)
        pkg._lib.call("pty_inr_bwd", spec.descriptor(), desc, gdesc, table.data_ptr(), dtab.data_ptr(), pdesc,
                      g.data_ptr(), ws.data_ptr(), ws.numel(), pkg._lib.stream_ptr(dev))
        outs.append(dtab.clone())
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])   # deterministic
    assert rel_err(outs[0], outs[3]) < 1e-5           # gather == scatter
    params = oracle_params(model)
    side = canvas // 2
    ref = orc.inr_field(model.xy_coords.cpu(), params["obj_table"], oracle_table(model), params["obj_mlp"], side, side)
    (torch.real(torch.conj(g.cpu().to(torch.complex128)) * ref.reshape(-1))).sum().backward()
    assert rel_err(outs[0], params["obj_table"].grad) < 1e-4


def test_fused_training_matches_autograd_training(pkg, dev):
    """Several fused steps (Adam, gradient clearing) track the drop-in autograd path + torch Adam."""
    inst = pkg.synthetic.make_named("small", seed=4)
    M = inst.probe.shape[0]
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    m1 = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=9)
    m2 = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=9)
    eng = pkg.FusedReconstruction(m1, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX, batch_size=4)
    eng.set_lr(1e-3)
    opt = torch.optim.Adam(lr=1e-3, params=[p for p in m2.parameters() if p.requires_grad])
    st = pkg.prepare_stack(torch.from_numpy(inst.stack).to(dev))
    tl = pkg.positions_tensor(inst.tlY, inst.tlX, dev)
    batches = [[0, 3, 8, 11], [5, 6, 20, 30], [1, 2, 9, 33], [4, 7, 10, 12]]
    for b in batches:
        bt = torch.tensor(b, dtype=torch.int32, device=dev)
        l1 = eng.step(bt).sum().item()
        opt.zero_grad()
        obj, probe = m2()
        l2 = pkg.farfield_loss(obj, pkg.normalize_probe(probe), st, bt, tl)
        l2.backward()
        opt.step()
        assert abs(l1 - l2.item()) < 1e-3 * abs(l2.item()), (b, l1, l2.item())
    # Adam turns FP32-level differences of near-zero gradients into +-lr steps,
    # so raw parameters are not comparable entry by entry; the fields are.
    with torch.no_grad():
        o1, p1 = m1()
        o2, p2 = m2()
    assert rel_err(o1, o2) < 1e-3 and rel_err(p1, p2) < 1e-3


def test_graph_replay_matches_eager(pkg, dev):
    """A captured step replayed several times (eager work and syncs in between) equals eager steps."""
    inst = pkg.synthetic.make_named("small", seed=6)
    M = inst.probe.shape[0]
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    m1 = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=11)
    m2 = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=11)
    stack = torch.from_numpy(inst.stack).to(dev)
    e1 = pkg.FusedReconstruction(m1, stack, inst.tlY, inst.tlX, batch_size=6)
    e2 = pkg.FusedReconstruction(m2, stack, inst.tlY, inst.tlX, batch_size=6)
    e1.set_lr(1e-3)
    e2.set_lr(1e-3)
    e2.capture()
    gen = torch.Generator().manual_seed(0)
    for it in range(5):
        b = torch.randperm(inst.stack.shape[2], generator=gen)[:6].to(torch.int32).to(dev)
        l1 = e1.step(b).clone()
        l2 = e2.graph_step(b).clone()
        torch.cuda.synchronize()
        _ = (torch.randn(1000, device=dev) ** 2).sum().item()     # unrelated eager work between replays
        # the fused step has no atomics and only fixed-order reductions, so the
        # graph replay is bitwise equal to the eager step
        assert torch.equal(l1, l2), (it, l1, l2)
        o1, _ = e1.current_fields()
        o2, _ = e2.current_fields()
        assert torch.equal(o1, o2), it


def test_graph_replay_after_step_without_update(pkg, dev):
    """An eager step without the update (gradients left in the buffer) followed
    by graph replays gives the bits of the same sequence run eagerly."""
    inst = pkg.synthetic.make_named("small", seed=12)
    M = inst.probe.shape[0]
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    engines = []
    for _ in range(2):
        m = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=23)
        e = pkg.FusedReconstruction(m, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX, batch_size=6)
        e.set_lr(1e-3)
        engines.append(e)
    e1, e2 = engines
    e2.capture()
    b0 = torch.arange(6, dtype=torch.int32, device=dev)
    e1.step(b0, update=False)
    e2.step(b0, update=False)
    for it in range(3):
        b = (b0 + 3 * it) % inst.stack.shape[2]
        l1 = e1.step(b).clone()
        l2 = e2.graph_step(b).clone()
        assert torch.equal(l1, l2), it
    assert torch.equal(e1.flat.data, e2.flat.data)


def test_batch_next_walks_an_order(pkg, dev):
    """pty_batch_next: batch c of the order, the previous batch's losses into
    the history, wrap to the start when a batch would pass the end, and the
    Adam count advancing only when hyper is given."""
    L = pkg._lib
    s = L.stream_ptr(dev)
    n, B = 23, 5
    order = torch.randperm(n, generator=torch.Generator().manual_seed(1)).to(torch.int32).to(dev)
    cursor = torch.zeros(1, dtype=torch.int32, device=dev)
    idx = torch.full((B,), -1, dtype=torch.int32, device=dev)
    hyper = torch.tensor([1e-3, 0.0], device=dev)
    loss = torch.zeros(B, device=dev)
    hist = torch.full((n,), -1.0, device=dev)
    starts = [0, 5, 10, 15, 0, 5]          # batch 4 (20..24) would pass 23: back to the start
    expect_hist = torch.full((n,), -1.0)
    for c, st in enumerate(starts):
        L.call("pty_batch_next", order.data_ptr(), n, cursor.data_ptr(), B, idx.data_ptr(),
               hyper.data_ptr() if c % 2 == 0 else None, loss.data_ptr(), hist.data_ptr(), n, s)
        assert torch.equal(idx.cpu(), order[st:st + B].cpu()), c
        prev = starts[c - 1] if c else None
        if c and st != 0:
            # the previous batch (number st / B - 1) recorded its losses
            expect_hist[st - B:st] = torch.arange(B, dtype=torch.float32) + 10 * (c - 1)
        assert torch.equal(hist.cpu(), expect_hist), (c, prev)
        assert int(cursor.item()) == st // B + 1
        loss.copy_(torch.arange(B, dtype=torch.float32, device=dev) + 10 * c)
    assert float(hyper[1].item()) == 3.0 and float(hyper[0].item()) == pytest.approx(1e-3)
    # without a history nothing is recorded
    hist2 = hist.clone()
    L.call("pty_batch_next", order.data_ptr(), n, cursor.data_ptr(), B, idx.data_ptr(), None, loss.data_ptr(),
           None, 0, s)
    assert torch.equal(hist, hist2) and torch.equal(idx.cpu(), order[10:15].cpu())


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
def test_run_epoch_graph_matches_eager_with_partial_batch(pkg, dev, precision):
    """run_epoch (graph replays for full batches, the trailing partial batch
    eager) gives the bits of the all-eager epoch, over two epochs; for FP8 this
    includes the delayed-scaling state, which the capture's warm-up pass must
    leave as it found it."""
    inst = pkg.synthetic.make_named("small", seed=13)
    M = inst.probe.shape[0]
    frames = inst.stack.shape[2]
    batch = 7
    assert frames % batch != 0
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    engines = []
    for _ in range(2):
        m = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=29, mlp_precision=precision)
        engines.append(pkg.FusedReconstruction(m, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX,
                                               batch_size=batch))
    g1 = torch.Generator(device=dev)
    g2 = torch.Generator(device=dev)
    g1.manual_seed(5)
    g2.manual_seed(5)
    for _ in range(2):
        t1 = engines[0].run_epoch(g1)
        t2 = engines[1].run_epoch(g2, use_graph=False)
        assert torch.equal(t1, t2)
    assert engines[0].graph is not None and engines[1].graph is None
    assert torch.equal(engines[0].flat.data, engines[1].flat.data)
    for a, b in zip(engines[0].fp8_states, engines[1].fp8_states):
        assert torch.equal(a, b)
    # the epoch used every full batch of its order: one more replay has no
    # batch left in it and is refused on the host
    with pytest.raises(RuntimeError, match="no full batch left"):
        engines[0].graph_next()
    engines[0].begin_order(torch.arange(batch, dtype=torch.int32, device=dev))
    engines[0].graph_next()
    with pytest.raises(RuntimeError, match="no full batch left"):
        engines[0].graph_next()


def test_fused_step_is_deterministic(pkg, dev):
    """Same inputs, same bits: two engines from the same seed stay identical over several updates."""
    inst = pkg.synthetic.make_named("small", seed=7)
    M = inst.probe.shape[0]
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    engines = []
    for _ in range(2):
        m = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=13)
        e = pkg.FusedReconstruction(m, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX, batch_size=5,
                                    loss_type="FD_loss")
        e.set_lr(1e-3)
        engines.append(e)
    for b in ([0, 1, 2, 3, 4], [10, 11, 12, 13, 14], [30, 5, 17, 2, 21]):
        bt = torch.tensor(b, dtype=torch.int32, device=dev)
        losses = [e.step(bt).clone() for e in engines]
        assert torch.equal(losses[0], losses[1])
    assert torch.equal(engines[0].flat.data, engines[1].flat.data)


@pytest.mark.parametrize("shape", ["small", "cfg1"])
def test_table_adam_in_place_matches_flat_adam(pkg, dev, shape):
    """The table-gradient kernels' in-place Adam step gives the bits of
    storing the table gradient and running the flat Adam over it (same
    element update, same gradient), over steps with and without the update.
    The canvas-1024 case (INR grid 512^2) has the coarse dense levels whose
    entries are split over several lanes."""
    inst = pkg.synthetic.make_named(shape, seed=8)
    M = inst.probe.shape[0]
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    engines = []
    for fuse in (True, False):
        m = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=17)
        e = pkg.FusedReconstruction(m, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX, batch_size=5,
                                    fuse_table_adam=fuse)
        e.set_lr(1e-3)
        engines.append(e)
    assert engines[0].fuse_table_adam and not engines[1].fuse_table_adam
    for it, b in enumerate(([0, 1, 2, 3, 4], [10, 11, 12, 13, 14], [30, 5, 17, 2, 21], [7, 8, 9, 20, 22])):
        bt = torch.tensor(b, dtype=torch.int32, device=dev)
        losses = [e.step(bt, update=(it != 1)).clone() for e in engines]
        assert torch.equal(losses[0], losses[1]), it
    f, u = engines
    assert torch.equal(f.flat.data, u.flat.data)
    assert torch.equal(f.m, u.m) and torch.equal(f.v, u.v)
    assert torch.equal(f.hyper, u.hyper)


def test_one_stream_matches_two_streams(pkg, dev):
    """The probe chain on its own stream (default) or on the caller's stream
    (overlap = False) gives the same bits: the two chains share no buffers."""
    inst = pkg.synthetic.make_named("small", seed=10)
    M = inst.probe.shape[0]
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    engines = []
    for overlap in (True, False):
        m = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=21)
        e = pkg.FusedReconstruction(m, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX, batch_size=5,
                                    loss_type="Poisson_likelihood_loss")
        e.overlap = overlap
        e.set_lr(1e-3)
        engines.append(e)
    for b in ([0, 1, 2, 3, 4], [9, 8, 7, 6, 5], [30, 5, 17, 2, 21]):
        bt = torch.tensor(b, dtype=torch.int32, device=dev)
        losses = [e.step(bt).clone() for e in engines]
        assert torch.equal(losses[0], losses[1])
    assert torch.equal(engines[0].flat.data, engines[1].flat.data)


def test_table_adam_in_place_contract(pkg, dev):
    """pty_inr_bwd_finish_adam: after an updating fused step the tables' slots
    of the flat gradient stay zero (the MLP slots are consumed by Adam too);
    the entry point rejects explicit coordinates and misaligned buffers with
    an error and no launch."""
    inst = pkg.synthetic.make_named("small", seed=9)
    M = inst.probe.shape[0]
    pinit = torch.from_numpy(inst.probe.astype(np.complex64))
    m = _model(pkg, dev, inst.canvas, M, residual=True, probe_init=pinit, seed=19)
    e = pkg.FusedReconstruction(m, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX, batch_size=5)
    e.set_lr(1e-3)
    before = m.xy_hash_encoder.params.detach().clone()
    e.step(torch.arange(5, dtype=torch.int32, device=dev))
    assert not torch.equal(before, m.xy_hash_encoder.params.detach())      # the table moved
    assert not e.flat.grad.any()                                            # nothing left behind
    L = pkg._lib
    n_tab = m.xy_hash_encoder.params.numel()
    buf = torch.zeros(3 * n_tab + 8, device=dev)
    p, mm, v = (buf[i * n_tab:(i + 1) * n_tab] for i in range(3))
    # explicit coordinates have no plan: rejected
    pts = pkg.complex_inr.CoordPoints(torch.rand(64, 2, device=dev)).descriptor(e.obj_spec)
    with pytest.raises(RuntimeError, match="grid"):
        L.call("pty_inr_bwd_finish_adam", e.obj_grid, e.obj_mlp, e.obj_gmlp, p.data_ptr(), mm.data_ptr(),
               v.data_ptr(), e.hyper.data_ptr(), 0.9, 0.999, 1e-8, pts, e.ws_obj.data_ptr(), e.ws_obj.numel(),
               L.stream_ptr(dev))
    # a moment buffer off its 8-byte alignment: rejected
    with pytest.raises(RuntimeError, match="aligned"):
        L.call("pty_inr_bwd_finish_adam", e.obj_grid, e.obj_mlp, e.obj_gmlp, p.data_ptr(), mm.data_ptr() + 4,
               v.data_ptr(), e.hyper.data_ptr(), 0.9, 0.999, 1e-8, e.obj_pdesc, e.ws_obj.data_ptr(), e.ws_obj.numel(),
               L.stream_ptr(dev))
    torch.cuda.synchronize()


def _mlp_grads_vs_oracle(model, params):
    """(name, error) per MLP tensor.  Weights and biases: relative error (norm).
    omega: omega * dL/domega = Re<G_W, W> + Re<G_b, b> sums terms that cancel,
    so its error is measured against the Cauchy-Schwarz scale
    (|G_W| |W| + |G_b| |b|) / |omega| of that identity."""
    mlp = params["obj_mlp"]
    ours = model.object_complex_mlp.tensors()
    out, i = [], 0
    for li, (w, b, om) in enumerate(mlp["layers"]):
        out.append((f"W{li}", rel_err(ours[i].grad, w.grad)))
        out.append((f"b{li}", rel_err(ours[i + 1].grad, b.grad)))
        scale = (w.grad.abs().pow(2).sum().sqrt() * w.detach().abs().pow(2).sum().sqrt()
                 + b.grad.abs().pow(2).sum().sqrt() * b.detach().abs().pow(2).sum().sqrt()).item() / abs(om.item())
        out.append((f"omega{li}", abs(ours[i + 2].grad.item() - om.grad.item()) / scale))
        i += 3
    out.append(("Wout", rel_err(ours[i].grad, mlp["out_weight"].grad)))
    out.append(("bout", rel_err(ours[i + 1].grad, mlp["out_bias"].grad)))
    return out


@pytest.mark.parametrize("hidden_layers", [1, 2])
def test_inr_field_fp8(pkg, dev, hidden_layers):
    """FP8 e4m3 MFMA path with per-tensor scales: the delayed-scaling state
    holds 224 / amax of what the previous pass saw, and field and gradients
    stay within the error of 3-bit mantissas against the FP64 oracle.

    Tolerance: e4m3 rounds to 2^-4 relative (about 3.6 % RMS per operand);
    each of the three forward GEMMs perturbs the phases by a few percent of a
    radian (omega_0 = 10 on layer 0), so the field is held to 1.2e-1 and the
    parameter gradients (two more GEMMs) to 3e-1."""
    model = _model(pkg, dev, 64, 16, log2T=12, residual=False, hidden_layers=hidden_layers)
    model.mlp_precision = "fp8"
    with torch.no_grad():   # features of order 0.1 as in the BF16 test (omega_0 = 10)
        model.xy_hash_encoder.params.uniform_(-0.1, 0.1)
    params = oracle_params(model)
    table = oracle_table(model)
    ref = orc.inr_field(model.xy_coords.cpu(), params["obj_table"], table, params["obj_mlp"], 32, 32)
    g = torch.randn(32 * 32, dtype=torch.complex64, device=dev)
    state = model.object_complex_mlp.fp8_state(dev)
    # pass 1 calibrates the scales (amax recorded, rolled after the backward)
    f = model._field(model.xy_hash_encoder, model.object_complex_mlp, "object")
    (torch.real(torch.conj(g) * f)).sum().backward()
    feats = orc.hashgrid_encode(model.xy_coords.cpu().numpy(), params["obj_table"].detach(), table)
    s = state.cpu()
    assert abs(s[0].item() - 224.0 / feats.abs().max().item()) < 1e-5 * s[0].item()
    assert s.numel() == 520 and (s[8:] == 0).all()
    assert (s[1:1 + hidden_layers] != 1).all() and (s[3:4 + hidden_layers] != 1).all()
    for p in model.parameters():
        p.grad = None
    f = model._field(model.xy_hash_encoder, model.object_complex_mlp, "object")
    err_f = rel_err(f.reshape(32, 32), ref)
    print(f"fp8 H={hidden_layers}: field rel err {err_f:.4f}")
    assert err_f < 1.2e-1, err_f
    (torch.real(torch.conj(g) * f)).sum().backward()
    (torch.real(torch.conj(g.cpu().to(torch.complex128)) * ref.reshape(-1))).sum().backward()
    e_t = rel_err(model.xy_hash_encoder.params.grad, params["obj_table"].grad)
    errs = _mlp_grads_vs_oracle(model, params)
    print(f"fp8 H={hidden_layers}: table grad {e_t:.4f}, mlp grads {[round(e, 4) for _, e in errs]}")
    assert e_t < 3e-1
    for shape, e in errs:
        assert e < 3e-1, (shape, e)


@pytest.mark.parametrize("hidden_layers,omegas", [(1, (10.0, 1.0)), (2, (10.0, 1.0)), (2, (13.0, 1.7))])
def test_inr_field_bf16(pkg, dev, hidden_layers, omegas):
    """BF16 MFMA path (fp32 accumulate): field and gradients against the FP64 oracle.

    Tolerance: bf16 operands carry 8 mantissa bits (relative 2^-9 per rounding);
    with features of order 0.1 and omega_0 = 10 the phases move by ~1e-3 rad, so
    the field is held to 2e-2 and the gradients to 5e-2 relative (norm).  The
    kernels fold each layer's frequency into the bf16 weight operands, so one
    case runs non-unit frequencies on every layer (first 13, hidden 1.7),
    where a frequency applied twice or not at all would show.
    """
    model = _model(pkg, dev, 64, 16, log2T=12, residual=False, hidden_layers=hidden_layers,
                   first_omega=omegas[0], hidden_omega=omegas[1])
    model.mlp_precision = "bf16"
    with torch.no_grad():
        model.xy_hash_encoder.params.uniform_(-0.1, 0.1)
    f = model._field(model.xy_hash_encoder, model.object_complex_mlp, "object")
    params = oracle_params(model)
    ref = orc.inr_field(model.xy_coords.cpu(), params["obj_table"], oracle_table(model), params["obj_mlp"], 32, 32)
    err_f = rel_err(f.reshape(32, 32), ref)
    assert err_f < 2e-2, err_f
    g = torch.randn(f.shape, dtype=torch.complex64, device=dev)
    (torch.real(torch.conj(g) * f)).sum().backward()
    (torch.real(torch.conj(g.cpu().to(torch.complex128)) * ref.reshape(-1))).sum().backward()
    assert rel_err(model.xy_hash_encoder.params.grad, params["obj_table"].grad) < 5e-2
    mlp = params["obj_mlp"]
    theirs = []
    for (w, b, om) in mlp["layers"]:
        theirs += [w, b, om]
    theirs += [mlp["out_weight"], mlp["out_bias"]]
    for a, b in zip(model.object_complex_mlp.tensors(), theirs):
        assert rel_err(a.grad, b.grad) < 5e-2, (a.shape, rel_err(a.grad, b.grad))
