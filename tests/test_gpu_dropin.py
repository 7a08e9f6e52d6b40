"""The module-level drop-in surface on the GPU against the FP64 oracle.

* ComplexGaborLayer2D.forward (models/complex_inr.py:38-40), the final complex
  Linear and ComplexMLP.net / .forward / _get_complex_output
  (complex_inr.py:73-82, 308-314), forward and backward, layer by layer;
* ComplexINRModel2D.forward on coordinates a caller has edited (the model
  then evaluates the INR on those coordinates, not on the default grid);
* the reference's own inline step (Recovery.py:262-316: COM roll, energy
  normalisation, per-position fftshift(fft2(.)), |.|^2, torch loss, batch
  sum, Haar term, backward) run with torch ops over the drop-in model, whose
  forward and backward are the HIP kernels: loss and every gradient against
  the oracle.

Bars: layers and MLP 1e-5 (FP32 FMA chains of length 32 against FP64);
fields 2e-6 / gradients 1e-4 as the fused-INR tests; the inline step with the
fused-step bars (loss 1e-4, gradients 2e-3, probe branch as there).
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import ptycho_oracle as orc
from tests.helpers import oracle_params, oracle_table, rel_err, step_reference

pytestmark = pytest.mark.gpu


def _mlp(pkg, dev, seed=0, hidden_layers=2):
    torch.manual_seed(seed)
    return pkg.ComplexMLP(32, 64, hidden_layers, 1, first_omega_0=10.0, hidden_omega_0=1.0,
                          trainable_omega0=True).to(dev)


def test_gabor_layers_and_linear_fwd_bwd(pkg, dev):
    mlp = _mlp(pkg, dev)
    layers = list(mlp.net)
    torch.manual_seed(1)
    h = (torch.rand(700, 32, device=dev) - 0.5) * 0.2
    # first (real) Gabor layer
    x = h.clone().requires_grad_(True)
    z0 = layers[0](x)
    assert z0.dtype == torch.complex64 and z0.shape == (700, 32)
    x64 = h.detach().cpu().double().requires_grad_(True)
    w0 = layers[0].linear.weight.detach().cpu().double().requires_grad_(True)
    b0 = layers[0].linear.bias.detach().cpu().double().requires_grad_(True)
    o0 = layers[0].omega_0.detach().cpu().double().requires_grad_(True)
    r0 = orc.gabor_layer(x64, w0, b0, o0)
    assert rel_err(z0, r0) < 1e-5
    g = torch.randn(700, 32, dtype=torch.complex64, device=dev)
    torch.real(torch.conj(g) * z0).sum().backward()
    torch.real(torch.conj(g.cpu().to(torch.complex128)) * r0).sum().backward()
    for a, b in ((x.grad, x64.grad), (layers[0].linear.weight.grad, w0.grad), (layers[0].linear.bias.grad, b0.grad),
                 (layers[0].omega_0.grad, o0.grad)):
        assert a.dtype == (torch.float32 if not b.is_complex() else torch.complex64)
        assert rel_err(a, b) < 1e-5, (a.shape, rel_err(a, b))
    # hidden (complex) layer and the final complex Linear on a complex input
    zin = z0.detach().clone().requires_grad_(True)
    z1 = layers[1](zin)
    o = layers[-1](z1)
    assert o.shape == (700, 1) and o.dtype == torch.complex64
    zin64 = z0.detach().cpu().to(torch.complex128).requires_grad_(True)
    w1 = layers[1].linear.weight.detach().cpu().to(torch.complex128).requires_grad_(True)
    b1 = layers[1].linear.bias.detach().cpu().to(torch.complex128).requires_grad_(True)
    o1 = layers[1].omega_0.detach().cpu().double().requires_grad_(True)
    wo = layers[-1].weight.detach().cpu().to(torch.complex128).requires_grad_(True)
    bo = layers[-1].bias.detach().cpu().to(torch.complex128).requires_grad_(True)
    r = F.linear(orc.gabor_layer(zin64, w1, b1, o1), wo, bo)
    assert rel_err(o, r) < 1e-5
    go = torch.randn(700, 1, dtype=torch.complex64, device=dev)
    torch.real(torch.conj(go) * o).sum().backward()
    torch.real(torch.conj(go.cpu().to(torch.complex128)) * r).sum().backward()
    for a, b in ((zin.grad, zin64.grad), (layers[1].linear.weight.grad, w1.grad), (layers[1].linear.bias.grad, b1.grad),
                 (layers[1].omega_0.grad, o1.grad), (layers[-1].weight.grad, wo.grad), (layers[-1].bias.grad, bo.grad)):
        assert rel_err(a, b) < 1e-5, (a.shape, rel_err(a, b))


@pytest.mark.parametrize("hidden_layers", [1, 2, 3])
def test_complex_mlp_forward_and_complex_output(pkg, dev, hidden_layers):
    """ComplexMLP.forward (.real of net) with autograd, and net(coords) /
    _get_complex_output; three hidden layers take the per-layer path in both
    modes (the fused kernel implements one or two)."""
    mlp = _mlp(pkg, dev, seed=3, hidden_layers=hidden_layers)
    torch.manual_seed(4)
    h = (torch.rand(500, 32, device=dev) - 0.5) * 0.2
    P = {"layers": [], "out_weight": None, "out_bias": None}
    ls = list(mlp.net)
    for i, l in enumerate(ls[:-1]):
        dt = torch.float64 if i == 0 else torch.complex128
        P["layers"].append(tuple(t.detach().cpu().to(d).requires_grad_(True) for t, d in
                                 ((l.linear.weight, dt), (l.linear.bias, dt), (l.omega_0, torch.float64))))
    P["out_weight"] = ls[-1].weight.detach().cpu().to(torch.complex128).requires_grad_(True)
    P["out_bias"] = ls[-1].bias.detach().cpu().to(torch.complex128).requires_grad_(True)
    ref = orc.complex_mlp(h.cpu().double(), P["layers"], P["out_weight"], P["out_bias"])
    real = mlp(h)
    assert real.dtype == torch.float32 and real.shape == (500, 1)
    assert rel_err(real, ref.real) < 1e-5
    y = torch.randn(500, 1, device=dev)
    (real * y).sum().backward()
    (ref.real * y.cpu().double()).sum().backward()
    theirs = [t for lay in P["layers"] for t in lay] + [P["out_weight"], P["out_bias"]]
    for a, b in zip(mlp.tensors(), theirs):
        assert rel_err(a.grad, b.grad) < 1e-4, (a.shape, rel_err(a.grad, b.grad))
    model = pkg.ComplexINRModel2D(32, 32, 2, probe_width=16, probe_height=16, log2_hashmap_size=10).to(dev)
    with torch.no_grad():
        full = model._get_complex_output(mlp, h)
        assert full.dtype == torch.complex64
        assert rel_err(full, ref.detach()) < 1e-5
        assert rel_err(mlp.net(h), ref.detach()) < 1e-5


@pytest.mark.parametrize("lo,hi", [(0.0, 1.0), (-0.25, 1.25)])
def test_forward_on_edited_coordinates(pkg, dev, lo, hi):
    """A caller who edits xy_coords gets the INR on those coordinates (the
    reference evaluates the encoder on the stored tensor, complex_inr.py:279),
    also outside [0, 1] (the encoder's general dense-index modulo)."""
    torch.manual_seed(5)
    model = pkg.ComplexINRModel2D(64, 64, 2, probe_width=16, probe_height=16, log2_hashmap_size=12,
                                  hidden_omega_0=1.0, trainable_omega0=True).to(dev)
    with torch.no_grad():
        model.xy_hash_encoder.params.uniform_(-1, 1)
        g = torch.Generator(device="cpu").manual_seed(6)
        new = (lo + (hi - lo) * torch.rand(model.xy_coords.shape, generator=g)).to(dev)
        model.xy_coords.copy_(new)
    assert not model.coords_are_grid("object") and model.coords_are_grid("probe")
    obj, _ = model()
    params = oracle_params(model)
    small = orc.inr_field(new.cpu(), params["obj_table"], oracle_table(model), params["obj_mlp"], 32, 32)
    big = orc.fourier_upsample(small[None, None], 2)[0, 0]
    assert rel_err(obj, big) < 2e-6
    gobj = torch.randn(64, 64, dtype=torch.complex64, device=dev)
    torch.real(torch.conj(gobj) * obj).sum().backward()
    torch.real(torch.conj(gobj.cpu().to(torch.complex128)) * big).sum().backward()
    assert rel_err(model.xy_hash_encoder.params.grad, params["obj_table"].grad) < 1e-4
    theirs = [t for lay in params["obj_mlp"]["layers"] for t in lay] + [params["obj_mlp"]["out_weight"],
                                                                       params["obj_mlp"]["out_bias"]]
    for a, b in zip(model.object_complex_mlp.tensors(), theirs):
        assert rel_err(a.grad, b.grad) < 1e-4, (a.shape, rel_err(a.grad, b.grad))


def _reference_inline_step(pkg, model, stack_mnf, tlY, tlX, batch, loss_type, sparsity_weight):
    """The loop body of Recovery.py:253-316, written with torch ops over the
    drop-in model (device tensors; stack in the reference layout [M, N, F])."""
    object_complex, probe_complex = model()
    with torch.no_grad():
        absP2 = probe_complex.abs() ** 2
        Mp, Np = absP2.shape
        tot = absP2.sum()
        size = torch.tensor([Mp, Np], device=probe_complex.device)
        com = torch.stack((absP2.sum(1).cumsum(0).mean(), absP2.sum(0).cumsum(0).mean()))
        cp = torch.trunc(size / 2 - size * com / tot + 1).long()
    if cp.any():
        probe_complex = torch.roll(probe_complex, (-cp[0].item(), -cp[1].item()), (0, 1))
    energy = (probe_complex.abs() ** 2).sum()
    probe_complex = probe_complex * (float(Mp * Np) / energy).sqrt()
    total = 0.0
    M, N = probe_complex.shape
    for i in batch:
        meas = stack_mnf[:, :, i]
        ew = probe_complex * object_complex[tlY[i]:tlY[i] + M, tlX[i]:tlX[i] + N]
        inten = torch.abs(torch.fft.fftshift(torch.fft.fft2(ew), dim=(-2, -1))) ** 2
        if loss_type == "smooth_L1_loss":
            loss = F.smooth_l1_loss(inten, meas)
        elif loss_type == "FD_loss":
            eps = 1e-8
            dW = torch.sqrt(meas + eps) - torch.sqrt(inten + eps)
            rx = torch.cat((dW[:, 1:] - dW[:, :-1], dW[:, :1] - dW[:, -1:]), dim=1)
            ry = torch.cat((dW[1:, :] - dW[:-1, :], dW[:1, :] - dW[-1:, :]), dim=0)
            loss = torch.sqrt(rx ** 2 + ry ** 2 + eps).mean()
        else:
            loss = (inten - meas * torch.log(inten + 1e-8)).mean()
        total = total + loss
    total = total + sparsity_weight * pkg.haar_wavelet_sparsity_loss(object_complex)
    total.backward()
    return float(total.detach())


@pytest.mark.parametrize("loss_type,sw", [("smooth_L1_loss", 0.0), ("FD_loss", 0.0),
                                          ("Poisson_likelihood_loss", 0.0), ("smooth_L1_loss", 1e-3)])
def test_reference_inline_step_over_dropin_model(pkg, dev, loss_type, sw):
    inst = pkg.synthetic.make_named("small", seed=1)
    M = inst.probe.shape[0]
    torch.manual_seed(0)
    model = pkg.ComplexINRModel2D(inst.canvas, inst.canvas, 2, probe_width=M, probe_height=M,
                                  log2_hashmap_size=12, hidden_omega_0=1.0, trainable_omega0=True, use_residual=True,
                                  probe_initial=torch.from_numpy(inst.probe.astype(np.complex64))).to(dev)
    params = oracle_params(model)
    batch = [0, 7, 13, 22, 35]
    stack = torch.from_numpy(inst.stack).to(dev)
    loss = _reference_inline_step(pkg, model, stack, inst.tlY, inst.tlX, batch, loss_type, sw)
    ref_loss, g = step_reference(model, params, inst, batch, loss_type, sparsity_weight=sw)
    assert abs(loss - ref_loss) < 1e-4 * abs(ref_loss), (loss, ref_loss)
    assert rel_err(model.xy_hash_encoder.params.grad, g["obj_table"]) < 2e-3
    for a, b in zip(model.object_complex_mlp.tensors(), g["obj_mlp"]):
        assert rel_err(a.grad, b) < 2e-3, ("obj", a.shape, rel_err(a.grad, b))
    _, g32 = step_reference(model, oracle_params(model, torch.float32), inst, batch, loss_type,
                            dtype=torch.float32, sparsity_weight=sw)
    tol_p = 6e-3 if loss_type == "Poisson_likelihood_loss" else 2e-3

    def bar(ours32, ref64):
        return max(tol_p, 4 * rel_err(ours32, ref64))

    assert rel_err(model.probe_hash_encoder.params.grad, g["prb_table"]) < bar(g32["prb_table"], g["prb_table"])
    for a, b, c in zip(model.probe_complex_mlp.tensors(), g["prb_mlp"], g32["prb_mlp"]):
        assert rel_err(a.grad, b) < bar(c, b), ("prb", a.shape, rel_err(a.grad, b))
    assert rel_err(model.k_probe.grad, g["k_probe"]) < bar(g32["k_probe"], g["k_probe"])


def test_reconstruct_writes_results_like_recovery(pkg, dev, tmp_path):
    """recovery.reconstruct with result_dir / vis_interval (Recovery.py:140-141,
    352-383): result.mat holds the last step's object and normalised probe,
    trained_models.pth the trained parameters."""
    from scipy.io import loadmat
    syn = pkg.synthetic
    inst = syn.make_named("tiny", seed=4)
    data = {"xlocation": inst.xlocation, "ylocation": inst.ylocation, "probe": inst.probe, "imRaw": inst.stack,
            "initProbe": inst.probe}
    cfg = pkg.recovery.normalize_config({
        "physics": {"wavelength": syn.WAVELENGTH, "camera_length": syn.CAMERA_LENGTH,
                    "camera_pixel_pitch": syn.PIXEL_PITCH},
        "model": {"init_type": "initProbe", "use_residual": True, "log2_hashmap_size": 10, "canvas_size": inst.canvas},
        "training": {"iters": 3, "batch_size": 4, "lr": 1e-3, "seed": 2},
        "output": {"vis_interval": 2, "result_dir": str(tmp_path / "result")},
    })
    model, eng, geom, losses = pkg.recovery.reconstruct(cfg, data, device=dev)
    assert losses.shape == (3,) and np.isfinite(losses).all()
    m = loadmat(str(tmp_path / "result" / "result.mat"))
    assert np.array_equal(m["object_complex"], eng.obj.cpu().numpy())
    assert np.array_equal(m["probe_complex"], eng.probe_n.cpu().numpy())
    saved = torch.load(str(tmp_path / "result" / "trained_models.pth"), weights_only=True)
    trained = [p for _, p in model.named_parameters() if p.requires_grad]
    assert len(saved) == len(trained)
    for a, b in zip(saved, trained):
        assert torch.equal(a, b.detach().cpu())
