"""Learned warm start: host-side shapes on the CPU; on the GPU, a small net
trained on simulated instances must start a reconstruction closer to the
object than the plain initProbe start and stay ahead after a few epochs."""
import numpy as np
import pytest
import torch


def test_net_and_patch_shapes_cpu(pkg):
    ws = pkg.warmstart
    torch.manual_seed(0)
    net = ws.WarmStartNet(base=8)
    x = torch.randn(3, 3, 64, 64)
    assert net(x).shape == (3, 2, 32, 32)
    intens = torch.rand(5, 128, 128)
    inp = ws.net_inputs(intens, np.arange(5) * 10, np.arange(5) * 20, 128, 512)
    assert inp.shape == (5, 3, 64, 64)
    assert torch.allclose(inp[:, 0], torch.nn.functional.adaptive_avg_pool2d(torch.log1p(intens)[:, None], 64)[:, 0])
    obj = torch.polar(torch.ones(512, 512), torch.zeros(512, 512))
    t = ws.patch_targets(obj, np.arange(5) * 10, np.arange(5) * 20, 128)
    assert t.shape == (5, 2, 32, 32) and torch.allclose(t[:, 0], torch.ones(5, 32, 32))
    # targets are rotated so that the patch mean is real and positive
    obj2 = obj * torch.polar(torch.ones(()), torch.tensor(1.3))
    t2 = ws.patch_targets(obj2, np.arange(5) * 10, np.arange(5) * 20, 128)
    assert torch.allclose(t2, t, atol=1e-6)
    probe = torch.full((4, 4), 2.0 + 0j)
    assert ws.object_scale(probe) == pytest.approx(2.0)


@pytest.mark.gpu
def test_warm_start_gains_over_plain_start(pkg, dev):
    ws = pkg.warmstart
    shape = dict(canvas=256, M=64, n_side=10, step=16, jitter=2, photons=1e4)
    x, y = ws.build_dataset(range(100, 140), device=dev, **shape)
    torch.manual_seed(0)
    net = ws.WarmStartNet(base=16).to(dev)
    hist = ws.train(net, x, y, epochs=6, batch=128)
    assert hist[-1] < hist[0]
    net.eval()
    inst = ws.simulate(7, jitter_probe=False, device=dev, **shape)
    truth = inst.obj.cpu().numpy()
    ty, tx = inst.tlY.cpu().numpy(), inst.tlX.cpu().numpy()
    mask = pkg.metrics.evaluation_mask(ty, tx, 64, 64, truth.shape)
    stack = inst.intens.permute(1, 2, 0).contiguous()

    def err(o):
        return pkg.metrics.aligned_nrmse(o.detach().cpu().numpy(), truth, mask, max_shift=2)

    def run(init, epochs=4):
        torch.manual_seed(0)
        model = pkg.ComplexINRModel2D(256, 256, 2, probe_width=64, probe_height=64, log2_hashmap_size=14,
                                      hidden_omega_0=1.0, trainable_omega0=True, use_residual=True,
                                      object_initial=init, probe_initial=inst.probe.to(torch.complex64)).to(dev)
        eng = pkg.FusedReconstruction(model, stack, ty, tx, batch_size=10)
        e0 = err(eng.current_fields()[0])
        gen = torch.Generator(device=dev)
        gen.manual_seed(0)
        for _ in range(epochs):
            eng.run_epoch(gen)
        return e0, err(eng.current_fields()[0])

    init = ws.warm_start_object(net, inst.intens, ty, tx, 64, 256)
    p0, p1 = run(None)
    w0, w1 = run(init)
    print(f"plain start {p0:.3f} -> {p1:.3f}; warm start {w0:.3f} -> {w1:.3f}")
    # the gain: after the same number of epochs the warm-started run is
    # closer to the object, and the warm start is not worse at epoch 0
    assert w0 < p0 + 0.02
    assert w1 < p1 - 0.02
