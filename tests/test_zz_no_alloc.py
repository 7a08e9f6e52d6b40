"""pty_farfield_fwd_bwd_noalloc (static device scratch, no workspace argument)
against pty_farfield_fwd_bwd with a workspace, through the C ABI: two calls in
a row, the second with a smaller input, each equal bit for bit to the
workspace entry point (the same kernels over the same data); and the size
limit rejected on the host with a clear error."""
import numpy as np
import pytest
import torch

FAR_SCRATCH = 16 << 20


def _case(pkg, dev, M, F, B, H, loss_type, seed):
    rng = np.random.default_rng(seed)
    obj = ((0.5 + 0.5 * rng.random((H, H))) * np.exp(1j * rng.uniform(-1, 1, (H, H)))).astype(np.complex64)
    probe = ((rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))) * 0.3).astype(np.complex64)
    tlY = rng.integers(0, H - M + 1, F)
    tlX = rng.integers(0, H - M + 1, F)
    stack = rng.poisson(5.0, (M, M, F)).astype(np.float32)
    st = pkg.prepare_stack(torch.from_numpy(stack).to(dev))
    tl = pkg.positions_tensor(tlY, tlX, dev)
    bidx = torch.from_numpy(rng.choice(F, B, replace=False).astype(np.int32)).to(dev)
    return (torch.from_numpy(obj).to(dev), torch.from_numpy(probe).to(dev), st, tl, bidx,
            pkg.farfield.far_descriptor(M, M, H, H, B, F, loss_type))


def _with_workspace(pkg, dev, o, p, st, tl, bidx, d):
    L = pkg._lib
    B = bidx.numel()
    loss = torch.empty(B, device=dev)
    dobj = torch.empty_like(o)
    dprb = torch.empty_like(p)
    ws = L.workspace(L.query("pty_farfield_workspace", d), dev)
    L.call("pty_farfield_fwd_bwd", d, o.data_ptr(), p.data_ptr(), st.data_ptr(), bidx.data_ptr(), tl.data_ptr(),
           loss.data_ptr(), dobj.data_ptr(), dprb.data_ptr(), ws.data_ptr(), ws.numel(), L.stream_ptr(dev))
    return loss, dobj, dprb


def _noalloc(pkg, dev, o, p, st, tl, bidx, d):
    L = pkg._lib
    B = bidx.numel()
    loss = torch.full((B,), float("nan"), device=dev)
    dobj = torch.full_like(o, float("nan"))
    dprb = torch.full_like(p, float("nan"))
    L.call("pty_farfield_fwd_bwd_noalloc", d, o.data_ptr(), p.data_ptr(), st.data_ptr(), bidx.data_ptr(),
           tl.data_ptr(), loss.data_ptr(), dobj.data_ptr(), dprb.data_ptr(), L.stream_ptr(dev))
    return loss, dobj, dprb


@pytest.mark.gpu
@pytest.mark.parametrize("loss_type", ["FD_loss", "smooth_L1_loss", "Poisson_likelihood_loss"])
def test_noalloc_farfield_matches_workspace_entry_point(pkg, dev, loss_type):
    # a batch that fills most of the scratch, then a smaller one (smaller
    # probe, canvas and batch): the second call must not see the first's data
    cases = [(256, 30, 20, 640, 1), (64, 9, 4, 200, 2)]
    for k, (M, F, B, H, seed) in enumerate(cases):
        o, p, st, tl, bidx, d = _case(pkg, dev, M, F, B, H, loss_type, seed)
        assert pkg._lib.query("pty_farfield_workspace", d) <= FAR_SCRATCH
        ref = _with_workspace(pkg, dev, o, p, st, tl, bidx, d)
        got = _noalloc(pkg, dev, o, p, st, tl, bidx, d)
        torch.cuda.synchronize()
        for name, a, b in zip(("loss", "dobject", "dprobe"), got, ref):
            assert torch.isfinite(torch.view_as_real(a) if a.is_complex() else a).all(), (k, name)
            assert torch.equal(a, b), (k, name)


def test_noalloc_rejects_batches_beyond_the_scratch(pkg):
    """Checked on the host before any device work (no GPU needed)."""
    L = pkg._lib
    d = pkg.farfield.far_descriptor(256, 256, 1024, 1024, 22, 400, "FD_loss")
    assert L.query("pty_farfield_workspace", d) > FAR_SCRATCH
    with pytest.raises(RuntimeError, match="more than the 16777216 of the static scratch"):
        L.call("pty_farfield_fwd_bwd_noalloc", d, 1, 1, 1, 1, 1, 1, 1, 1, None)
    d = pkg.farfield.far_descriptor(256, 256, 1024, 1024, 21, 400, "FD_loss")
    assert L.query("pty_farfield_workspace", d) <= FAR_SCRATCH
    with pytest.raises(RuntimeError, match="null pointer"):
        L.call("pty_farfield_fwd_bwd_noalloc", d, None, 1, 1, 1, 1, 1, 1, 1, None)
