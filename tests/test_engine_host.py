"""Host-side input checks of the fused step (no GPU: they run before any
device work).  The kernels index the measured stack, the scan windows and the
residual initial fields without bounds checks, so a mismatch must be a
ValueError on the host, as the reference's own torch ops would raise."""
import numpy as np
import pytest
import torch


def _model(pkg, canvas=32, M=16, probe_initial=True, object_initial=None):
    torch.manual_seed(0)
    pi = torch.ones(M, M, dtype=torch.complex64) if probe_initial is True else probe_initial
    return pkg.ComplexINRModel2D(canvas, canvas, 2, probe_width=M, probe_height=M, log2_hashmap_size=10,
                                 use_residual=True, probe_initial=pi, object_initial=object_initial)


def _windows(n):
    return np.arange(n) % 8, (np.arange(n) * 3) % 8


def test_good_inputs_reach_the_device_check(pkg):
    m = _model(pkg)
    tlY, tlX = _windows(5)
    with pytest.raises(RuntimeError, match="on the GPU"):
        pkg.FusedReconstruction(m, torch.zeros(16, 16, 5), tlY, tlX, batch_size=2)


@pytest.mark.parametrize("shape", [(16, 16), (16, 8, 5), (8, 16, 5), (5, 16, 16)])
def test_stack_shape_is_checked(pkg, shape):
    m = _model(pkg)
    tlY, tlX = _windows(5)
    with pytest.raises(ValueError, match="measured stack"):
        pkg.FusedReconstruction(m, torch.zeros(*shape), tlY, tlX)


@pytest.mark.parametrize("ny,nx", [(4, 5), (5, 4), (6, 6)])
def test_window_count_is_checked(pkg, ny, nx):
    m = _model(pkg)
    tlY, _ = _windows(ny)
    _, tlX = _windows(nx)
    with pytest.raises(ValueError, match="scan windows"):
        pkg.FusedReconstruction(m, torch.zeros(16, 16, 5), tlY, tlX)


def test_windows_outside_the_canvas_are_rejected(pkg):
    m = _model(pkg)
    with pytest.raises(ValueError, match="outside the canvas"):
        pkg.FusedReconstruction(m, torch.zeros(16, 16, 2), [0, 17], [0, 0])


def test_initial_shapes_are_checked(pkg):
    m = _model(pkg, probe_initial=torch.ones(16, 8, dtype=torch.complex64))
    tlY, tlX = _windows(5)
    with pytest.raises(ValueError, match="probe_initial has shape"):
        pkg.FusedReconstruction(m, torch.zeros(16, 16, 5), tlY, tlX)
    m = _model(pkg, object_initial=torch.ones(30, 32, dtype=torch.complex64))
    with pytest.raises(ValueError, match="object_initial has shape"):
        pkg.FusedReconstruction(m, torch.zeros(16, 16, 5), tlY, tlX)


def test_complex128_initial_is_accepted(pkg):
    """A MATLAB double probe (the reference's initProbe fallback) arrives as
    complex128: it is rounded once to complex64, not misread."""
    m = _model(pkg, probe_initial=torch.ones(16, 16, dtype=torch.complex128))
    tlY, tlX = _windows(5)
    with pytest.raises(RuntimeError, match="on the GPU"):
        pkg.FusedReconstruction(m, torch.zeros(16, 16, 5), tlY, tlX)
    from importlib import import_module
    eng = import_module(pkg.__name__ + ".engine")
    t = eng._check_initial("probe_initial", torch.full((4, 4), 1 / 3 + 0.25j, dtype=torch.complex128), (4, 4))
    assert t.dtype == torch.complex64 and t.is_contiguous()
    assert torch.equal(t, torch.full((4, 4), 1 / 3 + 0.25j, dtype=torch.complex128).to(torch.complex64))


def test_batch_size_is_checked(pkg):
    m = _model(pkg)
    tlY, tlX = _windows(5)
    with pytest.raises(ValueError, match="batch_size"):
        pkg.FusedReconstruction(m, torch.zeros(16, 16, 5), tlY, tlX, batch_size=0)


def test_changed_coordinates_are_rejected_by_the_fused_step(pkg):
    m = _model(pkg)
    assert m.coords_are_grid("object") and m.coords_are_grid("probe")
    orig = m.xy_coords[5, 0].item()
    with torch.no_grad():
        m.xy_coords[5, 0] += 0.01
    assert not m.coords_are_grid("object") and m.coords_are_grid("probe")
    tlY, tlX = _windows(5)
    with pytest.raises(ValueError, match="object coordinates were changed"):
        pkg.FusedReconstruction(m, torch.zeros(16, 16, 5), tlY, tlX)
    with torch.no_grad():
        m.xy_coords[5, 0] = orig
    assert m.coords_are_grid("object")


def test_fp8_is_flagged_experimental(pkg):
    """FP8 single steps meet their bars, but FP8 reconstructions do not converge
    (the precision sweep along the FP64 trajectory): selecting it warns."""
    import warnings
    for prec, warns in (("fp8", True), ("bf16", False), ("fp32", False)):
        with warnings.catch_warnings(record=True) as rec:
            warnings.simplefilter("always")
            torch.manual_seed(0)
            pkg.ComplexINRModel2D(32, 32, 2, probe_width=16, probe_height=16, log2_hashmap_size=10,
                                  mlp_precision=prec)
        hit = [w for w in rec if "experimental" in str(w.message)]
        assert bool(hit) == warns, (prec, [str(w.message) for w in rec])
