"""Host logic of the Recovery.py-compatible driver (CPU only): config loading
(Recovery.py:44-72), dataset bundles (utils/io.py:72-122) and checkpoints
(utils/io.py:11-22)."""
import numpy as np
import pytest
import torch

CONFIG = """
data:
  data_dir: data/plant
  bundle_file: dataset-plant.mat
  gap: '2'
physics:
  wavelength: 0.675e-6
  camera_length: 0.0295
  camera_pixel_pitch: '2.96e-5'
model:
  downsample_factor: 2
  log2_hashmap_size: '18'
  per_level_scale: 1.5
  init_type: initProbe
  use_residual: true
training:
  iters: 2000
  batch_size: 10
  lr: '1e-3'
  loss_type: FD_loss
  sparsity_weight: 0.0
output:
  vis_interval: 100
"""


def _write(tmp_path, text):
    p = tmp_path / "cfg.yaml"
    p.write_text(text)
    return str(p)


def test_load_config_flattens_and_coerces(pkg, tmp_path):
    cfg = pkg.recovery.load_config(_write(tmp_path, CONFIG))
    assert cfg["data_dir"] == "data/plant" and cfg["loss_type"] == "FD_loss"
    assert cfg["gap"] == 2 and isinstance(cfg["gap"], int)
    assert cfg["log2_hashmap_size"] == 18 and isinstance(cfg["log2_hashmap_size"], int)
    assert cfg["lr"] == 1e-3 and isinstance(cfg["lr"], float)
    assert cfg["camera_pixel_pitch"] == pytest.approx(2.96e-5)
    assert cfg["wavelength"] == pytest.approx(0.675e-6)
    assert cfg["use_residual"] is True


@pytest.mark.parametrize("edit, message", [
    (("initProbe", "initRandom"), "init_type"),
    (("FD_loss", "l2_loss"), "loss_type"),
    (("use_residual: true", "use_residual: true\n  model_type: gaussian_field"), "gaussian_field"),
    (("use_residual: true", "use_residual: true\n  table_precision: fp8"), "table_precision"),
    (("use_residual: true", "use_residual: true\n  mlp_precision: fp4"), "mlp_precision"),
])
def test_load_config_rejects(pkg, tmp_path, edit, message):
    with pytest.raises(ValueError, match=message):
        pkg.recovery.load_config(_write(tmp_path, CONFIG.replace(*edit)))


def test_load_config_keeps_sparsity_weight(pkg, tmp_path):
    cfg = pkg.recovery.load_config(_write(tmp_path, CONFIG.replace("sparsity_weight: 0.0", "sparsity_weight: 5e-3")))
    assert cfg["sparsity_weight"] == pytest.approx(5e-3) and isinstance(cfg["sparsity_weight"], float)


def test_bundle_loader_keys_squeeze_and_gap(pkg, tmp_path):
    from scipy.io import savemat
    rng = np.random.default_rng(0)
    x = rng.random((1, 12))
    y = rng.random((12, 1))
    raw = rng.random((8, 8, 12)).astype(np.float32)
    probe = (rng.random((8, 8)) + 1j * rng.random((8, 8))).astype(np.complex64)
    savemat(str(tmp_path / "bundle.mat"), {"xlocation": x, "ylocation": y, "imRaw": raw, "probe": probe})
    d = pkg.recovery.load_ptychography_data(str(tmp_path), bundle_file="bundle.mat")
    assert d["xlocation"].shape == (12,) and d["ylocation"].shape == (12,)
    np.testing.assert_array_equal(d["imRaw"], raw)
    np.testing.assert_array_equal(d["probe"], probe)
    assert "obj" not in d
    d2 = pkg.recovery.load_ptychography_data(str(tmp_path), bundle_file="bundle.mat", gap=3)
    np.testing.assert_array_equal(d2["xlocation"], x.squeeze()[::3])
    np.testing.assert_array_equal(d2["imRaw"], raw[:, :, ::3])


def test_checkpoint_round_trip(pkg, tmp_path):
    """save_model writes the requires_grad parameters in named_parameters order;
    load_model restores them into a freshly built model."""
    def make(seed):
        torch.manual_seed(seed)
        return pkg.ComplexINRModel2D(32, 32, 2, probe_width=16, probe_height=16, log2_hashmap_size=10,
                                     trainable_omega0=True, use_residual=True,
                                     probe_initial=torch.ones(16, 16, dtype=torch.complex64))
    a, b = make(1), make(2)
    with torch.no_grad():
        a.xy_hash_encoder.params.uniform_(-1, 1)
    path = str(tmp_path / "trained_models.pth")
    pkg.recovery.save_model(a, path)
    saved = torch.load(path, weights_only=True)
    names = [n for n, p in a.named_parameters() if p.requires_grad]
    assert len(saved) == len(names)
    assert "xy_coords" not in names and names[0] == "k_probe"
    pkg.recovery.load_model(b, path)
    for (na, pa), (nb, pb) in zip(a.named_parameters(), b.named_parameters()):
        assert na == nb
        if pa.requires_grad:
            assert torch.equal(pa, pb), na


def test_write_results_round_trip(pkg, tmp_path):
    """result.mat (object_complex, probe_complex) and trained_models.pth as
    Recovery.py:377-383 writes them; read back with scipy and load_model."""
    from types import SimpleNamespace

    from scipy.io import loadmat
    torch.manual_seed(0)
    model = pkg.ComplexINRModel2D(32, 32, 2, probe_width=16, probe_height=16, log2_hashmap_size=10,
                                  use_residual=True, probe_initial=torch.ones(16, 16, dtype=torch.complex64))
    g = torch.Generator().manual_seed(1)
    obj = torch.randn(32, 32, dtype=torch.complex64, generator=g)
    probe = torch.randn(16, 16, dtype=torch.complex64, generator=g)
    eng = SimpleNamespace(obj=obj, probe_n=probe)
    pkg.recovery.write_results(str(tmp_path / "res"), eng, model)
    m = loadmat(str(tmp_path / "res" / "result.mat"))
    assert m["object_complex"].dtype == np.complex64 and m["object_complex"].shape == (32, 32)
    assert np.array_equal(m["object_complex"], obj.numpy())
    assert np.array_equal(m["probe_complex"], probe.numpy())
    torch.manual_seed(5)
    other = pkg.ComplexINRModel2D(32, 32, 2, probe_width=16, probe_height=16, log2_hashmap_size=10,
                                  use_residual=True, probe_initial=torch.ones(16, 16, dtype=torch.complex64))
    pkg.recovery.load_model(other, str(tmp_path / "res" / "trained_models.pth"))
    for (n1, a), (n2, b) in zip(model.named_parameters(), other.named_parameters()):
        assert n1 == n2 and torch.equal(a, b), n1
