"""The C-ABI library loads and exports every symbol the header declares (CPU only)."""
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "ptycho_hip.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|size_t|const char\*)\s+(pty_\w+)\(", text, flags=re.M)))


def test_library_exports_header_symbols(pkg):
    lib = pkg._lib.load()
    assert lib.pty_abi_version() == pkg._lib.ABI_VERSION
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", str(pkg._lib.LIB_PATH)], capture_output=True, text=True)
    exported = set(re.findall(r" T (pty_\w+)", out.stdout))
    assert set(names) <= exported
    # the ctypes table binds exactly the declared entry points
    assert sorted(pkg._lib.exported_symbols()) == names


def test_errors_are_reported_without_a_gpu(pkg):
    """Descriptor validation runs on the host and reports through pty_last_error."""
    import pytest
    spec = pkg.level_table(16, 2, 14, 16, 1.5)
    d = spec.descriptor()
    d.n_features = 3
    with pytest.raises(RuntimeError, match="n_features_per_level must be 2"):
        pkg._lib.call("pty_hashgrid_fwd", d, 1, 1, 1, 1, None)
    # every corner of an input in [0, 1] must stay within res on each axis
    d = spec.descriptor()
    d.scale[3] = float(d.res[3])          # floor(res + 0.5) + 1 > res
    with pytest.raises(RuntimeError, match="level 3: scale"):
        pkg._lib.call("pty_hashgrid_fwd", d, 1, 1, 1, 1, None)
    d = spec.descriptor()
    d.scale[0] = -1.0
    with pytest.raises(RuntimeError, match="level 0: scale"):
        pkg._lib.call("pty_hashgrid_fwd", d, 1, 1, 1, 1, None)
    d = spec.descriptor()
    d.scale[0], d.res[0], d.size[0] = 0.0, 1, 1   # res (res + 1) = 2 = 2 size
    with pytest.raises(RuntimeError, match="dense level 0"):
        pkg._lib.call("pty_hashgrid_fwd", d, 1, 1, 1, 1, None)
    with pytest.raises(RuntimeError, match="prime factors"):
        pkg._lib.call("pty_upsample_fwd", 1, 67, 2, 1, 1, 1 << 30, None)   # 67 is prime > 64
    far = pkg.farfield.far_descriptor(64, 64, 256, 256, 10, 100, "FD_loss")
    assert pkg._lib.query("pty_farfield_workspace", far) > 0
    # any probe shape whose prime factors are at most 64 and whose LDS tiles fit
    for M, N in ((45, 63), (100, 128), (57, 114), (384, 384), (2048, 2048), (8, 4096)):
        far = pkg.farfield.far_descriptor(M, N, 4096, 4096, 4, 10, "FD_loss")
        assert pkg._lib.query("pty_farfield_workspace", far) >= 4 * M * N * 12, (M, N)
    for M, N, what in ((134, 134, "prime factor"), (64, 67, "prime factor"), (4, 64, "outside"),
                       (64, 5000, "outside")):
        far = pkg.farfield.far_descriptor(M, N, 8192, 8192, 4, 10, "FD_loss")
        assert pkg._lib.query("pty_farfield_workspace", far) == 0, (M, N)
        with pytest.raises(RuntimeError, match=what):
            pkg._lib.call("pty_farfield_fwd_bwd", far, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1 << 30, None)
    with pytest.raises(RuntimeError, match="outside"):
        pkg._lib.call("pty_probe_normalize_fwd", 1, 1, 64, 1, 1, 1, 1 << 20, None)
    with pytest.raises(ValueError):
        pkg.farfield.far_descriptor(64, 64, 256, 256, 10, 100, "l2")
    # batch_next: a batch longer than the order, an empty batch, a missing cursor
    with pytest.raises(RuntimeError, match="batch 8 with an order of 5"):
        pkg._lib.call("pty_batch_next", 1, 5, 1, 8, 1, None, None, None, 0, None)
    with pytest.raises(RuntimeError, match="batch 0"):
        pkg._lib.call("pty_batch_next", 1, 5, 1, 0, 1, None, None, None, 0, None)
    with pytest.raises(RuntimeError, match="null pointer"):
        pkg._lib.call("pty_batch_next", 1, 5, None, 2, 1, None, None, None, 0, None)


def test_table_precision_plumbing(pkg):
    """FP16 table storage: descriptor code, encoder dtype mapping, host checks."""
    import pytest
    import torch
    spec = pkg.level_table(16, 2, 14, 16, 1.5)
    assert spec.descriptor().table_precision == 0
    assert spec.descriptor("fp16").table_precision == 1
    cfg = {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 2, "log2_hashmap_size": 10,
           "base_resolution": 16, "per_level_scale": 1.5}
    assert pkg.HashGridEncoding(2, cfg).table_precision == "fp32"
    enc = pkg.HashGridEncoding(2, cfg, dtype=torch.float16)
    assert enc.table_precision == "fp16" and enc.descriptor().table_precision == 1
    assert enc.params.dtype == torch.float32          # the f32 master stays the parameter
    with pytest.raises(ValueError):
        pkg.HashGridEncoding(2, cfg, dtype=torch.bfloat16)
    with pytest.raises(ValueError):
        pkg.ComplexINRModel2D(32, 32, 2, probe_width=16, probe_height=16, table_precision="fp8")
    d = spec.descriptor()
    d.table_precision = 7
    with pytest.raises(RuntimeError, match="table_precision"):
        pkg._lib.call("pty_hashgrid_fwd", d, 1, 1, 1, 1, None)
    with pytest.raises(RuntimeError, match="null pointer"):
        pkg._lib.call("pty_cast_f16", None, 8, None, None)


def test_model_registration_order(pkg):
    """named_parameters order of the reference (checkpoint format, utils/io.py:11-14)."""
    import torch
    torch.manual_seed(0)
    m = pkg.ComplexINRModel2D(64, 64, 2, probe_width=16, probe_height=16, log2_hashmap_size=10,
                              trainable_omega0=True, use_residual=True,
                              probe_initial=torch.ones(16, 16, dtype=torch.complex64))
    names = [n for n, _ in m.named_parameters()]
    assert names[:3] == ["xy_coords", "probe_coords", "k_probe"]
    assert names[3:5] == ["xy_hash_encoder.params", "probe_hash_encoder.params"]
    assert names[5:8] == ["object_complex_mlp.net.0.omega_0", "object_complex_mlp.net.0.linear.weight",
                          "object_complex_mlp.net.0.linear.bias"]
    assert names[-2:] == ["probe_complex_mlp.net.3.weight", "probe_complex_mlp.net.3.bias"]
    assert m.k_probe.item() == torch.tensor(1e-8, dtype=torch.float32).item()
    assert not m.has_object_residual and m.has_probe_residual
    assert m.xy_hash_encoder.params.numel() == pkg.level_table(16, 2, 10, 16, 1.5).n_params
    with __import__("pytest").raises(ValueError):
        pkg.ComplexINRModel2D(64, 64, 2, use_residual=True, probe_initial=torch.ones(64, 64))


def test_mlp_init_consumes_rng_like_the_reference(pkg):
    """Same seed -> same MLP weights as a plain nn.Linear stack built in reference order."""
    import torch
    import torch.nn as nn
    torch.manual_seed(123)
    m = pkg.ComplexINRModel2D(32, 32, 2, probe_width=16, probe_height=16, log2_hashmap_size=10)
    torch.manual_seed(123)
    expect = []
    for _ in range(2):   # object then probe MLP
        expect.append(nn.Linear(32, 32))
        expect.append(nn.Linear(32, 32, dtype=torch.cfloat))
        expect.append(nn.Linear(32, 32, dtype=torch.cfloat))
        expect.append(nn.Linear(32, 1, dtype=torch.cfloat))
    got = []
    for mlp in (m.object_complex_mlp, m.probe_complex_mlp):
        layers = list(mlp.net)
        got += [layers[0].linear, layers[1].linear, layers[2].linear, layers[3]]
    for a, b in zip(got, expect):
        assert torch.equal(a.weight, b.weight) and torch.equal(a.bias, b.bias)
