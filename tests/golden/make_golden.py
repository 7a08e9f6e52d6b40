"""Write tests/golden/oracle_small.npz: FP64 outputs of the oracle
(oracle/ptycho_oracle.py) for one small seeded problem.

The reference ships no golden vectors for this path (no tests, no data), so
these are regression fixtures of the oracle, which is itself pinned by
tests/test_oracle.py: tests/test_golden.py checks that the oracle still
reproduces them (CPU), and that the HIP kernels match them (GPU) without
running the oracle at test time.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import ptycho_oracle as orc  # noqa: E402

OUT = Path(__file__).resolve().parent / "oracle_small.npz"


def inputs():
    """The seeded inputs (everything the kernels need), as numpy arrays."""
    rng = np.random.default_rng(20261203)
    log2T = 10
    coords = np.concatenate([rng.random((200, 2)), [[0.0, 0.0], [0.999999, 0.5], [0.5, 0.25]]]).astype(np.float32)
    table = orc.level_table(16, log2T, 16, 1.5)
    params = rng.uniform(-1, 1, table["n_entries"] * 2).astype(np.float32)
    w0 = (rng.standard_normal((32, 32)) * 0.3).astype(np.float32)
    b0 = (rng.standard_normal(32) * 0.1).astype(np.float32)
    wh = [((rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))) * 0.15).astype(np.complex64)
          for _ in range(2)]
    bh = [((rng.standard_normal(32) + 1j * rng.standard_normal(32)) * 0.1).astype(np.complex64) for _ in range(2)]
    wo = ((rng.standard_normal((1, 32)) + 1j * rng.standard_normal((1, 32))) * 0.2).astype(np.complex64)
    bo = ((rng.standard_normal(1) + 1j * rng.standard_normal(1)) * 0.1).astype(np.complex64)
    omegas = np.array([10.0, 1.0, 1.0], np.float32)
    small = (rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))).astype(np.complex64)
    M, H, F = 16, 48, 6
    obj = ((0.5 + 0.5 * rng.random((H, H))) * np.exp(1j * rng.uniform(-1, 1, (H, H)))).astype(np.complex64)
    probe = ((rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))) * 0.3).astype(np.complex64)
    tlY = rng.integers(0, H - M + 1, F).astype(np.int64)
    tlX = rng.integers(0, H - M + 1, F).astype(np.int64)
    stack = rng.poisson(5.0, (M, M, F)).astype(np.float32)
    batch = np.array([4, 1, 5, 2], np.int64)
    return dict(log2T=np.int64(log2T), coords=coords, params=params, w0=w0, b0=b0, wh0=wh[0], wh1=wh[1],
                bh0=bh[0], bh1=bh[1], wo=wo, bo=bo, omegas=omegas, small=small, obj=obj, probe=probe, tlY=tlY,
                tlX=tlX, stack=stack, batch=batch)


def outputs(d):
    """The oracle's FP64 outputs for the inputs d."""
    table = orc.level_table(16, int(d["log2T"]), 16, 1.5)
    idx, _ = orc.hash_corners(d["coords"], table)
    feats = orc.hashgrid_encode(d["coords"], torch.from_numpy(d["params"]).double(), table)
    layers = [(torch.from_numpy(d["w0"]).double(), torch.from_numpy(d["b0"]).double(),
               torch.tensor(float(d["omegas"][0]), dtype=torch.float64))]
    for k in range(2):
        layers.append((torch.from_numpy(d[f"wh{k}"]).to(torch.complex128),
                       torch.from_numpy(d[f"bh{k}"]).to(torch.complex128),
                       torch.tensor(float(d["omegas"][k + 1]), dtype=torch.float64)))
    field = orc.complex_mlp(feats, layers, torch.from_numpy(d["wo"]).to(torch.complex128),
                            torch.from_numpy(d["bo"]).to(torch.complex128))[:, 0]
    up = orc.fourier_upsample(torch.from_numpy(d["small"]).to(torch.complex128), 2)
    obj = torch.from_numpy(d["obj"]).to(torch.complex128)
    probe = torch.from_numpy(d["probe"]).to(torch.complex128)
    intens = torch.stack([orc.farfield_intensity(obj, probe, int(d["tlY"][i]), int(d["tlX"][i]))
                          for i in d["batch"]])
    stack = torch.from_numpy(d["stack"]).double()
    losses = {}
    for lt in ("smooth_L1_loss", "FD_loss", "Poisson_likelihood_loss"):
        losses[lt] = np.array([float(orc.batch_loss(obj, probe, stack, [int(i)], d["tlY"], d["tlX"], lt))
                               for i in d["batch"]])
    haar = float(orc.haar_wavelet_sparsity_loss(obj))
    return dict(indices=idx.astype(np.uint32), features=feats.numpy(), field=field.numpy(), upsample=up.numpy(),
                intensities=intens.numpy(), loss_smooth_L1_loss=losses["smooth_L1_loss"],
                loss_FD_loss=losses["FD_loss"], loss_Poisson_likelihood_loss=losses["Poisson_likelihood_loss"],
                haar=np.float64(haar))


def main():
    d = inputs()
    o = outputs(d)
    np.savez_compressed(OUT, **{f"in_{k}": v for k, v in d.items()}, **{f"out_{k}": v for k, v in o.items()})
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
