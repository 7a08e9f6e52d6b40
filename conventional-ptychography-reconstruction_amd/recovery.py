"""Recovery.py-compatible configuration and reconstruction driver.

Reference: Recovery.py:39-72 (load_config: one-level flatten of the YAML
sections, float/int coercion lists, enum validation), :80-158 (defaults,
init_type logic), :201-222 (model construction), :227-350 (training loop:
per-epoch randperm, ceil(frames/batch) Adam steps, StepLR per epoch).
Additive keys: ``canvas_size`` (explicit square canvas; the scan is centred
on it), ``mlp_precision`` ("fp32" | "bf16" | "fp8"), ``table_precision`` ("fp32" |
"fp16", the hash-table storage the kernels gather from), ``seed``.
"""
from __future__ import annotations

import os

import numpy as np
import torch
import yaml

from .complex_inr import ComplexINRModel2D
from .engine import FusedReconstruction
from .optics import center_windows, scan_positions_to_pixels

VALID_INIT_TYPES = ("initNone", "initProbe", "initQuadratic")
VALID_LOSS_TYPES = ("smooth_L1_loss", "FD_loss", "Poisson_likelihood_loss")
VALID_MODEL_TYPES = ("inr", "gaussian_field")

_FLOAT_KEYS = ("wavelength", "camera_length", "camera_pixel_pitch", "lr", "lr_decay_gamma", "per_level_scale",
               "first_omega_0", "hidden_omega_0", "sparsity_weight", "quadratic_focal_length")
_INT_KEYS = ("gap", "downsample_factor", "iters", "lr_decay_step", "n_levels", "n_features_per_level",
             "log2_hashmap_size", "base_resolution", "hidden_features", "hidden_layers", "vis_interval",
             "batch_size", "zoom_size", "canvas_size", "seed")


def normalize_config(raw: dict) -> dict:
    cfg = {}
    for section in raw.values():
        if isinstance(section, dict):
            cfg.update(section)
    for k in _FLOAT_KEYS:
        if k in cfg and cfg[k] is not None:
            cfg[k] = float(cfg[k])
    for k in _INT_KEYS:
        if k in cfg and cfg[k] is not None:
            cfg[k] = int(cfg[k])
    init_type = cfg.get("init_type", "initNone")
    if init_type not in VALID_INIT_TYPES:
        raise ValueError(f"init_type must be one of {VALID_INIT_TYPES}, got {init_type!r}")
    loss_type = cfg.get("loss_type", "smooth_L1_loss")
    if loss_type not in VALID_LOSS_TYPES:
        raise ValueError(f"loss_type must be one of {VALID_LOSS_TYPES}, got {loss_type!r}")
    model_type = cfg.get("model_type", "inr")
    if model_type not in VALID_MODEL_TYPES:
        raise ValueError(f"model_type must be one of {VALID_MODEL_TYPES}, got {model_type!r}")
    if model_type != "inr":
        raise ValueError("model_type 'gaussian_field' is not part of the gfx950 build")
    for key in ("mlp_precision",):
        if cfg.get(key, "fp32") not in ("fp32", "bf16", "fp8"):
            raise ValueError(f"{key} must be 'fp32', 'bf16' or 'fp8'")
    if cfg.get("table_precision", "fp32") not in ("fp32", "fp16"):
        raise ValueError("table_precision must be 'fp32' or 'fp16'")
    return cfg


def load_config(path: str) -> dict:
    with open(path, "r", encoding="utf-8") as f:
        raw = yaml.safe_load(f)
    return normalize_config(raw)


def quadratic_phase_probe(probe, dx, wavelength, focal_length):
    """exp(i pi / (lambda L) (X^2 + Y^2)) * probe (utils/optics.py:115-148), host numpy."""
    arr = probe.detach().cpu().numpy() if torch.is_tensor(probe) else np.asarray(probe)
    M, N = arr.shape
    dy = dx[0] if hasattr(dx, "__len__") else float(dx)
    dxx = dx[1] if hasattr(dx, "__len__") else float(dx)
    yp = (np.arange(M, dtype=np.float32) - M / 2) * dy
    xp = (np.arange(N, dtype=np.float32) - N / 2) * dxx
    X, Y = np.meshgrid(xp, yp)
    phase = np.pi / (wavelength * focal_length) * (X ** 2 + Y ** 2)
    return (np.exp(1j * phase).astype(np.complex64) * arr.astype(np.complex64))


def build(cfg: dict, data: dict, device="cuda", object_initial=None):
    """Model + fused step for a dataset dict (keys of load_ptychography_data)."""
    probe = np.asarray(data["probe"])
    M, N = probe.shape
    tlX, tlY, brX, brY, dx, canvas = scan_positions_to_pixels(
        data["xlocation"], data["ylocation"], probe.shape, cfg["wavelength"], cfg["camera_length"],
        cfg["camera_pixel_pitch"])
    if cfg.get("canvas_size"):
        canvas = int(cfg["canvas_size"])
        tlX, tlY = center_windows(tlX, tlY, probe.shape, canvas)
    init_type = cfg.get("init_type", "initNone")
    use_residual = cfg.get("use_residual", False)
    init_probe = data.get("initProbe")
    if init_type == "initQuadratic":
        base = init_probe if init_probe is not None else probe
        probe_initial = torch.from_numpy(quadratic_phase_probe(base, dx, cfg["wavelength"],
                                                               cfg.get("quadratic_focal_length", 0.1)))
    elif init_type == "initNone":
        probe_initial = None
        use_residual = False
    else:
        probe_initial = torch.from_numpy(np.asarray(init_probe if init_probe is not None else probe)
                                         .astype(np.complex64))
    if cfg.get("seed") is not None:
        torch.manual_seed(int(cfg["seed"]))
    model = ComplexINRModel2D(
        output_width=canvas, output_height=canvas, downsample_factor=cfg.get("downsample_factor", 2),
        update_probe=True, probe_width=N, probe_height=M, use_residual=use_residual,
        object_initial=object_initial, probe_initial=probe_initial,
        n_levels=cfg.get("n_levels", 16), n_features_per_level=cfg.get("n_features_per_level", 2),
        log2_hashmap_size=cfg.get("log2_hashmap_size", 18), base_resolution=cfg.get("base_resolution", 16),
        per_level_scale=cfg.get("per_level_scale", 1.5), first_omega_0=cfg.get("first_omega_0", 10.0),
        hidden_omega_0=cfg.get("hidden_omega_0", 1.0), hidden_features=cfg.get("hidden_features", 64),
        hidden_layers=cfg.get("hidden_layers", 2), trainable_omega0=cfg.get("trainable_omega0", True),
        mlp_precision=cfg.get("mlp_precision", "fp32"), table_precision=cfg.get("table_precision", "fp32"),
    ).to(device)
    stack = torch.as_tensor(np.asarray(data["imRaw"]), dtype=torch.float32).to(device)
    engine = FusedReconstruction(model, stack, tlY, tlX, loss_type=cfg.get("loss_type", "smooth_L1_loss"),
                                 lr=cfg.get("lr", 1e-3), batch_size=cfg.get("batch_size", 10),
                                 lr_decay_step=cfg.get("lr_decay_step", 10000),
                                 lr_decay_gamma=cfg.get("lr_decay_gamma", 0.1),
                                 sparsity_weight=cfg.get("sparsity_weight", 0.0) or 0.0)
    geom = dict(tlX=np.asarray(tlX), tlY=np.asarray(tlY), dx=dx, canvas=canvas, M=M, N=N)
    return model, engine, geom


def write_results(result_dir, engine, model):
    """The files Recovery.py:377-383 writes every ``vis_interval`` epochs (its
    PNG plots aside): ``result.mat`` with ``object_complex`` and
    ``probe_complex`` -- the object and the centred, energy-normalised probe
    of the epoch's last step, as the reference's loop variables hold them
    then (computed before that step's Adam update) -- and
    ``trained_models.pth`` (save_model: the requires_grad parameters in
    named_parameters order)."""
    from scipy.io import savemat
    os.makedirs(result_dir, exist_ok=True)
    savemat(os.path.join(result_dir, "result.mat"), {
        "object_complex": engine.obj.detach().squeeze().cpu().numpy(),
        "probe_complex": engine.probe_n.detach().squeeze().cpu().numpy(),
    })
    save_model(model, os.path.join(result_dir, "trained_models.pth"))


def reconstruct(cfg: dict, data: dict, device="cuda", epochs: int | None = None, callback=None):
    """Run ``iters`` epochs of the fused step; returns (model, engine, geom, losses).

    With ``result_dir`` in the config (Recovery.py:140-141), the results are
    written there after epoch 0, every ``vis_interval`` epochs (default 100)
    and after the last epoch (Recovery.py:352), as ``write_results`` does."""
    model, engine, geom = build(cfg, data, device)
    epochs = cfg.get("iters", 2000) if epochs is None else epochs
    result_dir = cfg.get("result_dir")
    vis_interval = int(cfg.get("vis_interval", 100) or 100)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(cfg.get("seed", 0) or 0))
    losses = []
    for ep in range(epochs):
        losses.append(engine.run_epoch(gen))
        if result_dir and (ep % vis_interval == 0 or ep == epochs - 1):
            write_results(result_dir, engine, model)
        if callback is not None:
            callback(ep, engine)
    return model, engine, geom, torch.stack(losses).cpu().numpy() if losses else np.zeros(0)


def read_data(file_path):
    """MATLAB v5 .mat reader (utils/io.py:25-69); v7.3 needs h5py, which is not bundled."""
    from scipy.io import loadmat
    try:
        return loadmat(file_path)
    except Exception as exc:  # v7.3 (HDF5) files
        raise ImportError("h5py is required to read MATLAB v7.3 files") from exc


def load_ptychography_data(data_dir, bundle_file=None, location_file=None, raw_file=None, probe_file=None,
                           gt_file=None, init_probe_file=None, loc_x_key="xlocation", loc_y_key="ylocation",
                           raw_key="imRaw", probe_key="probe", gt_key="obj", init_probe_key="initProbe", gap=1):
    """Dataset loader with the key names and precedence of utils/io.py:72-122."""
    out = {}
    if bundle_file is not None:
        bundle = read_data(os.path.join(data_dir, bundle_file))
        for k in (loc_x_key, loc_y_key, probe_key, raw_key, gt_key, init_probe_key):
            if k in bundle:
                out[k] = bundle[k]
    if location_file is not None:
        loc = read_data(os.path.join(data_dir, location_file))
        out[loc_x_key] = loc[loc_x_key]
        out[loc_y_key] = loc[loc_y_key]
    if raw_file is not None:
        out[raw_key] = read_data(os.path.join(data_dir, raw_file))[raw_key]
    if probe_file is not None:
        out[probe_key] = read_data(os.path.join(data_dir, probe_file))[probe_key]
    if gt_file is not None:
        gt = read_data(os.path.join(data_dir, gt_file))
        out[gt_key] = gt.get(gt_key, next(iter(v for k, v in gt.items() if not k.startswith("__")), None))
    if init_probe_file is not None:
        ip = read_data(os.path.join(data_dir, init_probe_file))
        out[init_probe_key] = ip.get(init_probe_key,
                                     next(iter(v for k, v in ip.items() if not k.startswith("__")), None))
    out[loc_x_key] = np.asarray(out[loc_x_key]).squeeze()
    out[loc_y_key] = np.asarray(out[loc_y_key]).squeeze()
    if gap and gap > 1:
        out[loc_x_key] = out[loc_x_key][::gap]
        out[loc_y_key] = out[loc_y_key][::gap]
        out[raw_key] = out[raw_key][:, :, ::gap]
    return out


def save_model(model, save_path):
    """List of requires_grad parameters in named_parameters order (utils/io.py:11-14)."""
    tensors = [p.detach().cpu().clone() for _, p in model.named_parameters() if p.requires_grad]
    torch.save(tensors, save_path)


def load_model(model, load_path):
    """Inverse of save_model (utils/io.py:17-22); loads tensors only (weights_only)."""
    tensors = torch.load(load_path, weights_only=True)
    for _, param in model.named_parameters():
        if param.requires_grad:
            param.data.copy_(tensors.pop(0).to(param.device))
