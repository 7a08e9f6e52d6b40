"""Shared test helpers: copy a product model's parameters into oracle form and
evaluate the oracle step (FP64, CPU) with autograd gradients."""
from __future__ import annotations

import numpy as np
import torch

from oracle import ptycho_oracle as orc


def _leaf(t, dtype):
    return t.detach().cpu().to(dtype).clone().requires_grad_(True)


def mlp_params(mlp, dtype=torch.float64):
    cdtype = torch.complex128 if dtype == torch.float64 else torch.complex64
    layers = list(mlp.net)
    out = {"layers": [], "out_weight": _leaf(layers[-1].weight, cdtype), "out_bias": _leaf(layers[-1].bias, cdtype)}
    for i, l in enumerate(layers[:-1]):
        wd = dtype if i == 0 else cdtype
        out["layers"].append((_leaf(l.linear.weight, wd), _leaf(l.linear.bias, wd), _leaf(l.omega_0, dtype)))
    return out


def oracle_params(model, dtype=torch.float64):
    p = {
        "obj_table": _leaf(model.xy_hash_encoder.params, dtype),
        "prb_table": _leaf(model.probe_hash_encoder.params, dtype),
        "obj_mlp": mlp_params(model.object_complex_mlp, dtype),
        "prb_mlp": mlp_params(model.probe_complex_mlp, dtype),
    }
    if model.has_probe_residual:
        p["k_probe"] = _leaf(model.k_probe, dtype)
        p["probe_initial"] = model.probe_initial.detach().cpu().to(torch.complex128)
    if model.has_object_residual:
        p["k_object"] = _leaf(model.k_object, dtype)
        p["object_initial"] = model.object_initial.detach().cpu().to(torch.complex128)
    return p


def oracle_table(model):
    e = model.xy_hash_encoder.encoding_config
    return orc.level_table(e["n_levels"], e["log2_hashmap_size"], e["base_resolution"], e["per_level_scale"])


def oracle_geom(model, tlY, tlX):
    return dict(obj_coords=model.xy_coords.detach().cpu(), prb_coords=model.probe_coords.detach().cpu(),
                ds_w=model.downsampled_width, ds_h=model.downsampled_height, probe_w=model.probe_width,
                probe_h=model.probe_height, ds=model.downsample_factor, tlY=np.asarray(tlY), tlX=np.asarray(tlX))


def flat_grads(params: dict) -> dict:
    g = {"obj_table": params["obj_table"].grad, "prb_table": params["prb_table"].grad}
    for key in ("obj_mlp", "prb_mlp"):
        m = params[key]
        lst = []
        for (w, b, om) in m["layers"]:
            lst += [w.grad, b.grad, om.grad]
        lst += [m["out_weight"].grad, m["out_bias"].grad]
        g[key] = lst
    for key in ("k_probe", "k_object"):
        if key in params:
            g[key] = params[key].grad
    return g


def step_reference(model, params, inst, batch, loss_type, sparsity_weight=0.0, dtype=torch.float64):
    """Oracle loss and gradients of one step (Recovery.py:253-316), in FP64
    (the parity reference) or, with dtype=float32 and FP32 parameter leaves,
    the same autograd computation in FP32 (its rounding floor)."""
    table = oracle_table(model)
    geom = oracle_geom(model, inst.tlY, inst.tlX)
    stack = torch.from_numpy(np.asarray(inst.stack)).to(dtype)
    loss, _, _ = orc.step_loss(params, table, geom, stack, batch, loss_type, dtype=dtype,
                               sparsity_weight=sparsity_weight)
    loss.backward()
    return float(loss.detach()), flat_grads(params)


def rel_err(a, b) -> float:
    a = torch.as_tensor(a).detach().cpu().to(torch.complex128)
    b = torch.as_tensor(b).detach().cpu().to(torch.complex128)
    den = b.abs().pow(2).sum().sqrt().item()
    num = (a - b).abs().pow(2).sum().sqrt().item()
    return num / den if den > 0 else num


def boundary_coords(table, per_level=40, seed=0):
    """f32 inputs whose position x*scale + 0.5 lies within a few ulps of an
    integer (a cell boundary) on some level: where a second rounding of the
    product could move floor() to the other cell."""
    rng = np.random.default_rng(seed)
    xs = []
    for l in range(table["n_levels"]):
        s = float(table["scale"][l])
        k = rng.integers(1, max(2, int(s)), per_level)
        x0 = ((k - 0.5) / s).astype(np.float32)
        for d in range(-3, 4):
            v = x0.copy()
            for _ in range(abs(d)):
                v = np.nextafter(v, np.float32(np.inf if d > 0 else -np.inf))
            xs.append(v)
    x = np.concatenate(xs)
    x = x[(x > 0) & (x < 1)]
    return x


def omega_error(g_ours, g_ref, gW_ref, W, gb_ref, b, omega) -> float:
    """Error of a frequency gradient dL/domega.  By the homogeneity of a
    Gabor layer in its weights, omega dL/domega = Re<G_W, W> + Re<G_b, b>: a
    sum of terms that cancel, so the error is measured against the
    Cauchy-Schwarz scale (|G_W| |W| + |G_b| |b|) / |omega| of that identity
    rather than against the (possibly tiny) result."""
    def nrm(t):
        return torch.as_tensor(t).detach().cpu().to(torch.complex128).abs().pow(2).sum().sqrt().item()
    scale = (nrm(gW_ref) * nrm(W) + nrm(gb_ref) * nrm(b)) / abs(float(torch.as_tensor(omega).detach().cpu()))
    d = abs(float(torch.as_tensor(g_ours).detach().cpu().double()) - float(torch.as_tensor(g_ref).detach().cpu().double()))
    return d / scale if scale > 0 else d
