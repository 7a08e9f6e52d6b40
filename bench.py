"""Throughput of the fused INR-ptychography training step on one MI355X.

Default workload (the reference-style "plant-like" reconstruction): canvas
1024x1024 (INR evaluated on 512x512 and Fourier-upsampled), probe 256x256,
20x20 scan with 40 px step and +/-4 px jitter (400 positions), 1e5 photons
per pattern with Poisson noise, HashGrid 16 levels x 2 features with
log2_hashmap_size 18, complex-exp MLP 64 wide with 2 hidden layers (omega
10 / 1), initProbe + use_residual, smooth-L1 loss, Adam lr 1e-3, batch 10,
FP32.  Synthetic seeded data (no dataset is shipped with the reference).
--workload cfg3 runs the large case (canvas 3648, INR 1824x1824, probe 512,
50x50 scan, batch 64, FD loss, log2_hashmap_size 20); cfg0 the small one.

A "step" is one optimizer step of Recovery.py:247-321 (whole-canvas INR
forward, one batch of positions, backward, Adam); the reference's "iters"
are epochs = ceil(positions/batch) steps.  value = iters/s = epochs per
second.

Two timed regions of K steps each, both bracketed by a synchronisation:
(1) eager steps (the host enqueues every launch) with the object and probe
chains on one stream and HIP events around each launch of the dominant
kernel; (2) the same steps replayed from the captured HIP graph with the
probe chain on its own stream, which is how FusedReconstruction.run_epoch
(the reconstruction loop) runs full batches; value and ms_per_step come from
(2) (from (1) with --no-graph), and eager_ms_per_step reports (1).

Also reported: the roofline of the dominant kernel (the MLP backward with
recompute, k_inr_bwd16 in FP32, launched once for the object grid and once for the
probe grid per step; its average launch duration from the events of region
(1)), its HBM traffic from the committed rocprofv3 PMC summary, and the CPU
baseline (the FP32 oracle restatement on the host cores, a bounded sample of
steps).  With the CPU baseline the line also carries `nrmse_leg`, the
wall-clock half of the metric measured after the timed regions on the
canvas-256 shapes (the plain PyTorch loop takes about 10 s per epoch at
canvas 1024): the plain loop for --nrmse-epochs epochs, then the fused engine
from the same initial model until it reaches the plain run's final object
NRMSE (within 1 %).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

WORKLOAD = dict(canvas=1024, M=256, n_side=20, step=40, jitter=4, photons=1e5)
LOG2T = 18
BATCH = 10
# throughput workloads: (generator shapes, log2 hashmap size, batch, loss, name)
WORKLOADS = {
    "cfg0": (dict(canvas=256, M=64, n_side=10, step=16, jitter=2, photons=1e4), 14, 10, "smooth_L1_loss",
             "canvas256_inr128_probe64_scan10x10_batch10_log2T14_smoothL1"),
    "cfg1": (WORKLOAD, LOG2T, BATCH, "smooth_L1_loss",
             "canvas1024_inr512_probe256_scan20x20_batch10_log2T18_smoothL1"),
    "cfg3": (dict(canvas=3648, M=512, n_side=50, step=64, jitter=6, photons=1e6), 20, 64, "FD_loss",
             "canvas3648_inr1824_probe512_scan50x50_batch64_log2T20_FD"),
}
MLP_FLOP_BWD_PER_POINT = 37376      # dX + dW of the complex MLP per point (recompute not counted)
MLP_FLOP_FWD_PER_POINT = 18688
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0}     # MI355X dense
MLP_BWD_KERNEL = {"fp32": "k_inr_bwd16<2>", "bf16": "k_inr_bwd_b16<2>",
                  "fp8": "k_inr_bwd_lp<pty::FmtFp8, 2>"}
PROFILES = ROOT / "profiles"


def measured_traffic(workload, precision, table_precision="fp32"):
    """HBM bytes per launch of the MLP-backward kernel from the committed
    rocprofv3 --pmc summary of this workload (FETCH_SIZE and WRITE_SIZE,
    scaled by a known-size copy, f32 tables), or None when there is none."""
    if table_precision != "fp32":
        return None
    path = PROFILES / f"pmc_mlp_bwd_{workload}_{precision}.json"
    if not path.exists():
        return None
    d = json.loads(path.read_text())
    return d.get("hbm_bytes_per_launch")


def pmc_summary(fetch_csv, write_csv, workload, precision, command):
    """Write profiles/pmc_mlp_bwd_<workload>_<precision>.{json,md} from two
    rocprofv3 --pmc counter_collection CSVs (FETCH_SIZE, WRITE_SIZE; KiB per
    dispatch).  Each counter is scaled by what it reads for the 256 MiB
    calibration copy (--copy-calibration) against the copy's true bytes."""
    import csv
    true_kib = 256 * 1024

    def read(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        cal = [float(r["Counter_Value"]) for r in rows
               if "copyBuffer" in r["Kernel_Name"] and float(r["Counter_Value"]) > true_kib / 4]
        if not cal:
            raise SystemExit(f"{path}: no calibration copy dispatch found")
        factor = true_kib / (sum(cal) / len(cal))
        import re
        key = re.compile(r"\b" + re.escape(MLP_BWD_KERNEL[precision]) + r"\(")
        # each step launches the kernel once for the object grid and once for
        # the probe grid (on two streams, so in either order); the object grid
        # has the larger point count, so the larger dispatch of each pair is
        # the object's
        per = {"object": [], "probe": []}
        hits = [r for r in rows if key.search(r["Kernel_Name"])]
        if not hits or len(hits) % 2:
            raise SystemExit(f"{path}: expected pairs of {key.pattern} dispatches, found {len(hits)}")
        hits.sort(key=lambda r: int(r["Dispatch_Id"]))
        for i in range(0, len(hits), 2):
            a, b = (float(hits[i + q]["Counter_Value"]) * factor * 1024 for q in range(2))
            per["object"].append(max(a, b))
            per["probe"].append(min(a, b))
        return factor, {g: sum(v) / len(v) for g, v in per.items()}, len(hits)

    ff, fetch, nf = read(fetch_csv, "FETCH_SIZE")
    wf, write, nw = read(write_csv, "WRITE_SIZE")
    grids = ["object", "probe"]
    per_launch = {g: {"fetch_bytes": fetch[g], "write_bytes": write[g]} for g in grids}
    mean = sum(fetch[g] + write[g] for g in grids) / len(grids)
    out = {"kernel": MLP_BWD_KERNEL[precision], "workload": workload, "precision": precision,
           "hbm_bytes_per_launch": mean,
           "note": "mean over the object-grid and probe-grid launches",
           "per_launch": per_launch, "dispatches": {"FETCH_SIZE": nf, "WRITE_SIZE": nw},
           "calibration": {"copy_bytes": true_kib * 1024, "fetch_factor": ff, "write_factor": wf},
           "command": command}
    stem = PROFILES / f"pmc_mlp_bwd_{workload}_{precision}"
    stem.with_suffix(".json").write_text(json.dumps(out, indent=1) + "\n")
    lines = [f"# HBM traffic of {MLP_BWD_KERNEL[precision]} ({workload}, {precision} MLP)", "",
             "Two rocprofv3 PMC passes (one counter each), run from the repository root on one MI355X:", "",
             "```", command, "```", "",
             f"Calibration: a 256 MiB device copy reads FETCH_SIZE x{ff:.4f} and WRITE_SIZE x{wf:.4f} of its "
             "true bytes; the kernel's counters are scaled by these factors.", "",
             "| launch | FETCH MB/launch | WRITE MB/launch |", "|---|---|---|"]
    for g in grids:
        lines.append(f"| {g} | {fetch[g] / 1e6:.2f} | {write[g] / 1e6:.2f} |")
    lines += ["", f"Mean HBM bytes per launch (fetch + write): {mean / 1e6:.2f} MB"]
    stem.with_suffix(".md").write_text("\n".join(lines) + "\n")
    print(json.dumps(out))


def copy_calibration(dev, mib=256, reps=3):
    """A plain device copy of known size, for scaling the HBM byte counters."""
    import torch
    a = torch.empty(mib * 1024 * 1024 // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def build_problem(pkg, device, seed=0, precision="fp32", workload="cfg1", table_precision="fp32"):
    import numpy as np
    import torch
    shapes, log2T, _, _, _ = WORKLOADS[workload]
    inst = pkg.synthetic.make_instance(seed=seed, **shapes)
    M = inst.probe.shape[0]
    torch.manual_seed(seed)
    model = pkg.ComplexINRModel2D(inst.canvas, inst.canvas, 2, probe_width=M, probe_height=M,
                                  log2_hashmap_size=log2T, hidden_omega_0=1.0, trainable_omega0=True,
                                  use_residual=True, mlp_precision=precision, table_precision=table_precision,
                                  probe_initial=torch.from_numpy(inst.probe.astype(np.complex64)))
    return inst, model.to(device)


def cpu_baseline(inst, seconds=15.0):
    """Oracle step (autograd + Adam) in FP32 on the host cores, bounded sample."""
    import numpy as np
    import torch
    import __graft_entry__
    from oracle import ptycho_oracle as orc
    from tests.helpers import oracle_geom, oracle_params, oracle_table

    pkg = __graft_entry__.load_package()
    torch.set_num_threads(os.cpu_count() or 1)
    _, model = build_problem(pkg, "cpu")
    params = oracle_params(model, dtype=torch.float32)
    if "probe_initial" in params:
        params["probe_initial"] = params["probe_initial"].to(torch.complex64)
    leaves = [params["obj_table"], params["prb_table"], params["k_probe"]]
    for key in ("obj_mlp", "prb_mlp"):
        for (w, b, om) in params[key]["layers"]:
            leaves += [w, b, om]
        leaves += [params[key]["out_weight"], params[key]["out_bias"]]
    opt = torch.optim.Adam(leaves, lr=1e-3)
    table = oracle_table(model)
    geom = oracle_geom(model, inst.tlY, inst.tlX)
    stack = torch.from_numpy(inst.stack)
    perm = torch.randperm(stack.shape[2])
    times = []
    t_end = time.perf_counter() + seconds
    k = 0
    while time.perf_counter() < t_end or k < 2:
        batch = perm[(k * BATCH) % 400:(k * BATCH) % 400 + BATCH].tolist()
        t0 = time.perf_counter()
        opt.zero_grad()
        loss, _, _ = orc.step_loss(params, table, geom, stack, batch, "smooth_L1_loss", dtype=torch.float32)
        loss.backward()
        opt.step()
        times.append(time.perf_counter() - t0)
        k += 1
    steps_per_s = len(times[1:]) / sum(times[1:]) if len(times) > 1 else 1.0 / times[0]
    return dict(value=steps_per_s / 40.0, unit="iters/s", cores=torch.get_num_threads(), kind="port",
                sample=f"{len(times) - 1} timed steps (+1 warm-up) of the FP32 oracle (pure-torch hash grid, "
                       f"autograd, torch Adam) at the same shapes; steps/s / 40",
                steps_per_s=steps_per_s)


# ----------------------------------------------------------------------------
# "plain" leg: the reference loop (Recovery.py:243-350) as plain PyTorch-ROCm
# FP32 on the GPU, with a pure-torch hash grid in place of tiny-cuda-nn (which
# has no ROCm build) and the reference's own per-position loop and host syncs.
# Used only by --mode nrmse, to time "wall-s to the plain run's final NRMSE".
# ----------------------------------------------------------------------------
def _pos_fma(x, s):
    """f32 x * s + 0.5 with one rounding, like the kernels' fmaf (tcnn's
    pos_fract): the product of two f32 values is exact in f64 and, for
    x * s >= 2^-5, so is the sum; below that the f64 sum rounds at 2^-54,
    which changes the f32 result only where it lands exactly on an f32
    rounding midpoint."""
    return (x.double() * s.double() + 0.5).float()


def _plain_hashgrid(coords, params, spec):
    import torch
    x = coords.float()
    outs = []
    tab = params.view(-1, 2)
    for l in range(spec.n_levels):
        s = torch.tensor(spec.scale[l], dtype=torch.float32, device=x.device)
        pos = _pos_fma(x, s)
        g = torch.floor(pos)
        f = pos - g
        gi = g.to(torch.int64)
        acc = 0.0
        for c in range(4):
            cx = gi[:, 0] + (c & 1)
            cy = gi[:, 1] + ((c >> 1) & 1)
            if spec.dense[l]:
                idx = (cx + cy * spec.res[l]) % spec.size[l]
            else:
                idx = ((cx & 0xFFFFFFFF) ^ ((cy * 2654435761) & 0xFFFFFFFF)) % spec.size[l]
            wx = f[:, 0] if c & 1 else 1 - f[:, 0]
            wy = f[:, 1] if c & 2 else 1 - f[:, 1]
            acc = acc + (wx * wy)[:, None] * tab[spec.offset[l] + idx]
        outs.append(acc)
    return torch.cat(outs, 1)


def _plain_mlp(h, mlp):
    import torch
    import torch.nn.functional as F
    layers = list(mlp.net)
    z = h
    for layer in layers[:-1]:
        z = torch.exp(1j * layer.omega_0 * F.linear(z, layer.linear.weight, layer.linear.bias))
    return F.linear(z, layers[-1].weight, layers[-1].bias)


def _plain_upsample(x, sf):
    import torch
    import torch.nn.functional as F
    H, W = x.shape[-2:]
    X = torch.fft.fftshift(torch.fft.fft2(x, dim=(-2, -1)), dim=(-2, -1))
    nH, nW = int(H * sf), int(W * sf)
    ph, pw = (nH - H) // 2, (nW - W) // 2
    X = F.pad(X, (pw, nW - W - pw, ph, nH - H - ph), mode="constant", value=0)
    return torch.fft.ifft2(torch.fft.ifftshift(X, dim=(-2, -1)), dim=(-2, -1)) * (sf ** 2)


def plain_reconstruct(model, inst, epochs, loss_type="smooth_L1_loss", lr=1e-3, batch=BATCH, seed=0,
                      check_every=None, on_check=None):
    """The reference training loop on the GPU in plain PyTorch (FP32, use_amp false)."""
    import torch
    import torch.nn.functional as F
    dev = model.xy_hash_encoder.params.device
    inp = torch.from_numpy(inst.stack).float().to(dev)
    frames = inp.shape[2]
    tlY, tlX = inst.tlY, inst.tlX
    M = inst.probe.shape[0]
    brY, brX = tlY + M, tlX + M
    opt = torch.optim.Adam(lr=lr, params=filter(lambda p: p.requires_grad, model.parameters()))
    sched = torch.optim.lr_scheduler.StepLR(opt, step_size=10000, gamma=0.1)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    ospec, pspec = model.xy_hash_encoder.spec, model.probe_hash_encoder.spec

    def fields():
        o = _plain_mlp(_plain_hashgrid(model.xy_coords, model.xy_hash_encoder.params, ospec),
                       model.object_complex_mlp).reshape(model.downsampled_width, model.downsampled_height)
        o = _plain_upsample(o[None, None], model.downsample_factor)[0, 0]
        p = _plain_mlp(_plain_hashgrid(model.probe_coords, model.probe_hash_encoder.params, pspec),
                       model.probe_complex_mlp).reshape(model.probe_width, model.probe_height)
        if model.has_probe_residual:
            p = model.k_probe * p + model.probe_initial
        return o, p

    t0 = time.perf_counter()
    for epoch in range(epochs):
        perm = torch.randperm(frames, device=dev, generator=gen)
        for bs in range(0, frames, batch):
            idx = perm[bs:bs + batch]
            opt.zero_grad()
            obj, probe = fields()
            with torch.no_grad():
                absP2 = probe.abs() ** 2
                Mp, Np = absP2.shape
                tot = absP2.sum()
                if tot == 0:
                    cp = None
                else:
                    mn = torch.tensor([Mp, Np], device=dev)
                    cp = torch.trunc(mn / 2 - mn * torch.stack((absP2.sum(1).cumsum(0).mean(),
                                                                absP2.sum(0).cumsum(0).mean())) / tot + 1).long()
            if cp is not None and cp.any():
                probe = torch.roll(probe, (-cp[0].item(), -cp[1].item()), (0, 1))
            energy = (probe.abs() ** 2).sum()
            if energy > 0:
                probe = probe * (float(Mp * Np) / energy).sqrt()
            loss_b = 0.0
            for i in idx:
                ew = probe * obj[tlY[i]:brY[i], tlX[i]:brX[i]]
                I = torch.abs(torch.fft.fftshift(torch.fft.fft2(ew), dim=(-2, -1))) ** 2
                loss_b = loss_b + F.smooth_l1_loss(I, inp[:, :, i])
            loss_b.item()
            loss_b.backward()
            opt.step()
        sched.step()
        if check_every and on_check and ((epoch + 1) % check_every == 0 or epoch == epochs - 1):
            torch.cuda.synchronize()
            with torch.no_grad():
                o, _ = fields()
            on_check(epoch + 1, time.perf_counter() - t0, o)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


# ----------------------------------------------------------------------------
# precision sweep: fused variants (MLP fp32 / bf16 / fp8 x table fp32 / fp16)
# against one FP64 run of the same math (plain PyTorch, batched far field, on
# the GPU), every run on the same seeded batch sequence.
# ----------------------------------------------------------------------------
def _fp64_params(model):
    import torch
    c128 = torch.complex128

    def leaf(t, dtype):
        return t.detach().to(dtype).clone().requires_grad_(True)

    out = {"obj_table": leaf(model.xy_hash_encoder.params, torch.float64),
           "prb_table": leaf(model.probe_hash_encoder.params, torch.float64)}
    for key, mlp in (("obj", model.object_complex_mlp), ("prb", model.probe_complex_mlp)):
        layers = list(mlp.net)
        lay = []
        for i, l in enumerate(layers[:-1]):
            lay.append((leaf(l.linear.weight, torch.float64 if i == 0 else c128),
                        leaf(l.linear.bias, torch.float64 if i == 0 else c128), leaf(l.omega_0, torch.float64)))
        out[key] = {"layers": lay, "wo": leaf(layers[-1].weight, c128), "bo": leaf(layers[-1].bias, c128)}
    out["k_probe"] = leaf(model.k_probe, torch.float64)
    return out


def _fp64_leaves(P):
    leaves = [P["obj_table"], P["prb_table"], P["k_probe"]]
    for key in ("obj", "prb"):
        for (w, b, om) in P[key]["layers"]:
            leaves += [w, b, om]
        leaves += [P[key]["wo"], P[key]["bo"]]
    return leaves


def _fp64_hashgrid(coords, table, spec):
    """Hash-grid features in FP64; f32 positions as in the kernels."""
    import torch
    x = coords.float()
    tab = table.view(-1, 2)
    outs = []
    for l in range(spec.n_levels):
        sc = torch.tensor(spec.scale[l], dtype=torch.float32, device=x.device)
        pos = _pos_fma(x, sc)
        g = torch.floor(pos)
        f = (pos - g).double()
        gi = g.to(torch.int64)
        acc = 0.0
        for c in range(4):
            cx = gi[:, 0] + (c & 1)
            cy = gi[:, 1] + ((c >> 1) & 1)
            if spec.dense[l]:
                idx = (cx + cy * spec.res[l]) % spec.size[l]
            else:
                idx = ((cx & 0xFFFFFFFF) ^ ((cy * 2654435761) & 0xFFFFFFFF)) % spec.size[l]
            wx = f[:, 0] if c & 1 else 1 - f[:, 0]
            wy = f[:, 1] if c & 2 else 1 - f[:, 1]
            acc = acc + (wx * wy)[:, None] * tab[spec.offset[l] + idx]
        outs.append(acc)
    return torch.cat(outs, 1)


def _fp64_mlp(h, M):
    import torch
    import torch.nn.functional as F
    z = h
    for (w, b, om) in M["layers"]:
        z = torch.exp(1j * om * F.linear(z, w, b))
    return F.linear(z, M["wo"], M["bo"])


def _fp64_step_loss(model, P, inp_bmn, tlY, tlX, idx, loss_type):
    import torch
    o = _fp64_mlp(_fp64_hashgrid(model.xy_coords, P["obj_table"], model.xy_hash_encoder.spec), P["obj"])
    o = o.reshape(model.downsampled_width, model.downsampled_height)
    o = _plain_upsample(o[None, None], model.downsample_factor)[0, 0]
    p = _fp64_mlp(_fp64_hashgrid(model.probe_coords, P["prb_table"], model.probe_hash_encoder.spec), P["prb"])
    p = P["k_probe"] * p.reshape(model.probe_width, model.probe_height) + model.probe_initial.to(torch.complex128)
    M, N = p.shape
    with torch.no_grad():   # Recovery.py:262-278 (device-side, no host syncs)
        a2 = p.abs() ** 2
        tot = a2.sum()
        mn = torch.tensor([M, N], device=p.device, dtype=torch.float64)
        cp = torch.trunc(mn / 2 - mn * torch.stack((a2.sum(1).cumsum(0).mean(), a2.sum(0).cumsum(0).mean())) / tot
                         + 1).long()
    p = torch.roll(p, (-int(cp[0]), -int(cp[1])), (0, 1))
    p = p * torch.sqrt(M * N / (p.abs() ** 2).sum())
    wins = torch.stack([o[tlY[i]:tlY[i] + M, tlX[i]:tlX[i] + N] for i in idx])
    I = torch.abs(torch.fft.fftshift(torch.fft.fft2(p * wins), dim=(-2, -1))) ** 2
    meas = inp_bmn[idx]
    eps = 1e-8
    if loss_type == "smooth_L1_loss":
        d = I - meas
        ad = d.abs()
        per = torch.where(ad < 1, 0.5 * d * d, ad - 0.5).mean(dim=(-2, -1))
    elif loss_type == "FD_loss":
        dW = torch.sqrt(meas + eps) - torch.sqrt(I + eps)
        Rx = torch.roll(dW, -1, -1) - dW
        Ry = torch.roll(dW, -1, -2) - dW
        per = torch.sqrt(Rx * Rx + Ry * Ry + eps).mean(dim=(-2, -1))
    else:
        per = (I - meas * torch.log(I + eps)).mean(dim=(-2, -1))
    return per.sum(), o, p


_LOSS_TAGS = {"smooth_L1_loss": "smoothL1", "FD_loss": "FD", "Poisson_likelihood_loss": "Poisson"}


def _model_leaves(model):
    """The model's parameters in the order of _fp64_leaves."""
    out = [model.xy_hash_encoder.params, model.probe_hash_encoder.params, model.k_probe]
    for mlp in (model.object_complex_mlp, model.probe_complex_mlp):
        layers = list(mlp.net)
        for l in layers[:-1]:
            out += [l.linear.weight, l.linear.bias, l.omega_0]
        out += [layers[-1].weight, layers[-1].bias]
    return out


def _leaf_groups(model):
    """Index ranges of _model_leaves grouped as object table, probe table,
    k_probe, object MLP, probe MLP."""
    n_o = 3 * (len(model.object_complex_mlp.net) - 1) + 2
    n_p = 3 * (len(model.probe_complex_mlp.net) - 1) + 2
    return {"object table": [0], "probe table": [1], "k_probe": [2],
            "object MLP": list(range(3, 3 + n_o)), "probe MLP": list(range(3 + n_o, 3 + n_o + n_p))}


def sweep_mode(args):
    """Every precision variant (MLP fp32 / bf16 / fp8 x table fp32 / fp16)
    against an FP64 run of the same math on the same batches, two ways.
    Free-running: each variant trains on its own, and its final loss and
    fields are compared with the FP64 run's (the trajectories diverge
    chaotically, so this mixes precision error with chaos).  Along the FP64
    trajectory: every --sweep-every steps the FP64 run's parameters, batch,
    loss and gradients are recorded; each variant then evaluates its loss
    and gradients at exactly those parameters (no update), which isolates the
    error of one step's arithmetic."""
    import numpy as np
    import torch
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    dev = torch.device("cuda:0")
    shapes, log2T, batch, _, wl_name = WORKLOADS[args.workload]
    loss_type = args.loss or "Poisson_likelihood_loss"
    wl_name = wl_name.rsplit("_", 1)[0] + "_" + _LOSS_TAGS[loss_type]
    inst = pkg.synthetic.make_instance(seed=0, **shapes)
    M = inst.probe.shape[0]
    frames = inst.stack.shape[2]
    mask = pkg.metrics.evaluation_mask(inst.tlY, inst.tlX, M, M, inst.obj.shape)
    rng = np.random.default_rng(1)
    batches = [[perm[b:b + batch] for b in range(0, frames, batch)]
               for perm in (rng.permutation(frames) for _ in range(args.epochs))]

    def make(mlp_p="fp32", tab_p="fp32"):
        torch.manual_seed(0)
        return pkg.ComplexINRModel2D(inst.canvas, inst.canvas, 2, probe_width=M, probe_height=M,
                                     log2_hashmap_size=log2T, hidden_omega_0=1.0, trainable_omega0=True,
                                     use_residual=True, mlp_precision=mlp_p, table_precision=tab_p,
                                     probe_initial=torch.from_numpy(inst.probe.astype(np.complex64))).to(dev)

    def nrmse(a, b, shift):
        return float(pkg.metrics.aligned_nrmse(np.asarray(a), np.asarray(b), mask if a.shape == mask.shape else None,
                                               max_shift=shift))

    # FP64 reference (Adam in FP64, same batches), with trajectory snapshots
    print(f"sweep: FP64 reference, {args.epochs} epochs of {len(batches[0])} steps", file=sys.stderr, flush=True)
    model64 = make()
    groups = _leaf_groups(model64)
    P = _fp64_params(model64)
    leaves = _fp64_leaves(P)
    opt = torch.optim.Adam(leaves, lr=1e-3)
    inp = torch.from_numpy(inst.stack).to(dev).permute(2, 0, 1).contiguous().double()
    t0 = time.perf_counter()
    ref_losses = []
    snaps = []
    k = 0
    for ep, eb in enumerate(batches):
        tot = 0.0
        for idx in eb:
            opt.zero_grad()
            loss, _, _ = _fp64_step_loss(model64, P, inp, inst.tlY, inst.tlX, list(idx), loss_type)
            loss.backward()
            if k % args.sweep_every == 0:
                snaps.append({"step": k, "idx": [int(i) for i in idx], "loss": float(loss.detach()),
                              "params": [t.detach().clone() for t in leaves],
                              "grads": [t.grad.detach().clone() for t in leaves]})
            opt.step()
            tot += float(loss)
            k += 1
        ref_losses.append(tot / len(eb))
        if (ep + 1) % 10 == 0:
            print(f"sweep: FP64 epoch {ep + 1} loss {ref_losses[-1]:.6g} ({time.perf_counter() - t0:.0f} s)",
                  file=sys.stderr, flush=True)
    with torch.no_grad():
        _, o64, p64 = _fp64_step_loss(model64, P, inp, inst.tlY, inst.tlX, [0], loss_type)
    o64, p64 = o64.cpu().numpy(), p64.cpu().numpy()
    rows = [{"variant": "fp64 reference", "final_epoch_loss": ref_losses[-1],
             "object_nrmse_vs_truth": nrmse(o64, inst.obj, 2), "wall_s": time.perf_counter() - t0}]

    def flat_real(ts):
        return torch.cat([(torch.view_as_real(t) if t.is_complex() else t).reshape(-1).double() for t in ts])

    for mlp_p, tab_p in (("fp32", "fp32"), ("bf16", "fp32"), ("fp8", "fp32"),
                         ("fp32", "fp16"), ("bf16", "fp16"), ("fp8", "fp16")):
        model = make(mlp_p, tab_p)
        eng = pkg.FusedReconstruction(model, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX,
                                      loss_type=loss_type, lr=1e-3, batch_size=batch)
        t0 = time.perf_counter()
        losses = []
        for eb in batches:
            acc = torch.zeros((), device=dev)
            for idx in eb:
                acc = acc + eng.step(torch.as_tensor(idx, dtype=torch.int32, device=dev)).sum()
            losses.append(acc)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        o, p = eng.current_fields()
        o, p = o.detach().cpu().numpy(), p.detach().cpu().numpy()
        fl = float(losses[-1]) / len(batches[-1])
        row = {"variant": f"mlp {mlp_p}, table {tab_p}", "final_epoch_loss": fl,
               "loss_rel_vs_fp64": abs(fl - ref_losses[-1]) / abs(ref_losses[-1]),
               "object_nrmse_vs_fp64": nrmse(o, o64, 0), "probe_nrmse_vs_fp64": nrmse(p, p64, 0),
               "object_nrmse_vs_truth": nrmse(o, inst.obj, 2), "wall_s": wall}
        # along the FP64 trajectory: loss and gradients at the recorded parameters
        mleaves = _model_leaves(model)
        traj = {"loss_rel": [], **{f"grad_rel {g}": [] for g in groups}}
        for sn in snaps:
            with torch.no_grad():
                for t, v in zip(mleaves, sn["params"]):
                    t.copy_(v.to(t.dtype))
            bi = torch.as_tensor(sn["idx"], dtype=torch.int32, device=dev)
            eng.step(bi, update=False)          # (FP8: records this point's amax for the scales)
            lv = float(eng.step(bi, update=False).sum())
            traj["loss_rel"].append(abs(lv - sn["loss"]) / abs(sn["loss"]))
            for g, ids in groups.items():
                gv = flat_real([eng.flat.g(mleaves[i]) for i in ids])
                gr = flat_real([sn["grads"][i] for i in ids])
                traj[f"grad_rel {g}"].append(float((gv - gr).norm() / gr.norm().clamp_min(1e-300)))
        row["along_fp64_trajectory"] = {key: {"median": float(np.median(v)), "max": float(np.max(v)),
                                              "values": [float(f"{x:.3g}") for x in v]}
                                        for key, v in traj.items()}
        rows.append(row)
        print(f"sweep: {row}", file=sys.stderr, flush=True)
    fp64_norms = {g: [float(f"{float(flat_real([sn['grads'][i] for i in ids]).norm()):.3g}") for sn in snaps]
                  for g, ids in groups.items()}
    print(json.dumps({"mode": "precision sweep", "workload": wl_name, "loss": loss_type, "epochs": args.epochs,
                      "steps_per_epoch": len(batches[0]), "trajectory_points": len(snaps),
                      "trajectory_every_steps": args.sweep_every,
                      "fp64_trajectory": {"step": [sn["step"] for sn in snaps],
                                          "loss": [float(f"{sn['loss']:.6g}") for sn in snaps],
                                          "grad_norm": fp64_norms},
                      "rows": rows,
                      "note": "object/probe NRMSE against the FP64 run: global complex scale fitted, no shift; "
                              "against the truth: shift <= 2 px, phase and scale fitted, coverage mask; "
                              "along_fp64_trajectory: each variant's batch loss and gradients (relative L2 "
                              "error per parameter group) at the FP64 run's parameters every "
                              "trajectory_every_steps steps"}))


# ----------------------------------------------------------------------------
# learned warm start: train the conv net on simulated instances, then compare
# warm-started and plain (initProbe) reconstructions on held-out seeds by
# epochs and wall-seconds to the plain run's final object NRMSE.
# ----------------------------------------------------------------------------
def warmstart_mode(args):
    import numpy as np
    import torch
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    ws = pkg.warmstart
    dev = torch.device("cuda:0")
    shapes, log2T, batch, loss_type, wl_name = WORKLOADS[args.workload]
    shape = {k: shapes[k] for k in ("canvas", "M", "n_side", "step", "jitter", "photons")}
    M, canvas = shape["M"], shape["canvas"]
    t0 = time.perf_counter()
    seeds = list(range(1000, 1000 + args.ws_train))
    x, y = ws.build_dataset(seeds, device=dev, **shape)
    t_data = time.perf_counter() - t0
    print(f"warmstart: {x.shape[0]} training pairs from {len(seeds)} instances ({t_data:.1f} s)",
          file=sys.stderr, flush=True)
    net = ws.HipWarmStartNet(base=32, device=dev, seed=0)
    t0 = time.perf_counter()
    hist = ws.train(net, x, y, epochs=args.ws_epochs,
                    log=lambda e, l: print(f"warmstart: net epoch {e} mse {l:.5g}", file=sys.stderr, flush=True))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    n_pairs = x.shape[0]
    del x, y
    rows = []
    for k in range(args.ws_heldout):
        inst = ws.simulate(900000 + k, jitter_probe=False, device=dev, **shape)
        truth = inst.obj.cpu().numpy()
        ty, tx = inst.tlY.cpu().numpy(), inst.tlX.cpu().numpy()
        mask = pkg.metrics.evaluation_mask(ty, tx, M, M, truth.shape)
        stack = inst.intens.permute(1, 2, 0).contiguous()          # reference layout [M, N, frames]

        def err(o):
            return pkg.metrics.aligned_nrmse(o.detach().cpu().numpy(), truth, mask, max_shift=2)

        def run(object_initial, target=None):
            torch.manual_seed(0)
            model = pkg.ComplexINRModel2D(canvas, canvas, 2, probe_width=M, probe_height=M, log2_hashmap_size=log2T,
                                          hidden_omega_0=1.0, trainable_omega0=True, use_residual=True,
                                          object_initial=object_initial,
                                          probe_initial=inst.probe.to(torch.complex64)).to(dev)
            eng = pkg.FusedReconstruction(model, stack, ty, tx, loss_type=loss_type, lr=1e-3, batch_size=batch)
            gen = torch.Generator(device=dev)
            gen.manual_seed(0)
            curve, hit = [], None
            t_start = time.perf_counter()
            for ep in range(args.epochs):
                eng.run_epoch(gen)
                if (ep + 1) % args.check_every == 0 or ep == args.epochs - 1:
                    torch.cuda.synchronize()
                    e = err(eng.current_fields()[0])
                    t = time.perf_counter() - t_start
                    curve.append((ep + 1, t, e))
                    if target is not None and hit is None and e <= target:
                        hit = (ep + 1, t)
            return curve, hit

        curve_plain, _ = run(None)
        target = curve_plain[-1][2] * 1.01
        t0 = time.perf_counter()
        init = ws.warm_start_object(net, inst.intens, ty, tx, M, canvas, probe=inst.probe)
        torch.cuda.synchronize()
        t_net = time.perf_counter() - t0
        curve_warm, hit = run(init, target)
        # control: the residual form with a constant initial canvas (the warm
        # canvas's mean over the covered pixels), separating what the net's
        # patches contribute from what the residual parametrisation does
        const = torch.full_like(init, complex(init[torch.from_numpy(mask).to(dev)].mean().item()))
        curve_const, hit_const = run(const, target)
        row = {"seed": 900000 + k, "plain_final_nrmse": curve_plain[-1][2], "plain_wall_s": curve_plain[-1][1],
               "plain_epochs": args.epochs, "warm_initial_nrmse": err(init), "warm_net_s": t_net,
               "warm_final_nrmse": curve_warm[-1][2],
               "const_final_nrmse": curve_const[-1][2],
               "const_epochs_to_plain_final": hit_const[0] if hit_const else None,
               "warm_epochs_to_plain_final": hit[0] if hit else None,
               "warm_wall_s_to_plain_final": (hit[1] + t_net) if hit else None,
               "plain_curve": curve_plain, "warm_curve": curve_warm, "const_curve": curve_const}
        rows.append(row)
        print(f"warmstart: seed {row['seed']} plain {row['plain_final_nrmse']:.4f} in {args.epochs} epochs; warm "
              f"init {row['warm_initial_nrmse']:.4f} final {row['warm_final_nrmse']:.4f}, reaches it at epoch "
              f"{row['warm_epochs_to_plain_final']}; constant canvas final {row['const_final_nrmse']:.4f}",
              file=sys.stderr, flush=True)
    hits = [r for r in rows if r["warm_epochs_to_plain_final"]]
    print(json.dumps({
        "mode": "learned warm start", "workload": wl_name, "train_instances": len(seeds),
        "train_pairs": n_pairs, "net": "HIP conv encoder-decoder (pty_wsnet_*), base 32, BF16 MFMA, batch 256",
        "net_epochs": args.ws_epochs, "net_mse": hist, "data_s": t_data, "train_s": t_train,
        "net_train_samples_per_s": n_pairs * args.ws_epochs / t_train,
        "reconstruction_epochs": args.epochs, "heldout": len(rows),
        "mean_plain_final_nrmse": float(np.mean([r["plain_final_nrmse"] for r in rows])),
        "mean_warm_initial_nrmse": float(np.mean([r["warm_initial_nrmse"] for r in rows])),
        "mean_warm_final_nrmse": float(np.mean([r["warm_final_nrmse"] for r in rows])),
        "mean_const_final_nrmse": float(np.mean([r["const_final_nrmse"] for r in rows])),
        "heldout_reaching_target": len(hits),
        "mean_epoch_ratio_warm_over_plain": (float(np.mean([r["warm_epochs_to_plain_final"] / args.epochs
                                                            for r in hits])) if hits else None),
        "mean_wall_ratio_warm_over_plain": (float(np.mean([r["warm_wall_s_to_plain_final"] / r["plain_wall_s"]
                                                           for r in hits])) if hits else None),
        "rows": rows}))


def nrmse_compare(workload, epochs, check_every):
    """Wall-clock to the plain run's final object NRMSE: the plain PyTorch
    restatement (the reference's loop) for `epochs` epochs, then the fused HIP
    engine from the same initial model until it reaches that NRMSE (within
    1 %); returns the record."""
    import numpy as np
    import torch
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    dev = torch.device("cuda:0")
    shapes = dict(pkg.synthetic.SHAPES[workload])
    inst = pkg.synthetic.make_instance(seed=0, **shapes)
    M = inst.probe.shape[0]
    log2T = 14 if workload == "cfg0" else LOG2T
    mask = pkg.metrics.evaluation_mask(inst.tlY, inst.tlX, M, M, inst.obj.shape)

    def make():
        torch.manual_seed(0)
        return pkg.ComplexINRModel2D(inst.canvas, inst.canvas, 2, probe_width=M, probe_height=M,
                                     log2_hashmap_size=log2T, hidden_omega_0=1.0, trainable_omega0=True,
                                     use_residual=True,
                                     probe_initial=torch.from_numpy(inst.probe.astype(np.complex64))).to(dev)

    def err(o):
        return pkg.metrics.aligned_nrmse(o.detach().cpu().numpy(), inst.obj, mask, max_shift=2)

    curve_plain = []

    def plain_check(e, t, o):
        curve_plain.append((e, t, err(o)))
        print(f"plain epoch {e}: {t:.1f} s, NRMSE {curve_plain[-1][2]:.4f}", file=sys.stderr, flush=True)

    wall_plain = plain_reconstruct(make(), inst, epochs, check_every=check_every, on_check=plain_check)
    final_plain = curve_plain[-1][2]
    target = final_plain * 1.01
    model = make()
    eng = pkg.FusedReconstruction(model, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX, batch_size=BATCH)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    curve_fused = []
    t0 = time.perf_counter()
    hit = None
    for ep in range(epochs):
        eng.run_epoch(gen)
        if (ep + 1) % check_every == 0 or ep == epochs - 1:
            torch.cuda.synchronize()
            o, _ = eng.current_fields()
            e = err(o)
            t = time.perf_counter() - t0
            curve_fused.append((ep + 1, t, e))
            print(f"fused epoch {ep + 1}: {t:.2f} s, NRMSE {e:.4f}", file=sys.stderr, flush=True)
            if hit is None and e <= target:
                hit = (ep + 1, t)
    wall_fused_total = time.perf_counter() - t0
    return {
        "metric": "wall-s to plain-final object NRMSE + iters/s: HIP fused vs plain, 1 MI355X",
        "workload": workload, "epochs": epochs,
        "plain": {"wall_s": wall_plain, "final_nrmse": final_plain, "iters_per_s": epochs / wall_plain,
                  "curve": curve_plain},
        "fused": {"wall_s_total": wall_fused_total, "final_nrmse": curve_fused[-1][2],
                  "iters_per_s": epochs / wall_fused_total, "curve": curve_fused,
                  "epochs_to_target": hit[0] if hit else None, "wall_s_to_target": hit[1] if hit else None},
        "target_nrmse": target,
        "speedup_to_target": (wall_plain / hit[1]) if hit else None,
        "note": "NRMSE checks (host transfer + alignment) are inside both wall clocks",
    }


def nrmse_mode(args):
    print(json.dumps(nrmse_compare(args.workload, args.epochs, args.check_every)))



# ----------------------------------------------------------------------------
# kernel mode: the gather and FFT kernels on their own, algorithmic GB/s
# ----------------------------------------------------------------------------
HBM_PEAK_GBS = 8000.0


def _time_launches(fn, reps=20, warmup=3):
    """Mean milliseconds of fn() between HIP events on the current stream."""
    import torch
    for _ in range(warmup):
        fn()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / reps


def kernels_mode(args):
    """Standalone kernels at the canvas-3648 shapes: the hash-grid encoder
    (pty_hashgrid_fwd, gathers of 4 corners x 16 levels per point), the
    Fourier upsample 1824 -> 3648 and its adjoint (pty_upsample_fwd/adj), and
    the far-field fused forward + adjoint for a batch of 64 probe-512
    patterns; then the object table's gradient + Adam stage of the fused step
    at canvas 1024 and 3648, and a plain device copy for scale.  Algorithmic bytes are the minimum the operation must move
    (inputs read once, outputs written once); achieved = bytes / mean launch
    time, against the 8 TB/s HBM peak (table gathers mostly hit L2 / MALL, so
    the gather kernel's figure can exceed it)."""
    import numpy as np
    import torch
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    _lib = pkg._lib
    dev = torch.device("cuda:0")
    rows = []

    def row(name, shape, nbytes, ms, extra=None):
        gbs = nbytes / (ms * 1e-3) / 1e9
        r = {"kernel": name, "shape": shape, "algorithmic_bytes": nbytes, "ms": ms, "GB_per_s": gbs,
             "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
        if extra:
            r.update(extra)
        rows.append(r)
        print(f"{name:34s} {shape:40s} {ms * 1e3:9.1f} us {gbs:9.0f} GB/s", file=sys.stderr, flush=True)

    # --- hash-grid encoder: n = 1824^2 points, log2T 20, fp32 and fp16 tables
    n = 1824 * 1824
    spec = pkg.level_table(16, 2, 20, 16, 1.5)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    # the INR's own points: the cell-centred 1824 x 1824 grid in raster order
    # (init_scale_grids), and the same number of uniform random points
    ax = (torch.arange(1824, device=dev, dtype=torch.float32) + 0.5) / 1824
    grid = torch.stack([ax.repeat(1824), ax.repeat_interleave(1824)], 1).contiguous()
    rnd = torch.rand(n, 2, device=dev, generator=g)
    params = (torch.rand(spec.n_entries * 2, device=dev, generator=g) - 0.5) * 2e-4
    out = torch.empty(n, 32, device=dev)
    s = _lib.stream_ptr(dev)
    for cname, coords in (("grid", grid), ("random", rnd)):
        for tp in ("fp32", "fp16"):
            desc = spec.descriptor(tp)
            tab = params if tp == "fp32" else params.to(torch.float16)
            ms = _time_launches(lambda: _lib.call("pty_hashgrid_fwd", desc, coords.data_ptr(), n, tab.data_ptr(),
                                                  out.data_ptr(), s))
            esz = 8 if tp == "fp32" else 4
            nbytes = n * (8 + 16 * 4 * esz + 32 * 4)
            row(f"k_hashgrid_fwd ({tp} table, {cname} points)", f"{n} points, 16 levels, log2T 20", nbytes, ms,
                {"bytes_per_point": nbytes / n})
    del grid, rnd, params, out

    # --- Fourier upsample and adjoint: 1824^2 -> 3648^2 complex64
    nsm, ds = 1824, 2
    m = nsm * ds
    small = torch.randn(nsm, nsm, dtype=torch.complex64, device=dev)
    big = torch.empty(m, m, dtype=torch.complex64, device=dev)
    ws = _lib.workspace(_lib.query("pty_upsample_workspace", nsm, ds), dev)
    nbytes = 8 * (nsm * nsm + m * m)
    ms = _time_launches(lambda: _lib.call("pty_upsample_fwd", small.data_ptr(), nsm, ds, big.data_ptr(),
                                          ws.data_ptr(), ws.numel(), s))
    row("upsample fwd (rows, T, rows, T)", f"{nsm}^2 -> {m}^2 c64", nbytes, ms)
    ms = _time_launches(lambda: _lib.call("pty_upsample_adj", big.data_ptr(), nsm, ds, small.data_ptr(),
                                          ws.data_ptr(), ws.numel(), s))
    row("upsample adjoint", f"{m}^2 -> {nsm}^2 c64", nbytes, ms)
    del small, big, ws

    # --- far field fused forward + adjoint: 64 patterns of 512^2, canvas 3648
    M, H, B, F = 512, 3648, 64, 128
    obj = torch.randn(H, H, dtype=torch.complex64, device=dev)
    prb = torch.randn(M, M, dtype=torch.complex64, device=dev)
    stack = torch.rand(F, M, M, device=dev) * 10
    tl = torch.randint(0, H - M, (F, 2), device=dev, dtype=torch.int32)
    bidx = torch.arange(B, device=dev, dtype=torch.int32)
    loss = torch.empty(B, device=dev)
    dobj = torch.empty_like(obj)
    dprb = torch.empty_like(prb)
    for lt in ("smooth_L1_loss", "FD_loss"):
        d = pkg.farfield.far_descriptor(M, M, H, H, B, F, lt)
        ws = _lib.workspace(_lib.query("pty_farfield_workspace", d), dev)
        ms = _time_launches(lambda: _lib.call("pty_farfield_fwd_bwd", d, obj.data_ptr(), prb.data_ptr(),
                                              stack.data_ptr(), bidx.data_ptr(), tl.data_ptr(), loss.data_ptr(),
                                              dobj.data_ptr(), dprb.data_ptr(), ws.data_ptr(), ws.numel(), s))
        # per pattern pixel: window read 8 B + measurement 4 B; canvas gradient
        # written once (8 B / canvas pixel); probe read + gradient 16 B / probe pixel
        nbytes = B * M * M * 12 + H * H * 8 + M * M * 16
        row(f"far field fwd+adjoint ({lt})", f"{B} x {M}^2 patterns, canvas {H}", nbytes, ms)
    del obj, prb, stack, dobj, dprb, ws

    # --- the table-gradient tail of the object INR's backward with the tables'
    # Adam step (pty_inr_bwd_finish_adam: gradient finish, dense levels in two
    # separable passes, hashed levels by CSR gather), in the fused step of both
    # workloads (eager, one stream, events around the launch)
    for wl in ("cfg1", "cfg3"):
        shapes_w, _, batch_w, loss_w, wl_name_w = WORKLOADS[wl]
        inst, model = build_problem(pkg, dev, workload=wl)
        eng = pkg.FusedReconstruction(model, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX,
                                      loss_type=loss_w, lr=1e-3, batch_size=batch_w)
        eng.overlap = False
        bi = torch.arange(batch_w, dtype=torch.int32, device=dev)
        for _ in range(3):
            eng.step(bi)
        torch.cuda.synchronize()
        eng.enable_timing(["obj_finish_adam"])
        for _ in range(10):
            eng.step(bi)
        torch.cuda.synchronize()
        t = sorted(eng.timings_ms()["obj_finish_adam"])
        ms = t[len(t) // 2]
        spec_w = model.xy_hash_encoder.spec
        n_pts = model.downsampled_width * model.downsampled_height
        n_hashed = sum(1 for d in spec_w.dense if not d)
        n_dense_entries = sum(sz for sz, d in zip(spec_w.size, spec_w.dense) if d)
        n_hashed_entries = spec_w.n_entries - n_dense_entries
        # feature gradients read once (8 B per point and level), the hashed
        # levels' CSR terms (4 per point and level, 4 B) and offsets (4 B per
        # entry), the dense levels' row sums A written and read once, Adam's
        # parameter and moments read and written (48 B per entry), the MLP
        # partials read once
        a_bytes = 8 * sum(model.downsampled_height * (r + 1) for r, d in zip(spec_w.res, spec_w.dense) if d)
        nbytes = (8 * spec_w.n_levels * n_pts + 16 * n_hashed * n_pts + 4 * n_hashed_entries + 2 * a_bytes
                  + 48 * spec_w.n_entries)
        row(f"table gradient + Adam, object ({wl_name_w.split('_')[0]})",
            f"{n_pts} points, {spec_w.n_entries} entries", nbytes, ms,
            {"dense_entries": n_dense_entries, "hashed_entries": n_hashed_entries})
        del eng, model, inst
        torch.cuda.empty_cache()

    # --- what a plain device copy achieves (read + write), for scale
    src = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    ms = _time_launches(lambda: dst.copy_(src))
    row("device copy (torch copy_, 1 GiB)", "read 1 GiB + write 1 GiB", 2 * src.numel() * 4, ms)
    del src, dst
    out = {"mode": "kernels", "device": torch.cuda.get_device_name(0), "hbm_peak_GBs": HBM_PEAK_GBS,
           "rows": rows}
    print(json.dumps(out))


def counters_mode(args):
    """The hot kernels once each (after one warm-up round) at the canvas-3648
    shapes, on inputs made on the GPU, for rocprofv3 --pmc passes: the fused
    INR forward (hash gather + MLP) and the MLP backward in FP32 and in BF16
    on the 1824^2 grid (log2T 20), the table-gradient + Adam stage, the
    Fourier upsample 1824 -> 3648 and its adjoint, the far field (FD, 64
    patterns of 512^2), the standalone encoder, and a 256 MiB device copy
    that calibrates the HBM byte counters.  --counters-groups picks a subset;
    every group synchronises and reports its time, so a slow pass shows where
    it is."""
    import torch
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    L = pkg._lib
    dev = torch.device("cuda:0")
    s = L.stream_ptr(dev)
    q = L.query
    g = torch.Generator(device=dev)
    g.manual_seed(0)

    def say(msg):
        print(f"counters: {msg}", file=sys.stderr, flush=True)

    def inr_group(prec):
        torch.manual_seed(1)
        model = pkg.ComplexINRModel2D(3648, 3648, 2, probe_width=512, probe_height=512, log2_hashmap_size=20,
                                      hidden_omega_0=1.0, trainable_omega0=True, mlp_precision=prec).to(dev)
        enc, mlp = model.xy_hash_encoder, model.object_complex_mlp
        with torch.no_grad():
            enc.params.copy_((torch.rand(enc.params.shape, device=dev, generator=g) - 0.5) * 0.2)
        spec = enc.spec
        grid = spec.descriptor("fp32")
        pts = model.grid_points("object")
        pd = pts.descriptor(spec)
        feat = torch.empty(q("pty_points_feature_floats", pts.n), device=dev)
        pd.features = feat.data_ptr()
        ts = [t.detach().contiguous() for t in mlp.tensors()]
        md = pkg.complex_inr.mlp_descriptor(ts, 32, prec)
        grads = [torch.zeros_like(t) for t in ts]
        gd = pkg.complex_inr.mlp_descriptor(grads, 32, prec)
        out = torch.empty(pts.n, dtype=torch.complex64, device=dev)
        dout = torch.randn(pts.n, dtype=torch.complex64, device=dev, generator=g)
        ws = L.workspace(q("pty_inr_bwd_workspace", grid, md, pd), dev)
        m1 = torch.zeros_like(enc.params)
        v1 = torch.zeros_like(enc.params)
        hyper = torch.tensor([1e-3, 1.0], device=dev)
        tab = enc.params.detach()

        def run():
            L.call("pty_inr_fwd", grid, md, tab.data_ptr(), pd, out.data_ptr(), s)
            L.call("pty_inr_bwd_mlp", grid, md, tab.data_ptr(), pd, dout.data_ptr(), ws.data_ptr(), ws.numel(), s)
            L.call("pty_inr_bwd_finish_adam", grid, md, gd, tab.data_ptr(), m1.data_ptr(), v1.data_ptr(),
                   hyper.data_ptr(), 0.9, 0.999, 1e-8, pd, ws.data_ptr(), ws.numel(), s)
        # the descriptors hold raw device pointers: keep their owners (the
        # model's weights, the grid's axes and plan, the feature cache and
        # the gradient buffers) alive as long as the closure
        run.owners = (model, pts, feat, ts, grads)
        return run

    def upsample_group():
        nsm, ds = 1824, 2
        small = torch.randn(nsm, nsm, dtype=torch.complex64, device=dev, generator=g)
        big = torch.empty(nsm * ds, nsm * ds, dtype=torch.complex64, device=dev)
        wsu = L.workspace(q("pty_upsample_workspace", nsm, ds), dev)

        def run():
            L.call("pty_upsample_fwd", small.data_ptr(), nsm, ds, big.data_ptr(), wsu.data_ptr(), wsu.numel(), s)
            L.call("pty_upsample_adj", big.data_ptr(), nsm, ds, small.data_ptr(), wsu.data_ptr(), wsu.numel(), s)
        return run

    def farfield_group():
        M, H, B, F = 512, 3648, 64, 128
        obj = torch.randn(H, H, dtype=torch.complex64, device=dev, generator=g)
        prb = torch.randn(M, M, dtype=torch.complex64, device=dev, generator=g)
        stack = torch.rand(F, M, M, device=dev, generator=g) * 10
        tl = torch.randint(0, H - M, (F, 2), device=dev, dtype=torch.int32, generator=g)
        bidx = torch.arange(B, device=dev, dtype=torch.int32)
        loss = torch.empty(B, device=dev)
        dobj = torch.empty_like(obj)
        dprb = torch.empty_like(prb)
        fd = pkg.farfield.far_descriptor(M, M, H, H, B, F, "FD_loss")
        wsf = L.workspace(q("pty_farfield_workspace", fd), dev)

        def run():
            L.call("pty_farfield_fwd_bwd", fd, obj.data_ptr(), prb.data_ptr(), stack.data_ptr(), bidx.data_ptr(),
                   tl.data_ptr(), loss.data_ptr(), dobj.data_ptr(), dprb.data_ptr(), wsf.data_ptr(), wsf.numel(), s)
        return run

    def hashgrid_group():
        n = 1824 * 1824
        spec = pkg.level_table(16, 2, 20, 16, 1.5)
        ax = (torch.arange(1824, device=dev, dtype=torch.float32) + 0.5) / 1824
        coords = torch.stack([ax.repeat(1824), ax.repeat_interleave(1824)], 1).contiguous()
        params = (torch.rand(spec.n_entries * 2, device=dev, generator=g) - 0.5) * 2e-4
        feats = torch.empty(n, 32, device=dev)
        desc = spec.descriptor("fp32")

        def run():
            L.call("pty_hashgrid_fwd", desc, coords.data_ptr(), n, params.data_ptr(), feats.data_ptr(), s)
        return run

    makers = {"inr_fp32": lambda: inr_group("fp32"), "inr_bf16": lambda: inr_group("bf16"),
              "upsample": upsample_group, "farfield": farfield_group, "hashgrid": hashgrid_group}
    names = [n for n in (args.counters_groups or ",".join(makers)).split(",") if n]
    bad = [n for n in names if n not in makers]
    if bad:
        raise SystemExit(f"unknown --counters-groups {bad}; choose from {sorted(makers)}")
    groups = []
    for n in names:
        groups.append((n, makers[n]()))
        torch.cuda.synchronize()
        say(f"{n} set up")
    for it in range(2):                 # warm-up round, then the profiled one
        for n, run in groups:
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            say(f"round {it} {n} {time.perf_counter() - t0:.3f} s")
    copy_calibration(dev, mib=256, reps=1)
    torch.cuda.synchronize()
    say("copy calibration done")
    print(json.dumps({"mode": "counters", "groups": names, "rounds": 2}))


# derived figures of the counter table.  Normalisation (MI355X: 8 XCDs, 256
# CUs, 1024 SIMDs): GRBM_GUI_ACTIVE is summed over the 8 XCDs, so a kernel's
# duration in cycles is GRBM_GUI_ACTIVE / 8 (checked against the kernel trace:
# 2.2-2.4 GHz for every kernel); SQ_VALU_MFMA_BUSY_CYCLES is per SIMD, summed
# over the 1024 SIMDs; SQ_WAVE_CYCLES, SQ_ACTIVE_INST_VALU and SQ_WAIT_INST_*
# are wave quad-cycles summed over waves; SQ_LDS_BANK_CONFLICT is stall cycles
# summed over the 256 CUs; SQ_INSTS_VALU counts MFMA instructions too.
GS_WORKLOADS = {
    # name: (n, canvas, parameterisation, max_patch_radius, min_scale)
    "gs50k_cholesky": (50000, 1024, "cholesky", 16, 1.0),
    "gs50k_rs": (50000, 1024, "rs", 8, 0.5),
}
GS_FLOP_FWD = 15     # per (Gaussian, pixel) evaluation: offsets, quadratic form, exp2, two accumulations
GS_FLOP_BWD = 30     # the same plus the seven gradient sums


def gscp_mode(args):
    """Gaussian-field rasterizer (pty_gs_*, the reference's gscp CUDA
    rasterizer) forward + backward through GaussianRasterize2D[Cholesky] on
    the synthetic scene of the reference's simulate_and_reconstruct.py:170
    (50,000 object Gaussians, canvas 1024^2 at downsample 1); Cholesky with
    max_patch_radius 16 / min_scale 1 (config/simulation_conventional_
    westwest_ng50000.yaml) and RS with the model defaults 8 / 0.5.  A step is
    one forward (including its count read-back, as the reference's) and one
    backward.  Throughput is reported as Gaussian-pixel evaluations per
    second; the CPU baseline is the oracle (numpy) on a sample of tiles."""
    import numpy as np
    import torch
    import __graft_entry__
    from oracle import gscp_oracle as go
    pkg = __graft_entry__.load_package()
    dev = torch.device("cuda:0")
    name = args.workload or "gs50k_cholesky"
    n, canvas, param, smax, smin = GS_WORKLOADS[name]
    rng = np.random.default_rng(0)
    xy = rng.uniform(0, canvas - 1, size=(n, 2))
    raw = rng.normal(0.6, 0.45, size=(n, 2))
    sig = smin + (smax - smin) / (1 + np.exp(-raw)) * 0.5
    p0 = np.log(sig)
    p1 = rng.normal(0, 0.1, size=(n, 1)) if param == "rs" else rng.normal(0, 1.0, size=(n,)) * sig[:, 0] * 0.5
    w = rng.normal(size=(n, 2))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    xy, p0, p1, w = map(f32, (xy, p0, p1, w))
    T = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in (xy, p0, p1, w)]
    fn = pkg.gscp.GaussianRasterize2D if param == "rs" else pkg.gscp.GaussianRasterize2DCholesky
    dout = torch.randn(2, canvas, canvas, device=dev, generator=torch.Generator(device=dev).manual_seed(1))
    fwd_ms, bwd_ms = [], []

    def step(record):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        out = fn.apply(*T, canvas, canvas, smax, smin)
        e[1].record()
        out.backward(dout)
        e[2].record()
        if record:
            fwd_ms.append(e)
        return out

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    f_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in fwd_ms]))
    b_ms = float(np.mean([b.elapsed_time(c) for _, b, c in fwd_ms]))
    R = pkg.gscp.rasterize_forward(*[t.detach() for t in T], canvas, canvas, smax, smin)[0] if param == "rs" else \
        pkg.gscp.rasterize_forward_cholesky(*[t.detach() for t in T], canvas, canvas, smax, smin)[0]
    evals = R * 256
    ms_step = wall * 1e3 / args.steps
    # CPU baseline: the oracle's render + backward on a sample of 32 tiles, scaled to every tile
    pre = go.preprocess_rs if param == "rs" else go.preprocess_cholesky
    t1 = time.perf_counter()
    geom = pre(xy, p0, p1, canvas, canvas, smax, smin)
    lists = go.tile_lists(geom)
    t_pre = time.perf_counter() - t1
    tiles = rng.choice(len(lists), size=32, replace=False)
    keep = set(tiles.tolist())
    sel = [l if t in keep else [] for t, l in enumerate(lists)]
    d64 = dout.cpu().numpy().astype(np.float64)
    t1 = time.perf_counter()
    go.render(geom, w, sel)
    go.render_backward(geom, w, d64, sel)
    t_tiles = time.perf_counter() - t1
    sel_pairs = sum(len(lists[t]) for t in tiles)
    cpu_s = t_pre + t_tiles * R / max(sel_pairs, 1)
    print(json.dumps({
        "mode": "gaussian-field rasterizer", "workload": name, "n_gaussians": n, "canvas": canvas,
        "param": param, "max_patch_radius": smax, "min_scale": smin, "steps": args.steps, "warmup": args.warmup,
        "pairs": R, "tiles_per_gaussian": R / n, "ms_per_step": ms_step, "fwd_ms": f_ms, "bwd_ms": b_ms,
        "evaluations_per_s": 2 * evals / (ms_step * 1e-3),
        "fwd_tflops": evals * GS_FLOP_FWD / (f_ms * 1e-3) / 1e12,
        "bwd_tflops": evals * GS_FLOP_BWD / (b_ms * 1e-3) / 1e12,
        "flop_per_evaluation": {"fwd": GS_FLOP_FWD, "bwd": GS_FLOP_BWD}, "peak_fp32_tflops": 157.3,
        "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "steps/s", "cores": 1, "kind": "port",
                         "sample": f"oracle (numpy FP64) geometry + tile lists of all {n} Gaussians, render + "
                                   f"backward of 32 tiles ({sel_pairs} pairs) scaled to {R} pairs"},
    }), flush=True)


def _cycles(d):
    return d["GRBM_GUI_ACTIVE"] / 8


_PMC_DERIVED = [
    ("waves per SIMD", lambda d: d["SQ_WAVE_CYCLES"] * 4 / (1024 * _cycles(d))),
    ("VALU insts / wave (no MFMA)", lambda d: (d["SQ_INSTS_VALU"] - d["SQ_INSTS_MFMA"]) / d["SQ_WAVES"]),
    ("MFMA insts / wave", lambda d: d["SQ_INSTS_MFMA"] / d["SQ_WAVES"]),
    ("transcendental insts / wave", lambda d: d["SQ_INSTS_VALU_TRANS_F32"] / d["SQ_WAVES"]),
    ("LDS insts / wave", lambda d: d["SQ_INSTS_LDS"] / d["SQ_WAVES"]),
    ("MFMA busy (% of SIMD cycles)", lambda d: 100 * d["SQ_VALU_MFMA_BUSY_CYCLES"] / (1024 * _cycles(d))),
    ("VALU active (% of SIMD cycles)", lambda d: 100 * d["SQ_ACTIVE_INST_VALU"] * 4 / (1024 * _cycles(d))),
    ("LDS bank-conflict stall (% of CU cycles)", lambda d: 100 * d["SQ_LDS_BANK_CONFLICT"] / (256 * _cycles(d))),
    ("LDS-issue wait (% of wave cycles)", lambda d: 100 * d["SQ_WAIT_INST_LDS"] / d["SQ_WAVE_CYCLES"]),
    ("any-issue wait (% of wave cycles)", lambda d: 100 * d["SQ_WAIT_INST_ANY"] / d["SQ_WAVE_CYCLES"]),
]


def pmc_table(out_stem, csvs, command):
    """Per-kernel mean of every counter over its dispatches in rocprofv3
    --pmc counter_collection CSVs (one CSV per pass), with the derived figures
    of _PMC_DERIVED; HBM bytes from FETCH_SIZE / WRITE_SIZE (KiB) scaled by
    what they read for the largest device-copy dispatch (256 MiB read + 256
    MiB written).  Writes <out_stem>.md and .json."""
    import csv
    import re
    from collections import defaultdict
    vals = defaultdict(lambda: defaultdict(list))
    for path in csvs:
        for r in csv.DictReader(open(path)):
            name = re.sub(r"\(.*", "", r["Kernel_Name"]).replace("void ", "").strip()
            vals[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    mean = {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in vals.items()}
    true_kib = 256 * 1024
    factors = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = [v for k, d in vals.items() if "copyBuffer" in k for v in d.get(c, []) if v > true_kib / 4]
        if cal:
            factors[c] = true_kib / max(cal)
    order = ["k_inr_fwd16", "k_inr_bwd16", "k_inr_fwd_b16", "k_inr_bwd_b16", "k_table_hashed", "k_table_dense",
             "k_dense_pass1", "k_inr_grad_finish", "k_up2_rows", "k_transpose", "k_far_rows_fwd", "k_far_cols",
             "k_far_fd", "k_far_rows_bwd", "k_far_dgrad", "k_hashgrid_fwd", "k_gs_"]
    names = [k for o in order for k in sorted(mean) if k.split("::")[-1].startswith(o)]
    counters = sorted({c for d in mean.values() for c in d})
    lines = [f"# Hardware counters per kernel ({', '.join(os.path.basename(os.path.dirname(c)) or c for c in csvs)})",
             "", "Per-dispatch means over the profiled dispatches, one rocprofv3 --pmc pass per counter group, "
             "run from the repository root on one MI355X:", "", "```", command, "```", ""]
    if factors:
        lines += ["HBM bytes: " + ", ".join(f"{c} reports {1 / f:.4f}x a 256 MiB device copy's true bytes, so "
                                              f"its readings are multiplied by {f:.4f}" for c, f in factors.items()),
                  ""]
    lines += ["Derived figures (MI355X: 8 XCDs, 256 CUs, 1024 SIMDs): a kernel's duration in cycles is "
              "GRBM_GUI_ACTIVE / 8 (the counter is summed over the XCDs; against the kernel trace this gives "
              "2.2-2.4 GHz for every kernel); MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / (1024 x cycles); VALU "
              "active = 4 SQ_ACTIVE_INST_VALU (quad-cycles) / (1024 x cycles); bank-conflict stall = "
              "SQ_LDS_BANK_CONFLICT / (256 x cycles); waits are SQ_WAIT_INST_* / SQ_WAVE_CYCLES; waves per "
              "SIMD = 4 SQ_WAVE_CYCLES / (1024 x cycles); VALU instructions exclude MFMA "
              "(SQ_INSTS_VALU - SQ_INSTS_MFMA). Read / write MB are the bytes L2 fetched from / wrote to the "
              "memory fabric (MALL hits count as fetches).", ""]
    hdr = ["kernel", "cycles"] + [d[0] for d in _PMC_DERIVED] + ["read MB", "write MB"]
    lines += ["| " + " | ".join(hdr) + " |", "|" + "---|" * len(hdr)]
    out = {"command": command, "factors": factors, "kernels": {}}
    for k in names:
        d = mean[k]
        row = [k.split("::")[-1], f"{_cycles(d):.3g}" if "GRBM_GUI_ACTIVE" in d else ""]
        derived = {}
        for nm, fn in _PMC_DERIVED:
            try:
                v = fn(d)
            except (KeyError, ZeroDivisionError):
                v = None
            derived[nm] = v
            row.append(f"{v:.3g}" if v is not None else "")
        hb = {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            v = d[c] * 1024 * factors[c] / 1e6 if c in d and c in factors else None
            hb[c] = v
            row.append(f"{v:.1f}" if v is not None else "")
        lines.append("| " + " | ".join(row) + " |")
        out["kernels"][k] = {"counters": d, "derived": derived, "hbm_MB": hb}
    lines += ["", "Raw per-dispatch means:", "", "| kernel | " + " | ".join(counters) + " |",
              "|" + "---|" * (len(counters) + 1)]
    for k in names:
        lines.append("| " + k.split("::")[-1] + " | " +
                     " | ".join(f"{mean[k][c]:.4g}" if c in mean[k] else "" for c in counters) + " |")
    Path(out_stem + ".md").write_text("\n".join(lines) + "\n")
    Path(out_stem + ".json").write_text(json.dumps(out, indent=1) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nrmse-epochs", type=int, default=100,
                    help="epochs of the default line's wall-to-NRMSE leg on the canvas-256 shapes (0: skip; "
                         "the leg runs with the CPU baseline, not under --no-cpu-baseline)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--precision", choices=["fp32", "bf16", "fp8"], default="fp32", help="MLP GEMM precision")
    ap.add_argument("--table-precision", choices=["fp32", "fp16"], default="fp32",
                    help="hash-table storage the kernels gather from (f32 master kept for Adam)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--mode", choices=["throughput", "nrmse", "sweep", "warmstart", "kernels", "counters", "gscp"],
                    default="throughput")
    ap.add_argument("--ws-train", type=int, default=200, help="warmstart: training instances")
    ap.add_argument("--ws-epochs", type=int, default=8, help="warmstart: net training epochs")
    ap.add_argument("--ws-heldout", type=int, default=3, help="warmstart: held-out seeds")
    ap.add_argument("--loss", default=None, help="sweep mode: loss type (default Poisson_likelihood_loss)")
    ap.add_argument("--workload", default=None,
                    help="throughput: cfg1 (default) | cfg0 | cfg3;  nrmse mode: cfg0 (default) | cfg1 | small;  "
                         "gscp mode: gs50k_cholesky (default) | gs50k_rs")
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--sweep-every", type=int, default=40,
                    help="sweep mode: steps between points of the FP64 trajectory the variants are evaluated at")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB device copy (reference dispatch for HBM byte counters)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise two rocprofv3 --pmc runs of this bench into profiles/ and exit")
    ap.add_argument("--pmc-command", default="", help="the profiling command, recorded in the summary")
    ap.add_argument("--counters-groups", default=None,
                    help="counters mode: comma list of inr_fp32,inr_bf16,upsample,farfield,hashgrid (default all)")
    ap.add_argument("--pmc-table", nargs="+", metavar=("OUT_STEM", "CSV"),
                    help="per-kernel table of rocprofv3 --pmc counter CSVs into OUT_STEM.{md,json} and exit")
    args = ap.parse_args()
    if args.pmc_table:
        pmc_table(args.pmc_table[0], args.pmc_table[1:], args.pmc_command)
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, args.workload or "cfg1", args.precision, args.pmc_command)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    if args.mode == "nrmse":
        args.workload = args.workload or "cfg0"
        nrmse_mode(args)
        return
    if args.mode == "sweep":
        args.workload = args.workload or "cfg1"
        sweep_mode(args)
        return
    if args.mode == "warmstart":
        args.workload = args.workload or "cfg1"
        warmstart_mode(args)
        return
    if args.mode == "counters":
        counters_mode(args)
        return
    if args.mode == "kernels":
        kernels_mode(args)
        return
    if args.mode == "gscp":
        gscp_mode(args)
        return
    args.workload = args.workload or "cfg1"
    shapes, _, batch, loss_type, wl_name = WORKLOADS[args.workload]

    import torch
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        copy_calibration(dev)
    print(f"bench: generating the {wl_name} instance on the host", file=sys.stderr, flush=True)
    inst, model = build_problem(pkg, dev, precision=args.precision, workload=args.workload,
                                table_precision=args.table_precision)
    print("bench: instance ready", file=sys.stderr, flush=True)
    eng = pkg.FusedReconstruction(model, torch.from_numpy(inst.stack).to(dev), inst.tlY, inst.tlX,
                                  loss_type=loss_type, lr=1e-3, batch_size=batch)
    steps_per_iter = (inst.stack.shape[2] + batch - 1) // batch
    frames = eng.frames
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)

    state = {"perm": None, "pos": frames}

    def next_batch():
        # full batches only (the trailing partial batch of an epoch is skipped
        # here so that every timed step has the same shape)
        if state["pos"] + batch > frames:
            state["perm"] = torch.randperm(frames, device=dev, generator=gen).to(torch.int32)
            state["pos"] = 0
        b = state["perm"][state["pos"]:state["pos"] + batch]
        state["pos"] += batch
        return b

    eng.set_lr(1e-3)
    for _ in range(args.warmup):
        eng.step(next_batch())
    torch.cuda.synchronize()
    # timed region 1: K eager steps (one launch sequence per step from the
    # host), HIP events on the launch stream around each MLP-backward launch.
    # The object and probe chains share one stream here, so each launch's
    # duration is the kernel's own (with two streams a launch shares the GPU
    # with the other chain's kernels for a run-dependent part of its time).
    eng.overlap = False
    eng.enable_timing(["obj_mlp_bwd", "prb_mlp_bwd"])
    t0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        eng.step(next_batch())
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    eager_ms = start.elapsed_time(stop) / args.steps
    tim = eng.timings_ms()
    eng.enable_timing([])
    eng.overlap = True
    # the MLP-backward kernel runs twice a step (object grid, probe grid);
    # average launch duration and average points per launch over both
    launches = tim["obj_mlp_bwd"] + tim["prb_mlp_bwd"]
    bwd_ms = sum(launches) / len(launches)
    n_obj = model.downsampled_width * model.downsampled_height
    n_prb = model.probe_width * model.probe_height
    n_avg = (n_obj + n_prb) / 2
    achieved = MLP_FLOP_BWD_PER_POINT * n_avg / (bwd_ms * 1e-3) / 1e12
    peak = MFMA_PEAK_TFLOPS[args.precision]
    loss_now = float(eng.loss.sum().item())
    # timed region 2: the same steps replayed from the captured HIP graph, as
    # FusedReconstruction.run_epoch (the reconstruction loop) runs them: each
    # epoch's permutation is drawn and loaded once (inside the region), and
    # every replay takes its batch from it on the device
    graph_ms = None
    if not args.no_graph:
        eng.capture()
        full = frames // batch
        left = {"n": 0}

        def next_replay():
            if left["n"] == 0:
                eng.begin_order(torch.randperm(frames, device=dev, generator=gen).to(torch.int32))
                left["n"] = full
            eng.graph_next()
            left["n"] -= 1

        for _ in range(5):
            next_replay()
        torch.cuda.synchronize()
        g0 = torch.cuda.Event(enable_timing=True)
        g1 = torch.cuda.Event(enable_timing=True)
        t1 = time.perf_counter()
        g0.record()
        for _ in range(args.steps):
            next_replay()
        g1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        graph_ms = g0.elapsed_time(g1) / args.steps
    ms_per_step = graph_ms if graph_ms is not None else eager_ms
    steps_per_s = 1000.0 / ms_per_step

    line = {
        "metric": "wall-s to plain-final object NRMSE + iters/s: HIP fused vs plain, 1 MI355X",
        "value": steps_per_s / steps_per_iter,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" + ("" if args.precision == "fp32" else f" (MLP GEMMs {args.precision}, fp32 accumulate)")
                 + ("" if args.table_precision == "fp32" else " (hash tables fp16)"),
        "data": f"synthetic (seeded 1/f^2 object, soft-aperture probe, Poisson {shapes['photons']:.0e} photons)",
        "config": {"workload": wl_name, "steps_per_iter": steps_per_iter, "global_batch": batch,
                   "positions": frames, "parallelism": "single-gpu"},
        "steps_per_s": steps_per_s,
        "timed_path": "graph replay (run_epoch)" if graph_ms is not None else "eager",
        "graph_ms_per_step": graph_ms,
        "eager_ms_per_step": eager_ms,
        "wall_s_timed": wall,
        "kernel_ms": {k: (sum(v) / len(v) if v else None) for k, v in tim.items()},
        "loss_last_step": loss_now,
        "roofline": {"bound": "mfma", "kernel": MLP_BWD_KERNEL[args.precision],
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(args.workload, args.precision, args.table_precision),
                     "avg_launch_ms": bwd_ms, "launches": len(launches),
                     "algorithmic": f"{MLP_FLOP_BWD_PER_POINT} FLOP/point x {n_avg:.0f} points per launch "
                                    f"(mean of the object grid {n_obj} and the probe grid {n_prb})"},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline and args.workload == "cfg1":
        line["cpu_baseline"] = cpu_baseline(inst, args.cpu_seconds)
    if (args.nrmse_epochs > 0 and not args.no_cpu_baseline and args.workload == "cfg1"
            and args.precision == "fp32" and args.table_precision == "fp32"):
        # the wall-clock half of the metric, measured after the timed region on
        # the canvas-256 shapes: the plain loop takes about 10 s per epoch at
        # canvas 1024 and 0.36 s at canvas 256
        print(f"bench: NRMSE leg (canvas 256, {args.nrmse_epochs} epochs)", file=sys.stderr, flush=True)
        rec = nrmse_compare("cfg0", args.nrmse_epochs, 10)
        line["nrmse_leg"] = {
            "workload": WORKLOADS["cfg0"][4], "epochs": rec["epochs"],
            "plain_wall_s": rec["plain"]["wall_s"], "plain_final_nrmse": rec["plain"]["final_nrmse"],
            "target_nrmse": rec["target_nrmse"], "fused_epochs_to_target": rec["fused"]["epochs_to_target"],
            "fused_wall_s_to_target": rec["fused"]["wall_s_to_target"],
            "wall_speedup_to_target": rec["speedup_to_target"],
            "note": "plain = FP32 PyTorch restatement of the reference loop (pure-torch hash grid, per-position "
                    "loop) on this GPU; NRMSE checks every 10 epochs are inside both wall clocks",
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
