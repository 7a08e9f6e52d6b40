"""Learned warm start for the object canvas.

The reference has no learned component: its object starts from the INR's
random initialisation (Recovery.py:201-222).  Here a small conv net maps each
diffraction pattern to a coarse object patch; the patches are overlap-averaged
into an initial canvas that seeds ``ComplexINRModel2D(object_initial=...,
use_residual=True)`` (complex_inr.py:209-250: object = k * INR + initial).

* Input per position: log(1 + I) average-pooled to 64 x 64, plus two constant
  planes holding the pattern's normalised scan offset (window centre / canvas
  - 1/2).
* Net: 4-level conv encoder-decoder, 32 base channels (``WarmStartNet``
  defines the architecture and draws the initial weights); forward, backward
  and Adam run on this package's HIP kernels (``HipWarmStartNet``,
  ``pty_wsnet_*``: im2col + BF16 MFMA GEMMs with FP32 accumulation, FP32
  master weights).
* Output per position: a 32 x 32 complex object patch (2 channels) = the
  object under the probe window, average-pooled by M / 32, in the scale the
  reconstruction converges to: the step normalises the probe to energy M*N
  (Recovery.py:280-285), so the object comes out multiplied by
  sqrt(E / (M*N)), E the energy of the photon-scaled probe.
* Training pairs are simulated from seeded instances with this package's
  far-field kernel (pty_farfield_intensity): random 1/f^2 objects, probes with
  the radius jittered by +-20 % and quadratic phase in [0, 0.02] (scaled to
  the probe size as in synthetic.py), Poisson noise at the instance photon
  count.  Fields are built on the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import _lib
from .farfield import farfield_intensity

IN_SIZE = 64
PATCH = 32


# ----------------------------------------------------------------------------
# simulated training data (GPU)
# ----------------------------------------------------------------------------
def _field_1f2(gen: torch.Generator, n: int, device) -> torch.Tensor:
    white = torch.randn(n, n, generator=gen, device=device, dtype=torch.float32)
    f = torch.fft.fftfreq(n, device=device)
    r = torch.sqrt(f[:, None] ** 2 + f[None, :] ** 2)
    filt = torch.where(r > 0, 1.0 / torch.clamp(r, min=1e-12), torch.zeros_like(r))
    x = torch.fft.ifft2(torch.fft.fft2(white) * filt).real
    return (x - x.min()) / (x.max() - x.min())


def random_object(gen, n, device) -> torch.Tensor:
    amp = 0.5 + 0.5 * _field_1f2(gen, n, device)
    phase = (_field_1f2(gen, n, device) - 0.5) * math.pi
    return torch.polar(amp, phase)


def soft_aperture_probe(M: int, radius: float, quad: float, device, sigma: float = 2.0) -> torch.Tensor:
    y = torch.arange(M, device=device, dtype=torch.float32)
    r = torch.sqrt((y[:, None] - M / 2) ** 2 + (y[None, :] - M / 2) ** 2)
    amp = torch.where(r <= radius, torch.ones_like(r), torch.exp(-((r - radius) ** 2) / (2 * sigma ** 2)))
    return torch.polar(amp, quad * r ** 2)


def raster(gen, n_side, step, jitter, M, canvas, device):
    g = torch.arange(n_side, device=device)
    gy, gx = torch.meshgrid(g, g, indexing="ij")
    ty = gy.reshape(-1) * step + torch.randint(-jitter, jitter + 1, (n_side * n_side,), generator=gen, device=device)
    tx = gx.reshape(-1) * step + torch.randint(-jitter, jitter + 1, (n_side * n_side,), generator=gen, device=device)
    ty, tx = ty - ty.min(), tx - tx.min()
    ty = torch.clamp(ty + (canvas - (ty.max() + M)) // 2, 0, canvas - M)
    tx = torch.clamp(tx + (canvas - (tx.max() + M)) // 2, 0, canvas - M)
    return ty, tx


@dataclass
class SimInstance:
    obj: torch.Tensor       # complex [canvas, canvas]
    probe: torch.Tensor     # complex [M, M] (photon-scaled)
    tlY: torch.Tensor
    tlX: torch.Tensor
    intens: torch.Tensor    # [frames, M, M] (reference orientation, fftshifted)


def simulate(seed: int, canvas=1024, M=256, n_side=20, step=40, jitter=4, photons=1e5, jitter_probe=True,
             device="cuda") -> SimInstance:
    """One seeded instance; |fftshift(fft2(P * O_window))|^2 by the HIP far-field kernel."""
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    obj = random_object(gen, canvas, device)
    u = torch.rand(2, generator=gen, device=device).tolist()
    radius = 20.0 * M / 64 * ((0.8 + 0.4 * u[0]) if jitter_probe else 1.0)
    quad = (0.02 * u[1] if jitter_probe else 0.01) * (64.0 / M) ** 2
    probe = soft_aperture_probe(M, radius, quad, device)
    ty, tx = raster(gen, n_side, step, jitter, M, canvas, device)
    tl = torch.stack([ty, tx], 1).to(torch.int32).contiguous()
    frames = tl.shape[0]
    idx = torch.arange(frames, dtype=torch.int32, device=device)
    I = farfield_intensity(obj, probe, idx, tl, frames)
    scale = photons / I.sum(dim=(1, 2)).mean()
    I = torch.poisson(I * scale, generator=gen)
    return SimInstance(obj=obj, probe=probe * torch.sqrt(scale), tlY=ty, tlX=tx, intens=I)


def _as_float(t, device) -> torch.Tensor:
    return t.to(device).float() if torch.is_tensor(t) else torch.as_tensor(np.asarray(t), device=device).float()


def net_inputs(intens: torch.Tensor, tlY, tlX, M: int, canvas: int) -> torch.Tensor:
    """[frames, 3, 64, 64]: pooled log(1 + I) and the two offset planes."""
    x = torch.log1p(intens.float()).unsqueeze(1)
    if M != IN_SIZE:
        x = F.adaptive_avg_pool2d(x, IN_SIZE) if M > IN_SIZE else F.interpolate(x, size=IN_SIZE, mode="bilinear")
    ty = _as_float(tlY, x.device)
    tx = _as_float(tlX, x.device)
    oy = ((ty + M / 2) / canvas - 0.5)[:, None, None, None].expand(-1, 1, IN_SIZE, IN_SIZE)
    ox = ((tx + M / 2) / canvas - 0.5)[:, None, None, None].expand(-1, 1, IN_SIZE, IN_SIZE)
    return torch.cat([x, oy, ox], 1)


def object_scale(probe: torch.Tensor) -> float:
    """Factor between the true object and the one a reconstruction with the
    energy-normalised probe converges to: sqrt(E / (M N))."""
    return float(torch.sqrt((probe.abs() ** 2).sum() / probe.numel()))


def patch_targets(obj: torch.Tensor, tlY, tlX, M: int) -> torch.Tensor:
    """[frames, 2, 32, 32]: the object under each window, pooled by M / 32 and
    rotated so that its mean has phase 0.  A single pattern does not
    determine the global phase of its window (|F(P O e^{i phi})| = |F(P O)|),
    so the net predicts the patch up to that phase; warm_start_object
    restores consistent phases by stitching."""
    ty = np.asarray(tlY.cpu() if torch.is_tensor(tlY) else tlY)
    tx = np.asarray(tlX.cpu() if torch.is_tensor(tlX) else tlX)
    wins = torch.stack([obj[int(a):int(a) + M, int(b):int(b) + M] for a, b in zip(ty, tx)])
    ri = F.adaptive_avg_pool2d(torch.stack([wins.real, wins.imag], 1).float(), PATCH)
    z = torch.complex(ri[:, 0], ri[:, 1])
    m = z.mean(dim=(1, 2), keepdim=True)
    z = z * torch.where(m.abs() > 0, m.conj() / m.abs().clamp_min(1e-30), torch.ones_like(m))
    return torch.stack([z.real, z.imag], 1)


# ----------------------------------------------------------------------------
# the net
# ----------------------------------------------------------------------------
def _block(cin, cout):
    return nn.Sequential(nn.Conv2d(cin, cout, 3, padding=1), nn.GELU(),
                         nn.Conv2d(cout, cout, 3, padding=1), nn.GELU())


class WarmStartNet(nn.Module):
    """4-level encoder-decoder: 64 -> 32 -> 16 -> 8 (channels c, 2c, 4c, 8c) and back to 32 x 32."""

    def __init__(self, base: int = 32):
        super().__init__()
        c = base
        self.e1, self.e2, self.e3, self.e4 = _block(3, c), _block(c, 2 * c), _block(2 * c, 4 * c), _block(4 * c, 8 * c)
        self.u3 = nn.ConvTranspose2d(8 * c, 4 * c, 2, stride=2)
        self.d3 = _block(8 * c, 4 * c)
        self.u2 = nn.ConvTranspose2d(4 * c, 2 * c, 2, stride=2)
        self.d2 = _block(4 * c, 2 * c)
        self.out = nn.Conv2d(2 * c, 2, 1)

    def forward(self, x):
        x1 = self.e1(x)                       # 64
        x2 = self.e2(F.avg_pool2d(x1, 2))     # 32
        x3 = self.e3(F.avg_pool2d(x2, 2))     # 16
        x4 = self.e4(F.avg_pool2d(x3, 2))     # 8
        y3 = self.d3(torch.cat([self.u3(x4), x3], 1))   # 16
        y2 = self.d2(torch.cat([self.u2(y3), x2], 1))   # 32
        return self.out(y2)                   # [B, 2, 32, 32]


def pack_params(module: WarmStartNet) -> torch.Tensor:
    """The module's parameters as the flat f32 vector of the HIP kernels:
    torch's parameter order, conv weights [Co][kh][kw][Ci], transposed-conv
    weights [kh][kw][Co][Ci]."""
    parts = []
    for mod in module.modules():
        if isinstance(mod, nn.ConvTranspose2d):
            parts += [mod.weight.detach().permute(2, 3, 1, 0).reshape(-1), mod.bias.detach().reshape(-1)]
        elif isinstance(mod, nn.Conv2d):
            parts += [mod.weight.detach().permute(0, 2, 3, 1).reshape(-1), mod.bias.detach().reshape(-1)]
    return torch.cat(parts).float().contiguous()


def _module_order_matches(module: WarmStartNet) -> bool:
    """modules() visits the convolutions in parameters() order."""
    convs = [m for m in module.modules() if isinstance(m, (nn.Conv2d, nn.ConvTranspose2d))]
    flat = [p for m in convs for p in (m.weight, m.bias)]
    return [id(p) for p in flat] == [id(p) for p in module.parameters()]


class HipWarmStartNet:
    """WarmStartNet trained and evaluated on the HIP kernels (pty_wsnet_*).

    FP32 master parameters in one flat buffer (``pack_params`` layout), their
    gradients and Adam's moments beside them; every GEMM in BF16 on MFMA with
    FP32 accumulation; Adam is ``pty_adam`` (torch.optim.Adam semantics)."""

    def __init__(self, base: int = 32, device="cuda", seed: int = 0, module: WarmStartNet | None = None):
        if module is None:
            with torch.random.fork_rng(devices=[]):
                torch.manual_seed(seed)
                module = WarmStartNet(base)
        assert _module_order_matches(module)
        self.base = base
        self.device = torch.device(device)
        n = _lib.query("pty_wsnet_param_count", base)
        self.params = pack_params(module).to(self.device)
        if self.params.numel() != n:
            raise ValueError(f"base {base}: {self.params.numel()} parameters, the kernels expect {n}")
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.hyper = torch.tensor([1e-3, 0.0], dtype=torch.float32, device=self.device)
        self._ws = {}
        self._mse_ws = _lib.workspace(4096 * 4, self.device)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.device)

    def _work(self, batch: int) -> torch.Tensor:
        if batch not in self._ws:
            self._ws = {batch: _lib.workspace(_lib.query("pty_wsnet_workspace", self.base, batch), self.device)}
        return self._ws[batch]

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """[B, 2, 32, 32] for inputs [B, 64, 64, 3] (NHWC f32, device)."""
        x = x.float().contiguous()
        B = x.shape[0]
        if x.shape[1:] != (64, 64, 3):
            raise ValueError(f"net input {tuple(x.shape)}, expected [B, 64, 64, 3]")
        out = torch.empty(B, 2, 32, 32, dtype=torch.float32, device=self.device)
        ws = self._work(B)
        _lib.call("pty_wsnet_forward", self.base, self.params.data_ptr(), x.data_ptr(), B, out.data_ptr(),
                  ws.data_ptr(), ws.numel(), _lib.stream_ptr(self.device))
        return out

    __call__ = forward

    def loss_and_grad(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        """MSE of the patches; leaves d loss / d params in ``grads``."""
        out = self.forward(x)
        y = y.float().contiguous()
        dout = torch.empty_like(out)
        s = _lib.stream_ptr(self.device)
        _lib.call("pty_wsnet_mse", out.data_ptr(), y.data_ptr(), out.numel(), self.loss.data_ptr(), dout.data_ptr(),
                  self._mse_ws.data_ptr(), self._mse_ws.numel(), s)
        ws = self._work(x.shape[0])
        _lib.call("pty_wsnet_backward", self.base, self.params.data_ptr(), x.shape[0], dout.data_ptr(),
                  self.grads.data_ptr(), ws.data_ptr(), ws.numel(), s)
        return self.loss

    def step(self, x, y, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8) -> torch.Tensor:
        """One Adam step on a batch; returns the batch loss (device)."""
        loss = self.loss_and_grad(x, y)
        s = _lib.stream_ptr(self.device)
        _lib.call("pty_fill", self.hyper.data_ptr(), 1, float(lr), s)
        _lib.call("pty_adam_tick", self.hyper.data_ptr(), s)
        _lib.call("pty_adam", self.params.data_ptr(), self.grads.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                  self.params.numel(), self.hyper.data_ptr(), float(betas[0]), float(betas[1]), float(eps), 1, s)
        return loss


def net_inputs_hip(intens: torch.Tensor, tlY, tlX, M: int, canvas: int) -> torch.Tensor:
    """[frames, 64, 64, 3] (NHWC) by the pty_wsnet_inputs kernel: pooled
    log(1 + I) and the window offset planes of ``net_inputs``."""
    dev = intens.device
    I = intens.float().contiguous()
    tl = torch.stack([_as_float(tlY, dev), _as_float(tlX, dev)], 1).to(torch.int32).contiguous()
    P = I.shape[0]
    x = torch.empty(P, 64, 64, 3, dtype=torch.float32, device=dev)
    _lib.call("pty_wsnet_inputs", I.data_ptr(), P, M, tl.data_ptr(), canvas, x.data_ptr(), _lib.stream_ptr(dev))
    return x


def build_dataset(seeds, device="cuda", **shape):
    """Stacked (inputs [N, 64, 64, 3], targets [N, 2, 32, 32]) f32 of the
    seeded instances.  The targets are the object patches themselves (values
    of order one); ``warm_start_object`` puts them on the reconstruction's
    scale with ``object_scale`` of the probe."""
    xs, ys = [], []
    M, canvas = shape.get("M", 256), shape.get("canvas", 1024)
    for s in seeds:
        inst = simulate(s, device=device, **shape)
        xs.append(net_inputs_hip(inst.intens, inst.tlY, inst.tlX, M, canvas))
        ys.append(patch_targets(inst.obj, inst.tlY, inst.tlX, M).float())
    return torch.cat(xs), torch.cat(ys)


def train(net: HipWarmStartNet, x: torch.Tensor, y: torch.Tensor, epochs: int = 30, lr: float = 1e-3,
          batch: int = 256, seed: int = 0, log=None) -> list[float]:
    """Adam on the MSE of the patches over shuffled minibatches (the last one
    may be smaller); returns the mean loss of each epoch."""
    gen = torch.Generator(device=x.device)
    gen.manual_seed(seed)
    n = x.shape[0]
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen, device=x.device)
        tot = torch.zeros((), device=x.device)
        for b in range(0, n, batch):
            i = perm[b:b + batch]
            tot = tot + net.step(x[i], y[i], lr=lr)[0] * i.numel()
        hist.append(float(tot) / n)
        if log:
            log(ep + 1, hist[-1])
    return hist


@torch.no_grad()
def warm_start_object(net: HipWarmStartNet, intens: torch.Tensor, tlY, tlX, M: int, canvas: int,
                      probe: torch.Tensor | None = None, batch: int = 512) -> torch.Tensor:
    """Initial complex canvas from the patterns [frames, M, M] (fftshifted,
    reference orientation): predicted patches upsampled to the window, each
    rotated to agree in phase with the canvas built so far over their overlap
    (positions in scan order), and overlap-averaged; uncovered pixels get the
    mean of the covered ones.  With the initial probe, the canvas is put on
    the scale the reconstruction converges to (``object_scale``)."""
    x = net_inputs_hip(intens, tlY, tlX, M, canvas)
    preds = [net(x[b:b + batch]) for b in range(0, x.shape[0], batch)]
    p = F.interpolate(torch.cat(preds), size=M, mode="bilinear", align_corners=False)
    pz = torch.complex(p[:, 0], p[:, 1])
    acc = torch.zeros(canvas, canvas, dtype=torch.complex64, device=x.device)
    cnt = torch.zeros(canvas, canvas, device=x.device)
    ty = np.asarray(tlY.cpu() if torch.is_tensor(tlY) else tlY)
    tx = np.asarray(tlX.cpu() if torch.is_tensor(tlX) else tlX)
    for k, (a, b) in enumerate(zip(ty, tx)):
        sl = (slice(int(a), int(a) + M), slice(int(b), int(b) + M))
        c = (acc[sl].conj() * pz[k]).sum()          # <canvas, patch> over the overlap
        rot = torch.where(c.abs() > 0, c.conj() / c.abs().clamp_min(1e-30), torch.ones_like(c))
        acc[sl] += pz[k] * rot
        cnt[sl] += 1
    cov = cnt > 0
    out = torch.empty(canvas, canvas, dtype=torch.complex64, device=x.device)
    out[cov] = acc[cov] / cnt[cov]
    out[~cov] = out[cov].mean() if cov.any() else 1.0
    if probe is not None:
        out = out * object_scale(probe)
    return out


def stack_to_patterns(stack_mnf: torch.Tensor) -> torch.Tensor:
    """Reference stack layout [M, N, frames] -> [frames, M, N]."""
    return stack_mnf.permute(2, 0, 1).contiguous()
