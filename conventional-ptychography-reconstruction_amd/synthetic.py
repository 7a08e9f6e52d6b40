"""Seeded synthetic ptychography instances (host, numpy FP64).

Template: scripts/simulate_and_reconstruct.py:85-156 of the reference (scan
positions, forward simulation |fftshift(fft2(P * O_window))|^2), extended
with photon scaling and Poisson noise.  Definitions:

* object: amplitude = real Gaussian random field with power spectrum 1/f^2
  (white noise times 1/|f| in Fourier space, DC removed), min-max mapped to
  [0.5, 1]; phase = independent field mapped to [-pi/2, pi/2];
* probe: circular aperture of radius R px with a Gaussian edge (sigma 2 px)
  and quadratic phase a*r^2 rad, centred at (M/2, M/2);
* scan: n x n raster with a fixed step plus U(-j, +j) jitter, rounded to
  integer pixels, centred on the canvas (clamped to it where the jitter
  would leave it);
* measurement: I = |fftshift(fft2(P * O_window))|^2 with one global scale on
  the probe so that mean_j sum I_j = photons, then Poisson noise; stored
  f32 [M, N, frames] (the reference layout).
Scan positions are also returned in metres for the reference geometry
(wavelength 0.675 um, camera length 0.0295 m, pitch 29.6 um).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

WAVELENGTH = 0.675e-6
CAMERA_LENGTH = 0.0295
PIXEL_PITCH = 2.96e-5


def gaussian_field_1f2(rng: np.random.Generator, n: int) -> np.ndarray:
    white = rng.standard_normal((n, n))
    fy = np.fft.fftfreq(n)[:, None]
    fx = np.fft.fftfreq(n)[None, :]
    f = np.sqrt(fx ** 2 + fy ** 2)
    filt = np.zeros_like(f)
    filt[f > 0] = 1.0 / f[f > 0]
    return np.real(np.fft.ifft2(np.fft.fft2(white) * filt))


def map_range(x: np.ndarray, lo: float, hi: float) -> np.ndarray:
    x = (x - x.min()) / (x.max() - x.min())
    return lo + (hi - lo) * x


def random_object(rng: np.random.Generator, n: int) -> np.ndarray:
    amp = map_range(gaussian_field_1f2(rng, n), 0.5, 1.0)
    phase = map_range(gaussian_field_1f2(rng, n), -np.pi / 2, np.pi / 2)
    return amp * np.exp(1j * phase)


def soft_aperture_probe(M: int, radius: float, quad: float, sigma: float = 2.0) -> np.ndarray:
    y, x = np.mgrid[0:M, 0:M].astype(np.float64)
    r = np.hypot(y - M / 2, x - M / 2)
    amp = np.where(r <= radius, 1.0, np.exp(-((r - radius) ** 2) / (2 * sigma ** 2)))
    return amp * np.exp(1j * quad * r ** 2)


def raster_scan(rng: np.random.Generator, n_side: int, step: int, jitter: int, M: int, canvas: int):
    gy, gx = np.mgrid[0:n_side, 0:n_side]
    ty = gy.ravel() * step + rng.integers(-jitter, jitter + 1, size=n_side * n_side)
    tx = gx.ravel() * step + rng.integers(-jitter, jitter + 1, size=n_side * n_side)
    ty = ty - ty.min()
    tx = tx - tx.min()
    ext_y, ext_x = ty.max() + M, tx.max() + M
    # centred on the canvas; jitter that would leave it is clamped
    ty = np.clip(ty + (canvas - ext_y) // 2, 0, canvas - M)
    tx = np.clip(tx + (canvas - ext_x) // 2, 0, canvas - M)
    return ty.astype(np.int64), tx.astype(np.int64)


def simulate(obj: np.ndarray, probe: np.ndarray, tlY, tlX) -> np.ndarray:
    M, N = probe.shape
    out = np.empty((M, N, len(tlY)), np.float64)
    for i, (ty, tx) in enumerate(zip(tlY, tlX)):
        ew = probe * obj[ty:ty + M, tx:tx + N]
        out[:, :, i] = np.abs(np.fft.fftshift(np.fft.fft2(ew))) ** 2
    return out


@dataclass
class Instance:
    obj: np.ndarray            # complex [canvas, canvas]
    probe: np.ndarray          # complex [M, M] (photon-scaled)
    tlY: np.ndarray
    tlX: np.ndarray
    stack: np.ndarray          # f32 [M, M, frames]
    xlocation: np.ndarray      # metres
    ylocation: np.ndarray
    canvas: int
    meta: dict = field(default_factory=dict)


# name -> (canvas, M, n_side, step, jitter, photons, radius_frac, quad_coef_at_64)
SHAPES = {
    "cfg0": dict(canvas=256, M=64, n_side=10, step=16, jitter=2, photons=1e4),
    "cfg1": dict(canvas=1024, M=256, n_side=20, step=40, jitter=4, photons=1e5),
    "cfg3": dict(canvas=3648, M=512, n_side=50, step=64, jitter=6, photons=1e6),
    "tiny": dict(canvas=64, M=16, n_side=4, step=8, jitter=1, photons=1e3),
    "small": dict(canvas=128, M=32, n_side=6, step=14, jitter=1, photons=1e3),
}


def make_instance(seed: int = 0, canvas: int = 256, M: int = 64, n_side: int = 10, step: int = 16,
                  jitter: int = 2, photons: float = 1e4, radius: float | None = None,
                  quad: float | None = None, noise: bool = True) -> Instance:
    rng = np.random.Generator(np.random.PCG64(seed))
    obj = random_object(rng, canvas)
    radius = 20.0 * M / 64 if radius is None else radius
    quad = 0.01 * (64.0 / M) ** 2 if quad is None else quad
    probe = soft_aperture_probe(M, radius, quad)
    tlY, tlX = raster_scan(rng, n_side, step, jitter, M, canvas)
    I = simulate(obj, probe, tlY, tlX)
    scale = photons / I.sum(axis=(0, 1)).mean()
    probe = probe * np.sqrt(scale)
    I = I * scale
    if noise:
        I = rng.poisson(I).astype(np.float64)
    dx = WAVELENGTH * CAMERA_LENGTH / (M * PIXEL_PITCH)
    return Instance(obj=obj, probe=probe, tlY=tlY, tlX=tlX, stack=I.astype(np.float32),
                    xlocation=tlX * dx, ylocation=tlY * dx, canvas=canvas,
                    meta=dict(seed=seed, M=M, n_side=n_side, step=step, jitter=jitter, photons=photons,
                              radius=radius, quad=quad, dx=dx))


def make_named(name: str, seed: int = 0, **kw) -> Instance:
    args = dict(SHAPES[name])
    args.update(kw)
    return make_instance(seed=seed, **args)
