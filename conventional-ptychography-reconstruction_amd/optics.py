"""Scan geometry on the host (FP64), utils/optics.py:68-99 of the reference."""
from __future__ import annotations

import math

import numpy as np


def scan_positions_to_pixels(xlocation, ylocation, probe_shape, wavelength, camera_length,
                             camera_pixel_pitch):
    """Physical scan positions (m) -> top-left / bottom-right pixel indices.

    dx = wavelength * camera_length / ([M, N] * pitch) in FP64; tl =
    round((pos - min) / dx) with numpy's round-half-to-even; br = tl + [M, N];
    canvas = ceil(max(br) / 6) * 6.
    Returns (tlX, tlY, brX, brY, dx, canvas_size).
    """
    M, N = probe_shape
    x_pos = np.asarray(xlocation, dtype=np.float64).squeeze() - np.min(xlocation)
    y_pos = np.asarray(ylocation, dtype=np.float64).squeeze() - np.min(ylocation)
    dx = wavelength * camera_length / (np.array([M, N], dtype=np.float64) * camera_pixel_pitch)
    tlY = np.round(y_pos / dx[0]).astype(int)
    tlX = np.round(x_pos / dx[1]).astype(int)
    brY = tlY + M
    brX = tlX + N
    max_size = max(int(np.max(brY)), int(np.max(brX)))
    canvas_size = math.ceil(max_size / 6) * 6
    return tlX, tlY, brX, brY, dx, canvas_size


def center_windows(tlX, tlY, probe_shape, canvas_size):
    """Shift the scan so that its footprint is centred on an explicit canvas.

    Used when the canvas is given explicitly (``canvas_size`` config key)
    instead of the reference's ceil(max/6)*6 rule.
    """
    M, N = probe_shape
    ext_y = int(np.max(tlY)) + M
    ext_x = int(np.max(tlX)) + N
    if ext_y > canvas_size or ext_x > canvas_size:
        raise ValueError(f"scan extent {ext_y}x{ext_x} exceeds canvas {canvas_size}")
    oy = (canvas_size - ext_y) // 2
    ox = (canvas_size - ext_x) // 2
    return np.asarray(tlX) + ox, np.asarray(tlY) + oy
