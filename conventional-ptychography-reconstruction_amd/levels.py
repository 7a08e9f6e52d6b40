"""Host-side hash-grid level table (built once per encoder).

Follows the multiresolution hash encoding configured at
models/complex_inr.py:130-137 ("HashGrid", n_levels, n_features_per_level,
log2_hashmap_size, base_resolution, per_level_scale):

  scale_l  = f32(exp2(f32(l) * f32(log2(per_level_scale)))) * base - 1   (f32 steps)
  res_l    = ceil(scale_l) + 1
  size_l   = min(next_multiple_of_8(res_l^2), 2^log2_hashmap_size)
  offset_l = sum of the sizes of the coarser levels
  dense when res_l^2 <= size_l, hashed otherwise

The table is computed once in f32 on the host and handed to the kernels as
is, so that the device never recomputes exp2 (the integer-scale levels would
otherwise flip floor() with one ulp of exp2f).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass(frozen=True)
class GridSpec:
    n_levels: int
    n_features: int
    log2_hashmap_size: int
    base_resolution: int
    per_level_scale: float
    scale: tuple
    res: tuple
    size: tuple
    offset: tuple
    dense: tuple
    n_entries: int

    @property
    def n_params(self) -> int:
        return self.n_entries * self.n_features

    @property
    def n_output_dims(self) -> int:
        return self.n_levels * self.n_features

    def descriptor(self, table_precision: str = "fp32") -> _lib.PtyGrid:
        """C descriptor; ``table_precision`` names the storage the kernels read
        ('fp32' params, or 'fp16' for the half copy made by pty_cast_f16)."""
        d = _lib.PtyGrid()
        d.table_precision = _lib.TABLE_CODES[table_precision]
        d.abi_version = _lib.ABI_VERSION
        d.n_levels = self.n_levels
        d.n_features = self.n_features
        d.log2_hashmap_size = self.log2_hashmap_size
        mask = 0
        for l in range(self.n_levels):
            d.scale[l] = float(self.scale[l])
            d.res[l] = int(self.res[l])
            d.size[l] = int(self.size[l])
            d.offset[l] = int(self.offset[l])
            if self.dense[l]:
                mask |= 1 << l
        d.dense_mask = mask
        d.n_entries = self.n_entries
        return d


def _next_multiple(v: int, m: int) -> int:
    return (v + m - 1) // m * m


def level_table(n_levels: int = 16, n_features: int = 2, log2_hashmap_size: int = 19,
                base_resolution: int = 16, per_level_scale: float = 1.5) -> GridSpec:
    if n_features != 2:
        raise ValueError("n_features_per_level must be 2 for the gfx950 encoder")
    if not 1 <= n_levels <= _lib.MAX_LEVELS:
        raise ValueError(f"n_levels must be in 1..{_lib.MAX_LEVELS}")
    f32 = np.float32
    log2_pls = f32(np.log2(f32(per_level_scale)))
    T = 1 << int(log2_hashmap_size)
    scale, res, size, offset, dense = [], [], [], [], []
    off = 0
    for l in range(n_levels):
        s = f32(f32(np.exp2(f32(l) * log2_pls)) * f32(base_resolution) - f32(1.0))
        r = int(math.ceil(float(s))) + 1
        # the 2^31-1 cap on r^2 only matters for absurd resolutions
        n = min(_next_multiple(min(r * r, (1 << 31) - 1), 8), T)
        scale.append(float(s))
        res.append(r)
        size.append(n)
        offset.append(off)
        dense.append(r * r <= n)
        off += n
    return GridSpec(n_levels, n_features, int(log2_hashmap_size), int(base_resolution),
                    float(per_level_scale), tuple(scale), tuple(res), tuple(size), tuple(offset),
                    tuple(dense), off)
