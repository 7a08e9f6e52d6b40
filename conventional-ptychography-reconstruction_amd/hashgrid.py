"""Drop-in for ``tcnn.Encoding(n_input_dims=2, encoding_config={"otype": "HashGrid", ...})``.

Reference call sites: models/complex_inr.py:130-145 (construction), :279 and
:294 (``encoder(coords)``).  The module owns one fp32 parameter ``params`` of
shape [n_entries * n_features] (the tcnn layout, entry-major with features
interleaved), initialised U(-1e-4, 1e-4) from its own generator (seed 1337)
so that the global torch RNG stream seen by the MLP initialisers is the same
as in the reference.  Forward and backward run the HIP kernels
``pty_hashgrid_fwd`` / ``pty_hashgrid_bwd``; output is float32 [N, L*F].

``dtype=torch.float16`` (tcnn's half-precision setting) keeps ``params`` as
the f32 master and has the forward kernels gather from an FP16 copy of it,
refreshed by ``pty_cast_f16`` on every forward; gradients stay f32.
"""
from __future__ import annotations

import torch
import torch.nn as nn

from . import _lib
from .levels import GridSpec, level_table


def table_storage(params: torch.Tensor, table_precision: str) -> torch.Tensor:
    """The tensor the forward kernels read: ``params`` itself (fp32) or a
    fresh FP16 copy of it (round to nearest even, on the device)."""
    if table_precision == "fp32":
        return params
    if table_precision != "fp16":
        raise ValueError(f"table_precision must be one of {tuple(_lib.TABLE_CODES)}")
    _lib.require_cuda(params)
    src = params.detach().contiguous()
    half = torch.empty(src.numel(), dtype=torch.float16, device=src.device)
    _lib.call("pty_cast_f16", src.data_ptr(), src.numel(), half.data_ptr(), _lib.stream_ptr(src.device))
    return half


class HashGridFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, params, spec: GridSpec, table_precision: str = "fp32"):
        _lib.require_cuda(coords, params)
        coords = coords.contiguous().float()
        n = coords.shape[0]
        out = torch.empty(n, spec.n_output_dims, device=coords.device, dtype=torch.float32)
        desc = spec.descriptor(table_precision)
        table = table_storage(params, table_precision)
        _lib.call("pty_hashgrid_fwd", desc, coords.data_ptr(), n, table.data_ptr(), out.data_ptr(),
                  _lib.stream_ptr(coords.device))
        ctx.save_for_backward(coords)
        ctx.spec = spec
        return out

    @staticmethod
    def backward(ctx, dout):
        (coords,) = ctx.saved_tensors
        spec = ctx.spec
        dout = dout.contiguous().float()
        dparams = torch.zeros(spec.n_params, device=coords.device, dtype=torch.float32)
        _lib.call("pty_hashgrid_bwd", spec.descriptor(), coords.data_ptr(), coords.shape[0], dout.data_ptr(),
                  dparams.data_ptr(), _lib.stream_ptr(coords.device))
        return None, dparams, None, None


def hashgrid_indices(coords: torch.Tensor, spec: GridSpec) -> torch.Tensor:
    """Corner entry indices [N, L, 4] (int64) computed by the device kernel."""
    _lib.require_cuda(coords)
    coords = coords.contiguous().float()
    n = coords.shape[0]
    idx = torch.empty(n, spec.n_levels, 4, dtype=torch.int32, device=coords.device)
    _lib.call("pty_hashgrid_indices", spec.descriptor(), coords.data_ptr(), n, idx.data_ptr(),
              _lib.stream_ptr(coords.device))
    return idx.to(torch.int64) & 0xFFFFFFFF


class HashGridEncoding(nn.Module):
    """HashGrid encoder with the constructor and ``params`` layout of tcnn.Encoding."""

    def __init__(self, n_input_dims: int, encoding_config: dict, seed: int = 1337, dtype=None):
        super().__init__()
        if n_input_dims != 2:
            raise ValueError("the gfx950 HashGrid encoder is 2-D (n_input_dims=2)")
        if encoding_config.get("otype", "HashGrid") != "HashGrid":
            raise ValueError(f"unsupported encoding {encoding_config.get('otype')!r}")
        self.n_input_dims = n_input_dims
        self.encoding_config = dict(encoding_config)
        self.spec = level_table(
            n_levels=int(encoding_config.get("n_levels", 16)),
            n_features=int(encoding_config.get("n_features_per_level", 2)),
            log2_hashmap_size=int(encoding_config.get("log2_hashmap_size", 19)),
            base_resolution=int(encoding_config.get("base_resolution", 16)),
            per_level_scale=float(encoding_config.get("per_level_scale", 2.0)),
        )
        if dtype is None or dtype == torch.float32:
            self.table_precision = "fp32"
        elif dtype == torch.float16:
            self.table_precision = "fp16"
        else:
            raise ValueError(f"HashGrid tables are float32 or float16, not {dtype}")
        gen = torch.Generator().manual_seed(seed)
        init = (torch.rand(self.spec.n_params, generator=gen, dtype=torch.float32) * 2 - 1) * 1e-4
        self.params = nn.Parameter(init)

    @property
    def n_output_dims(self) -> int:
        return self.spec.n_output_dims

    def n_params(self) -> int:
        return self.spec.n_params

    def descriptor(self) -> _lib.PtyGrid:
        return self.spec.descriptor(self.table_precision)

    def forward(self, coords):
        return HashGridFn.apply(coords, self.params, self.spec, self.table_precision)
