"""ctypes binding of the C ABI in ``include/ptycho_hip.h``.

The library is loaded from the package directory.  There is no fallback: if
``libptycho_hip.so`` is missing or a call fails, a RuntimeError is raised.
"""
from __future__ import annotations

import ctypes
from ctypes import c_float, c_int32, c_int64, c_size_t, c_uint32, c_void_p
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libptycho_hip.so"
ABI_VERSION = 7
MAX_LEVELS = 16
MAX_HIDDEN = 4
FAR_SCRATCH_BYTES = 16 << 20   # PTY_FAR_SCRATCH_BYTES

LOSS_CODES = {"smooth_L1_loss": 0, "FD_loss": 1, "Poisson_likelihood_loss": 2}
PREC_CODES = {"fp32": 0, "bf16": 1, "fp8": 2}
TABLE_CODES = {"fp32": 0, "fp16": 1}


class PtyGrid(ctypes.Structure):
    _fields_ = [
        ("abi_version", c_int32), ("n_levels", c_int32), ("n_features", c_int32),
        ("log2_hashmap_size", c_int32),
        ("scale", c_float * MAX_LEVELS), ("res", c_uint32 * MAX_LEVELS),
        ("size", c_uint32 * MAX_LEVELS), ("offset", c_uint32 * MAX_LEVELS),
        ("dense_mask", c_uint32), ("n_entries", c_uint32), ("table_precision", c_uint32),
    ]


class PtyMlp(ctypes.Structure):
    _fields_ = [
        ("n_in", c_int32), ("n_hidden", c_int32), ("precision", c_int32), ("pad_", c_int32),
        ("w0", c_void_p), ("b0", c_void_p),
        ("wh", c_void_p * MAX_HIDDEN), ("bh", c_void_p * MAX_HIDDEN),
        ("omega", c_void_p * (MAX_HIDDEN + 1)),
        ("wo", c_void_p), ("bo", c_void_p), ("fp8_state", c_void_p),
    ]


class PtyPoints(ctypes.Structure):
    _fields_ = [
        ("coords", c_void_p), ("axis0", c_void_p), ("axis1", c_void_p),
        ("n_axis0", c_int32), ("n_axis1", c_int32), ("n", c_int64), ("plan", c_void_p),
        ("features", c_void_p),
    ]


class PtyFar(ctypes.Structure):
    _fields_ = [
        ("M", c_int32), ("N", c_int32), ("canvas_h", c_int32), ("canvas_w", c_int32),
        ("n_batch", c_int32), ("n_frames", c_int32), ("loss_type", c_int32), ("pad_", c_int32),
    ]


class PtyLayer(ctypes.Structure):
    _fields_ = [
        ("n_in", c_int32), ("n_out", c_int32), ("complex_io", c_int32), ("activation", c_int32),
        ("weight", c_void_p), ("bias", c_void_p), ("omega", c_void_p),
    ]


class PtyGs(ctypes.Structure):
    _fields_ = [
        ("n", c_int32), ("width", c_int32), ("height", c_int32), ("max_patch_radius", c_int32),
        ("min_scale", c_float), ("param", c_int32),
    ]


GS_PARAM_CODES = {"rs": 0, "cholesky": 1}

P = ctypes.POINTER
_SIGNATURES = {
    "pty_abi_version": (c_int32, []),
    "pty_last_error": (ctypes.c_char_p, []),
    "pty_hashgrid_fwd": (c_int32, [P(PtyGrid), c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "pty_hashgrid_bwd": (c_int32, [P(PtyGrid), c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "pty_hashgrid_indices": (c_int32, [P(PtyGrid), c_void_p, c_int64, c_void_p, c_void_p]),
    "pty_inr_fwd": (c_int32, [P(PtyGrid), P(PtyMlp), c_void_p, P(PtyPoints), c_void_p, c_void_p]),
    "pty_inr_bwd_workspace": (c_size_t, [P(PtyGrid), P(PtyMlp), P(PtyPoints)]),
    "pty_inr_bwd": (c_int32, [P(PtyGrid), P(PtyMlp), P(PtyMlp), c_void_p, c_void_p, P(PtyPoints),
                              c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_inr_bwd_mlp": (c_int32, [P(PtyGrid), P(PtyMlp), c_void_p, P(PtyPoints), c_void_p, c_void_p, c_size_t,
                                  c_void_p]),
    "pty_inr_bwd_finish": (c_int32, [P(PtyGrid), P(PtyMlp), P(PtyMlp), c_void_p, P(PtyPoints), c_void_p, c_size_t,
                                     c_void_p]),
    "pty_inr_bwd_finish_adam": (c_int32, [P(PtyGrid), P(PtyMlp), P(PtyMlp), c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_float, c_float, c_float, P(PtyPoints), c_void_p, c_size_t, c_void_p]),
    "pty_points_plan_bytes": (c_size_t, [P(PtyGrid), P(PtyPoints)]),
    "pty_points_feature_floats": (c_size_t, [c_int64]),
    "pty_points_plan_build": (c_int32, [P(PtyGrid), P(PtyPoints), c_void_p, c_size_t, c_void_p]),
    "pty_mlp_fwd": (c_int32, [P(PtyMlp), c_void_p, c_int64, c_void_p, c_void_p]),
    "pty_layer_fwd": (c_int32, [P(PtyLayer), c_void_p, c_int64, c_void_p, c_void_p]),
    "pty_layer_bwd_workspace": (c_size_t, [P(PtyLayer), c_int64]),
    "pty_layer_bwd": (c_int32, [P(PtyLayer), c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_size_t, c_void_p]),
    "pty_upsample_workspace": (c_size_t, [c_int32, c_int32]),
    "pty_upsample_fwd": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_upsample_adj": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_residual_fwd": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "pty_residual_workspace": (c_size_t, [c_int64]),
    "pty_residual_bwd": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                   c_size_t, c_void_p]),
    "pty_probe_workspace": (c_size_t, [c_int32, c_int32]),
    "pty_probe_normalize_fwd": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                                          c_size_t, c_void_p]),
    "pty_probe_normalize_bwd": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_void_p, c_void_p,
                                          c_void_p, c_size_t, c_void_p]),
    "pty_stack_prepare": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p]),
    "pty_farfield_workspace": (c_size_t, [P(PtyFar)]),
    "pty_farfield_fwd_bwd": (c_int32, [P(PtyFar), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_farfield_fwd_bwd_noalloc": (c_int32, [P(PtyFar), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                               c_void_p, c_void_p, c_void_p, c_void_p]),
    "pty_farfield_intensity": (c_int32, [P(PtyFar), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_size_t, c_void_p]),
    "pty_adam": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_float, c_float,
                           c_float, c_int32, c_void_p]),
    "pty_adam_tick": (c_int32, [c_void_p, c_void_p]),
    "pty_batch_next": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_int64, c_void_p]),
    "pty_fill": (c_int32, [c_void_p, c_int64, c_float, c_void_p]),
    "pty_gs_geom_bytes": (c_size_t, [P(PtyGs)]),
    "pty_gs_preprocess": (c_int32, [P(PtyGs), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pty_gs_num_rendered": (c_int32, [P(PtyGs), c_void_p, P(c_int64), c_void_p]),
    "pty_gs_bin_bytes": (c_size_t, [P(PtyGs), c_int64, c_int32]),
    "pty_gs_render_fwd": (c_int32, [P(PtyGs), c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pty_gs_render_bwd": (c_int32, [P(PtyGs), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pty_wsnet_param_count": (c_size_t, [c_int32]),
    "pty_wsnet_workspace": (c_size_t, [c_int32, c_int32]),
    "pty_wsnet_inputs": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p]),
    "pty_wsnet_forward": (c_int32, [c_int32, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_wsnet_mse": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_wsnet_backward": (c_int32, [c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pty_cast_f16": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p]),
    "pty_fp8_roll": (c_int32, [c_void_p, c_void_p]),
    "pty_haar_workspace": (c_size_t, [c_int64, c_int32]),
    "pty_haar_sparsity": (c_int32, [c_void_p, c_int64, c_int32, c_float, c_float, c_void_p, c_void_p, c_void_p,
                                    c_size_t, c_void_p]),
}

_lib = None


def load():
    """Load the shared library (once).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH.name} is not built; run build() of the package "
                           f"(hipcc --offload-arch=gfx950) first")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pty_abi_version() != ABI_VERSION:
        raise RuntimeError(f"ABI version mismatch: library {lib.pty_abi_version()} != {ABI_VERSION}")
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGNATURES)


def call(name: str, *args):
    """Call an int-returning entry point; raise RuntimeError on a nonzero code."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.pty_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc


def query(name: str, *args) -> int:
    return int(getattr(load(), name)(*args))


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def require_cuda(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RuntimeError("the HIP path needs device tensors (got a CPU tensor)")


def workspace(nbytes: int, device) -> torch.Tensor:
    """Device scratch of at least nbytes (from torch's caching allocator)."""
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
