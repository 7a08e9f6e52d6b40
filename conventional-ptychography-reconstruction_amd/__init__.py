"""MI355X (gfx950) implementation of the INR far-field ptychography hot path.

Drop-in surface (same names and semantics as the reference):
  ComplexINRModel2D, ComplexMLP, ComplexGaborLayer2D   models/complex_inr.py
  HashGridEncoding (for tcnn.Encoding "HashGrid")      models/complex_inr.py:138-145
  normalize_probe, farfield_loss                       Recovery.py:262-307
  haar_wavelet_sparsity_loss                           utils/loss.py:60-83
  scan_positions_to_pixels                             utils/optics.py:68-99
  load_config / load_ptychography_data / save_model    Recovery.py:44-72, utils/io.py
  gscp.GaussianRasterize2D[Cholesky], gscp.rasterize_* gscp/cuda/_wrapper.py, gscp/csrc (the
                                                       Gaussian-field rasterizer)
The learned object warm start (``warmstart``) is an addition without a
counterpart in the reference.
Compute runs in the hand-written HIP kernels of ``libptycho_hip.so`` (C ABI in
include/ptycho_hip.h); there is no CPU or PyTorch fallback.
"""
from . import _lib, gscp, loss, metrics, recovery, synthetic, warmstart
from .complex_inr import (ComplexGaborLayer2D, ComplexINRModel2D, ComplexMLP, CoordPoints, GridPoints,
                          InrFieldFn, ResidualFn, UpsampleFn, fourier_upsample, mlp_descriptor, mlp_forward)
from .engine import FlatParams, FusedReconstruction
from .farfield import (VALID_LOSS_TYPES, farfield_intensity, farfield_loss, normalize_probe, positions_tensor,
                       prepare_stack)
from .hashgrid import HashGridEncoding, HashGridFn, hashgrid_indices
from .levels import GridSpec, level_table
from .loss import haar_wavelet_sparsity_loss
from .optics import center_windows, scan_positions_to_pixels

__all__ = [
    "ComplexGaborLayer2D", "ComplexINRModel2D", "ComplexMLP", "CoordPoints", "GridPoints", "InrFieldFn",
    "ResidualFn", "UpsampleFn", "fourier_upsample", "mlp_descriptor", "mlp_forward", "FlatParams",
    "FusedReconstruction", "VALID_LOSS_TYPES", "farfield_intensity", "farfield_loss", "normalize_probe",
    "positions_tensor", "prepare_stack", "HashGridEncoding", "HashGridFn", "hashgrid_indices", "GridSpec",
    "level_table", "center_windows", "scan_positions_to_pixels", "haar_wavelet_sparsity_loss",
]
