"""Drop-in for Kennett_Reflectivity/kennett_reflectivity_gpu.py.

``kennett_reflectivity_batch`` keeps the reference signature and return
convention (kennett_reflectivity_gpu.py:88-124): NumPy in, a new (np, nw)
complex128 NumPy array out, inputs untouched.  The recursion runs in the HIP
kernel ``kh_forward`` (one work-item per (p, omega) lane) instead of the
reference's full-grid NumPy temporaries; frequency grids of the form
(k0 + j)*dw + i*gamma (every grid the reference builds) take ``kh_forward_grid``.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _abi, engine
from .layer_model import LayerModel

__all__ = ["kennett_reflectivity_batch", "get_device"]


def get_device() -> torch.device:
    """The HIP device the kernels run on (kennett_reflectivity_gpu.py:39-53
    prefers MPS > CUDA > CPU; here a ROCm GPU is required and there is no CPU
    fallback)."""
    return _abi.device()


def kennett_reflectivity_batch(
    model: LayerModel,
    p_samples: np.ndarray,
    omega: np.ndarray,
    free_surface: bool = False,
) -> np.ndarray:
    """PP reflectivity for all slownesses and frequencies, shape (np, nw)."""
    p_samples = np.ascontiguousarray(np.asarray(p_samples, dtype=np.float64).reshape(-1))
    omega_c = np.ascontiguousarray(np.asarray(omega).astype(np.complex128).reshape(-1))
    dev = get_device()
    layers = engine.layers_tensor(model, dev)
    p_t = torch.from_numpy(p_samples).to(dev)
    grid = engine.uniform_grid(omega_c)
    if grid is not None:
        R = engine.forward_grid(layers, p_t, *grid, omega_c.shape[0], free_surface=free_surface)
    else:
        R = engine.forward(layers, p_t, torch.from_numpy(omega_c).to(dev),
                           free_surface=free_surface)
    return R.cpu().numpy()
