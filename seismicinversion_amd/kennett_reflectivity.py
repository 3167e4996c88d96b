"""Drop-in for Kennett_Reflectivity/kennett_reflectivity.py.

``kennett_reflectivity(model, p, omega, free_surface)`` (kennett_reflectivity.py:73-78)
is the batched kernel at a single slowness.  ``inv2x2``/``batch_inv2x2``
(kennett_reflectivity.py:17-51) are the reference's small host utilities,
kept for callers that import them.
"""

from __future__ import annotations

import numpy as np

from .kennett_reflectivity_gpu import kennett_reflectivity_batch
from .layer_model import LayerModel

__all__ = ["kennett_reflectivity", "inv2x2", "batch_inv2x2"]


def kennett_reflectivity(model: LayerModel, p: float, omega: np.ndarray,
                         free_surface: bool = False) -> np.ndarray:
    """PP reflectivity at one slowness for every frequency, shape (nw,)."""
    return kennett_reflectivity_batch(model, np.array([float(p)]), omega, free_surface)[0]


def inv2x2(M: np.ndarray) -> np.ndarray:
    a, b, c, d = M[0, 0], M[0, 1], M[1, 0], M[1, 1]
    det = a * d - b * c
    return np.array([[d, -b], [-c, a]], dtype=M.dtype) / det


def batch_inv2x2(M: np.ndarray) -> np.ndarray:
    a, b, c, d = M[..., 0, 0], M[..., 0, 1], M[..., 1, 0], M[..., 1, 1]
    det = a * d - b * c
    out = np.empty_like(M)
    out[..., 0, 0] = d / det
    out[..., 0, 1] = -b / det
    out[..., 1, 0] = -c / det
    out[..., 1, 1] = a / det
    return out
