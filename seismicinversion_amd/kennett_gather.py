"""Drop-in for the float64 gather of Kennett_Reflectivity/kennett_gather.py.

``compute_gather`` (kennett_gather.py:127-305) keeps its signature; the
reference spreads slowness batches over a multiprocessing pool, here the
whole sum is one device pass (``n_workers`` is accepted and ignored).  It
differs from ``compute_gather_gpu`` only in its defaults and in tapering with
the ``p_max`` argument instead of the last slowness (kennett_gather.py:228).
"""

from __future__ import annotations

import collections.abc

import numpy as np

from .kennett_gather_gpu import _gather
from .layer_model import LayerModel
from .workloads import default_ocean_crust_model

__all__ = ["compute_gather", "default_ocean_crust_model"]


def compute_gather(
    model: LayerModel,
    offsets: np.ndarray,
    T: float = 64.0,
    nw: int = 2048,
    np_slow: int = 2048,
    p_max: float = 1.0,
    gamma: float | None = None,
    source_func: collections.abc.Callable | None = None,
    n_workers: int | None = None,
    free_surface: bool = False,
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    return _gather(model, offsets, T, nw, np_slow, p_max, gamma, source_func, free_surface,
                   None, taper_p_max=p_max)
