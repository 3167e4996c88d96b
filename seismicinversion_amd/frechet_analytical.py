"""Drop-in for Kennett_Reflectivity/frechet_analytical.py.

``frechet_kennett`` (frechet_analytical.py:583-911) returns the analytical
Frechet Jacobian J[i, j] = dR(omega_i)/dm_j of the PP reflectivity for the
packed parameters [Vp_1..N, Vs_1..N, rho_1..N, h_1..N-1] (ocean fixed,
half-space thickness excluded).  The reference differentiates the recursion
parameter by parameter in NumPy; here every column is one tangent-linear
sweep of the device kernel kh_jvp, all columns in one launch.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _abi, derivatives, engine
from .kennett_torch import _basis

__all__ = ["frechet_kennett"]


def frechet_kennett(alpha, beta, rho, thickness, Q_alpha, Q_beta, p: float,
                    omega: np.ndarray) -> np.ndarray:
    """Complex (nfreq, 4N - 1) Jacobian; thickness uses np.inf for the half-space."""
    tab = np.stack([np.asarray(x, dtype=np.float64) for x in
                    (alpha, beta, rho, thickness, Q_alpha, Q_beta)], axis=1)
    dev = _abi.device()
    layers = engine.layers_tensor(tab, dev)
    _, dR = derivatives.jvp(layers, _basis(tab.shape[0]), torch.tensor([float(p)],
                                                                       dtype=torch.float64),
                            torch.from_numpy(np.asarray(omega, dtype=np.float64)
                                             .astype(np.complex128)))
    return dR[:, 0, :].T.cpu().numpy()
