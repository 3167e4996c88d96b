"""Layer model container for plane-stratified elastic media.

Mirrors the public surface of Kennett_Reflectivity/layer_model.py:
``LayerModel`` (same fields, same validation and error messages,
layer_model.py:77-167), ``complex_slowness`` (layer_model.py:15-45) and
``vertical_slowness`` (layer_model.py:48-74).

The scalar helpers here are host conveniences for callers that build models
by hand; the reflectivity kernels compute slownesses themselves, per
(layer, slowness) lane, on the device (see ``csrc/kh_physics.hpp``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["LayerModel", "complex_slowness", "vertical_slowness", "pack_layers", "check_table",
           "MAX_LAYERS", "MAX_LAYERS_DERIV"]

# Column order of the packed (n_layers, 6) float64 layer table that the
# C-ABI takes (include/kennett_hip.h).
LAYER_COLUMNS = ("alpha", "beta", "rho", "thickness", "Q_alpha", "Q_beta")

# Layer caps of the kernels (KH_MAX_LAYERS and KH_MAX_LAYERS_DERIV in
# include/kennett_hip.h): forward launches, derivative launches.
MAX_LAYERS = 240
MAX_LAYERS_DERIV = 120


def complex_slowness(velocity: float, Q: float) -> complex:
    """s = (4Q^2 + 2Qi) / ((1 + 4Q^2) v)  (layer_model.py:15-45)."""
    if velocity <= 0:
        raise ValueError(f"velocity must be positive, got {velocity}")
    if Q <= 0:
        raise ValueError(f"Q must be positive, got {Q}")
    two_q = 2.0 * Q
    two_q_sq = two_q * two_q
    denom = (1.0 + two_q_sq) * velocity
    return complex(two_q_sq / denom, two_q / denom)


def vertical_slowness(slowness: complex, p: complex) -> complex:
    """eta = sqrt((s+p)(s-p)), negated when Im(eta) <= 0 (layer_model.py:48-74)."""
    t = (slowness + p) * (slowness - p)
    eta = complex(np.sqrt(np.complex128(t)))
    if eta.imag <= 0.0:
        eta = -eta
    return eta


@dataclass
class LayerModel:
    """Plane-stratified elastic model; layer 0 is the acoustic water layer and
    the last layer is the half-space (thickness ``np.inf``)."""

    alpha: np.ndarray
    beta: np.ndarray
    rho: np.ndarray
    thickness: np.ndarray
    Q_alpha: np.ndarray
    Q_beta: np.ndarray

    def __post_init__(self) -> None:
        for name in LAYER_COLUMNS:
            setattr(self, name, np.asarray(getattr(self, name), dtype=np.float64))
        n = self.n_layers
        for name in LAYER_COLUMNS:
            if len(getattr(self, name)) != n:
                raise ValueError(f"{name} length mismatch: expected {n}")
        if np.any(self.alpha <= 0):
            raise ValueError("All alpha must be positive")
        if np.any(self.beta < 0):
            raise ValueError("All beta must be non-negative")
        if np.any(self.rho <= 0):
            raise ValueError("All rho must be positive")

    @property
    def n_layers(self) -> int:
        return len(self.alpha)

    @classmethod
    def from_arrays(cls, alpha, beta, rho, thickness, Q_alpha, Q_beta) -> "LayerModel":
        return cls(alpha=alpha, beta=beta, rho=rho, thickness=thickness,
                   Q_alpha=Q_alpha, Q_beta=Q_beta)

    def complex_slowness_p(self) -> np.ndarray:
        return np.array([complex_slowness(a, q) for a, q in zip(self.alpha, self.Q_alpha)],
                        dtype=np.complex128)

    def complex_slowness_s(self) -> np.ndarray:
        out = np.zeros(self.n_layers, dtype=np.complex128)
        for i, (b, q) in enumerate(zip(self.beta, self.Q_beta)):
            if b > 0:
                out[i] = complex_slowness(b, q)
        return out

    def complex_velocity_s(self) -> np.ndarray:
        s_s = self.complex_slowness_s()
        out = np.zeros(self.n_layers, dtype=np.complex128)
        nz = np.abs(s_s) > 0
        out[nz] = 1.0 / s_s[nz]
        return out


def check_table(table, max_layers: int = MAX_LAYERS, values: bool = True) -> np.ndarray:
    """The contiguous (n_layers, 6) float64 layer table of ``table`` (rows
    [alpha, beta, rho, h, Q_alpha, Q_beta]), after the checks every kernel
    relies on.  Raises ``ValueError`` for a table of another shape or with
    fewer than two or more than ``max_layers`` layers (the kernels index
    every row as six doubles and stage at most that many layers) and, with
    ``values``, for a NaN anywhere or a value LayerModel or the kernels
    reject (alpha <= 0, beta < 0, rho <= 0, Q_alpha <= 0, Q_beta <= 0 for a
    solid layer)."""
    tab = np.ascontiguousarray(np.asarray(table, dtype=np.float64))
    if tab.ndim != 2 or tab.shape[1] != len(LAYER_COLUMNS):
        raise ValueError(f"layer table must be (n_layers, 6), got {tab.shape}")
    nl = tab.shape[0]
    if nl < 2:
        raise ValueError("model needs at least 2 layers (water + half-space)")
    if nl > max_layers:
        cap = "KH_MAX_LAYERS_DERIV" if max_layers == MAX_LAYERS_DERIV else "KH_MAX_LAYERS"
        raise ValueError(f"model has {nl} layers; the kernels take at most {max_layers} ({cap})")
    if not values:
        return tab
    # fast path: one column-wise minimum (NaN propagates through it) decides
    # almost every table; the messages below are only built for a bad one
    mins = tab.min(axis=0)
    if (mins[0] > 0 and mins[1] >= 0 and mins[2] > 0 and mins[3] == mins[3] and mins[4] > 0
            and mins[5] == mins[5]
            and (mins[5] > 0 or bool(((tab[:, 5] > 0) | (tab[:, 1] <= 0)).all()))):
        return tab
    if np.isnan(tab).any():
        raise ValueError("model contains NaN")
    alpha, beta, rho, _, q_a, q_b = tab.T
    if np.any(alpha <= 0):
        raise ValueError("All alpha must be positive")
    if np.any(beta < 0):
        raise ValueError("All beta must be non-negative")
    if np.any(rho <= 0):
        raise ValueError("All rho must be positive")
    if np.any(q_a <= 0) or np.any((beta > 0) & (q_b <= 0)):
        raise ValueError("Q must be positive")
    return tab


def pack_layers(model: LayerModel, max_layers: int = MAX_LAYERS) -> np.ndarray:
    """(n_layers, 6) float64 table [alpha, beta, rho, h, Q_alpha, Q_beta] per row.

    Raises ``ValueError`` for inputs the kernels cannot take (``check_table``):
    fewer than two or more than ``max_layers`` layers, a non-positive Q, or a
    NaN anywhere."""
    return check_table(np.array([model.alpha, model.beta, model.rho, model.thickness,
                                 model.Q_alpha, model.Q_beta], dtype=np.float64).T.copy(),
                       max_layers)
