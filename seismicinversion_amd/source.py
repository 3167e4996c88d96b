"""Source wavelets (drop-in for Kennett_Reflectivity/source.py).

``ricker_spectrum`` (source.py:10-45) is evaluated on the host: it is a
handful of operations per frequency and callers may pass any Python
``source_func`` in its place, so the gather takes S(omega) as an input array.
"""

from __future__ import annotations

import numpy as np

__all__ = ["ricker_spectrum", "ricker_wavelet"]


def ricker_spectrum(omega: np.ndarray, omega_max: float) -> np.ndarray:
    """S = Z exp(-Z + i w T0), Z = w^2 / (4 a^2), a = 0.1 w_max, T0 = 5 / (a sqrt 2)."""
    a = 0.1 * omega_max
    t0 = 5.0 / (a * np.sqrt(2.0))
    z = (omega ** 2) / (4.0 * a ** 2)
    return z * np.exp(-z + 1j * omega * t0)


def ricker_wavelet(t: np.ndarray, f_peak: float) -> np.ndarray:
    """(1 - 2 tau^2) exp(-tau^2), tau = pi f (t - 1/(pi f))  (source.py:48-80)."""
    t0 = 1.0 / (np.pi * f_peak)
    tau = np.pi * f_peak * (t - t0)
    return (1.0 - 2.0 * tau ** 2) * np.exp(-(tau ** 2))
