"""Drop-in for Kennett_Reflectivity/scattering_matrices.py.

``solid_solid_interface`` and ``ocean_bottom_interface`` keep the reference
signatures (scattering_matrices.py:42-298) and return a
``ScatteringCoefficients`` of 2x2 complex128 NumPy arrays; the coefficients
are evaluated by the device kernel (kh_interfaces_direct) with the same
formulas the reflectivity kernels stage in LDS.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import engine

__all__ = ["ScatteringCoefficients", "solid_solid_interface", "ocean_bottom_interface"]


@dataclass
class ScatteringCoefficients:
    """2x2 P-SV coefficients at an interface (index 0 = P, 1 = S)."""

    Rd: np.ndarray
    Ru: np.ndarray
    Tu: np.ndarray
    Td: np.ndarray

    def __post_init__(self):
        for name in ("Rd", "Ru", "Tu", "Td"):
            arr = np.asarray(getattr(self, name), dtype=np.complex128)
            if arr.shape != (2, 2):
                raise ValueError(f"{name} must be 2x2, got {arr.shape}")
            setattr(self, name, arr)


def _pack(outs):
    Rd, Ru, Tu, Td = (o.cpu().numpy()[0] for o in outs)
    return ScatteringCoefficients(Rd=Rd, Ru=Ru, Tu=Tu, Td=Td)


def solid_solid_interface(p, eta1, neta1, rho1, beta1, eta2, neta2, rho2,
                          beta2) -> ScatteringCoefficients:
    """Modified (energy-normalised) coefficients of a solid-solid interface."""
    return _pack(engine.interfaces_direct(p, eta1, neta1, rho1, beta1, eta2, neta2, rho2, beta2,
                                          ocean=False))


def ocean_bottom_interface(p, eta1, rho1, eta2, neta2, rho2, beta2) -> ScatteringCoefficients:
    """Coefficients of the acoustic (water) over elastic interface."""
    return _pack(engine.interfaces_direct(p, eta1, None, rho1, None, eta2, neta2, rho2, beta2,
                                          ocean=True))
