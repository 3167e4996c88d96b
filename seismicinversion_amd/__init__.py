"""MI355X-native Kennett recursive P-SV reflectivity.

The public names of Kennett_Reflectivity/__init__.py:29-66 for the
reflectivity path, resolved lazily (the reference does the same,
__init__.py:69-179) so that importing the package does not load the HIP
library or torch until a name that needs them is used.  Every compute entry
point runs on the GPU through libkennett_hip.so; there is no CPU fallback.
"""

from __future__ import annotations

import importlib

__version__ = "0.1.0"

_EXPORTS = {
    "LayerModel": "layer_model",
    "complex_slowness": "layer_model",
    "vertical_slowness": "layer_model",
    "ScatteringCoefficients": "scattering_matrices",
    "solid_solid_interface": "scattering_matrices",
    "ocean_bottom_interface": "scattering_matrices",
    "kennett_reflectivity": "kennett_reflectivity",
    "inv2x2": "kennett_reflectivity",
    "batch_inv2x2": "kennett_reflectivity",
    "ricker_spectrum": "source",
    "ricker_wavelet": "source",
    "compute_gather": "kennett_gather",
    "plot_gather": "kennett_gather",
    "compute_seismogram": "kennett_seismogram",
    "default_ocean_crust_model": "workloads",
    "compute_gather_gpu": "kennett_gather_gpu",
    "kennett_reflectivity_batch": "kennett_reflectivity_gpu",
    "get_device": "kennett_reflectivity_gpu",
    "kennett_reflectivity_torch": "kennett_torch",
    "forward_model_torch": "kennett_torch",
    "jacobian": "kennett_torch",
    "hessian": "kennett_torch",
    "model_to_tensors": "kennett_torch",
    "frechet_kennett": "frechet_analytical",
    "InversionResult": "taup_inversion",
    "compute_taup_traces": "taup_inversion",
    "invert_taup": "taup_inversion",
    "taup_misfit_grad_hessian": "taup_inversion",
    "load_config": "inversion_config",
    "save_config": "inversion_config",
    "InversionConfig": "inversion_config",
    "gmm_reflectivity": "global_matrix",
    "gmm_reflectivity_batch": "global_matrix",
    "reflectivity_multi": "batch",
    "ModelSpec": "batch",
    "KennettServer": "serve",
    "InvertSpec": "serve",
}

__all__ = sorted(_EXPORTS)


def __getattr__(name: str):
    mod = _EXPORTS.get(name)
    if mod is None:
        raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
    return getattr(importlib.import_module(f".{mod}", __name__), name)
