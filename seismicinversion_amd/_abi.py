"""ctypes binding of libkennett_hip.so (declarations: include/kennett_hip.h).

torch is imported first on purpose: the library then binds to the HIP
runtime torch already loaded, so device pointers and streams are shared.
There is no fallback: if the library is missing, or no GPU is visible, every
entry point raises.
"""

from __future__ import annotations

import ctypes
import os

import torch

from ._build import LIB_PATH

__all__ = ["lib", "check", "device", "stream_handle", "ptr", "EXPORTS"]

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_vp = ctypes.c_void_p

# name -> argtypes (every entry point returns int unless listed in _RESTYPES)
EXPORTS: dict[str, list] = {
    "kh_version": [],
    "kh_last_error": [],
    "kh_vertical_slowness": [_vp, _c_int, _vp, _c_int, _vp, _vp, _vp],
    "kh_interface_coefficients": [_vp, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _vp],
    "kh_forward": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp, _vp],
    "kh_forward_ex": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp, _c_int, _c_int, _vp],
    "kh_forward_grid": [_vp, _c_int, _vp, _c_int, _c_double, _c_double, _c_double, _c_int, _c_int,
                        _vp, _vp],
    "kh_forward_grid_ex": [_vp, _c_int, _vp, _c_int, _c_double, _c_double, _c_double, _c_int,
                           _c_int, _vp, _c_int, _c_int, _vp],
    "kh_bessel_workspace_bytes": [_c_int, _c_int, _c_int],
    "kh_bessel_sum": [_vp, _c_int, _c_int, _vp, _c_double, _c_double, _c_double, _c_double, _vp,
                      _c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp],
    "kh_hermitian_traces": [_vp, _c_int, _c_int, _c_double, _c_double, _vp, _vp],
    "kh_deriv_workspace_bytes": [_c_int, _c_int, _c_int, _c_int, _c_int],
    "kh_deriv_workspace_bytes_min": [_c_int, _c_int, _c_int, _c_int, _c_int],
    "kh_jvp": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp, _vp, _vp],
    "kh_vjp": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp,
               ctypes.c_size_t, _vp],
    "kh_misfit_grad": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp,
                       ctypes.c_size_t, _vp],
    "kh_misfit_hvp": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp, _c_int,
                      _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp],
    "kh_misfit_hvp_partial": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp,
                              _c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp],
    "kh_vjp_hvp": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp, _vp, _vp,
                   _vp, _vp, _vp, ctypes.c_size_t, _vp],
    "kh_interfaces_direct": [_c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                             _vp, _vp, _vp, _vp],
    "kh_gram_workspace_bytes": [_c_int, ctypes.c_long],
    "kh_gram": [_vp, _c_int, ctypes.c_long, _c_double, _vp, _vp, ctypes.c_size_t, _vp],
    "kh_residual_norms": [_vp, _vp, _c_int, _c_int, _vp, _vp],
    "kh_gmm_forward": [_vp, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp, _vp],
    "kh_gmm_forward_grid": [_vp, _c_int, _vp, _c_int, _c_double, _c_double, _c_double, _c_int,
                            _c_int, _vp, _vp],
    "kh_multi_tiles": [_c_int, _c_int, _c_int, _c_int],
    "kh_multi_items": [_c_int, _c_int, _c_int],
    "kh_multi_workgroups": [_c_int, _c_int, _c_int, _c_int, _vp],
    "kh_multi_table_bytes": [ctypes.c_longlong],
    "kh_forward_multi": [_vp, _c_int, ctypes.c_longlong, _c_int, _vp, _vp, _vp, _c_int, _c_int,
                         _c_int, ctypes.c_longlong, _vp, ctypes.c_size_t, _vp, _vp, _vp],
    "kh_forward_multi_static": [_vp, _c_int, ctypes.c_longlong, _c_int, _vp, _vp, _vp, _c_int,
                                _c_int, _c_int, ctypes.c_longlong, _c_int, _vp, _vp],
    "kh_lm_tables": [_vp, _c_int, _vp, _c_int, _vp, _vp, _vp],
    "kh_lm_assemble": [_c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "kh_lm_prepare": [_c_int, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _c_double, _c_int, _vp,
                      _vp, _vp, _vp, _vp],
    "kh_lm_trials_workspace_bytes": [_c_int, _c_int],
    "kh_lm_trials": [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                     ctypes.c_size_t, _vp],
    "kh_lm_trial_misfits": [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _vp,
                            _vp, _vp, _vp],
    "kh_lm_decide": [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp],
}
_RESTYPES = {"kh_multi_items": ctypes.c_longlong, "kh_multi_table_bytes": ctypes.c_size_t,
             "kh_deriv_workspace_bytes": ctypes.c_size_t,
             "kh_deriv_workspace_bytes_min": ctypes.c_size_t,
             "kh_gram_workspace_bytes": ctypes.c_size_t,
             "kh_lm_trials_workspace_bytes": ctypes.c_size_t,
             "kh_version": ctypes.c_char_p, "kh_last_error": ctypes.c_char_p,
             "kh_bessel_workspace_bytes": ctypes.c_size_t}

KH_ERR_INVALID_ARGUMENT = 1

_LIB = None


def lib():
    """Load (once) and return the library handle."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m seismicinversion_amd._build` "
                "(or __graft_entry__.build())")
        handle = ctypes.CDLL(LIB_PATH)
        for name, argtypes in EXPORTS.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, _c_int)
        _LIB = handle
    return _LIB


def check(rc: int, name: str) -> None:
    """Raise ValueError (invalid argument) or RuntimeError for a non-zero status."""
    if rc == 0:
        return
    msg = lib().kh_last_error().decode(errors="replace")
    if rc == KH_ERR_INVALID_ARGUMENT:
        raise ValueError(f"{name}: {msg}")
    raise RuntimeError(f"{name} failed (status {rc}): {msg}")


def device(dev=None) -> torch.device:
    """The HIP device to run on; raises if there is none (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("seismicinversion_amd needs a ROCm GPU: no HIP device is visible")
    if dev is None:
        return torch.device("cuda", torch.cuda.current_device())
    dev = torch.device(dev)
    if dev.type != "cuda":
        raise RuntimeError(f"seismicinversion_amd runs on HIP devices, not {dev}")
    return dev


def stream_handle(dev: torch.device) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def ptr(t: torch.Tensor) -> int:
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()
