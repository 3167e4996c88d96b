"""Drop-in for Kennett_Reflectivity/kennett_gather_gpu.py.

``compute_gather_gpu`` keeps the reference signature and return value
(kennett_gather_gpu.py:169-215).  All three stages run on the device in
float64/complex128: the reflectivity on the damped frequency grid
(kh_forward_grid), the Bessel-weighted slowness sum (kh_bessel_sum) and the
Hermitian inverse FFT with the exp(-gamma t) factor (kh_hermitian_traces).
The reference instead runs the recursion in NumPy, evaluates J0 on the host
and sums in float32 on the device.

The trace is multiplied by exp(-gamma t), as the reference does
(kennett_gather_gpu.py:275,282); undoing the exp(-i w t)-convention damping
would take exp(+gamma t).  The factor is kept for parity.
"""

from __future__ import annotations

import collections.abc

import numpy as np
import torch

from . import _abi, engine
from .layer_model import LayerModel
from .source import ricker_spectrum
from .workloads import default_ocean_crust_model

__all__ = ["compute_gather_gpu", "default_ocean_crust_model", "gather_device", "main"]


def gather_device(layers: torch.Tensor, offsets: torch.Tensor, T: float, nw: int, np_slow: int,
                  p_max: float, gamma: float, S: torch.Tensor | None, free_surface: bool,
                  taper_p_max: float | None = None, keep: bool = False):
    """Device pipeline; returns traces (nr, 2*nw) float64 on the device, and
    with ``keep`` also (R, U)."""
    dev = layers.device
    dw = 2.0 * np.pi / T
    dp = p_max / np_slow
    p = torch.arange(1, np_slow + 1, dtype=torch.float64, device=dev) * dp
    nwm = nw - 1
    R = engine.forward_grid(layers, p, 1.0, dw, gamma, nwm, free_surface=free_surface)
    if taper_p_max is None:
        taper_p_max = float(np_slow * dp)  # p_samples[-1], as kennett_gather_gpu.py:100
    U = engine.bessel_sum(R, p, dp, taper_p_max, 1.0, dw, offsets, S)
    traces = engine.hermitian_traces(U, nw, gamma, T)
    return (traces, R, U) if keep else traces


def _gather(model, offsets, T, nw, np_slow, p_max, gamma, source_func, free_surface, device,
            taper_p_max):
    dev = _abi.device(device)
    offsets = np.asarray(offsets, dtype=np.float64)
    if gamma is None:
        gamma = np.pi / T
    dw = 2.0 * np.pi / T
    omega_real = np.arange(1, nw, dtype=np.float64) * dw
    wmax = nw * dw
    S = ricker_spectrum(omega_real, wmax) if source_func is None else source_func(omega_real, wmax)
    S_t = torch.from_numpy(np.ascontiguousarray(np.asarray(S, dtype=np.complex128))).to(dev)
    layers = engine.layers_tensor(model, dev)
    off_t = torch.from_numpy(np.ascontiguousarray(offsets)).to(dev)
    traces = gather_device(layers, off_t, T, nw, np_slow, p_max, gamma, S_t, free_surface,
                           taper_p_max=taper_p_max)
    nt = 2 * nw
    time = np.arange(nt, dtype=np.float64) * (T / float(nt))
    return time, offsets, traces.cpu().numpy()


def compute_gather_gpu(
    model: LayerModel,
    offsets: np.ndarray,
    T: float = 64.0,
    nw: int = 2048,
    np_slow: int = 4096,
    p_max: float = 1.2,
    gamma: float | None = None,
    source_func: collections.abc.Callable | None = None,
    free_surface: bool = False,
    device: torch.device | None = None,
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(time (nt,), offsets (nr,), gather (nr, nt)) with nt = 2*nw."""
    return _gather(model, offsets, T, nw, np_slow, p_max, gamma, source_func, free_surface,
                   device, taper_p_max=None)


def main(argv: list[str] | None = None) -> None:
    """CLI of kennett_gather_gpu.py (defaults: offsets to 50 km, 4096 slownesses up to 1.2 s/km)."""
    from .kennett_gather import gather_cli

    gather_cli(argv, description="GPU seismogram gather via Kennett reflectivity.",
               r_max=50.0, np_slow=4096, p_max=1.2, t_max=30.0, output="gather_gpu.png",
               workers_flag=False,
               compute=lambda m, off, a: compute_gather_gpu(
                   m, off, T=a.duration, nw=a.nw, np_slow=a.np_slow, p_max=a.p_max,
                   gamma=a.gamma, free_surface=a.free_surface,
                   device=torch.device(a.device) if a.device else None))


if __name__ == "__main__":
    main()
