"""Drop-in for Kennett_Reflectivity/kennett_seismogram.py.

``compute_seismogram`` keeps the reference signature and return value
(kennett_seismogram.py:57-150): the vertical-displacement seismogram of one
slowness, R(omega) * S(omega) on omega_k = k * 2*pi/T (k = 1..nw-1, real
frequencies), as a length 2*nw real trace through the Hermitian FFT,
multiplied by exp(+eps t) (kennett_seismogram.py:143, the opposite sign of
the gather's exp(-gamma t)).  On the device: the reflectivity is one
slowness row of kh_forward_grid, the trace one row of kh_hermitian_traces
(given -eps as its damping rate).

``main`` is the reference CLI (kennett_seismogram.py:153-300) with the same
flags: it writes ``seismogram_p{p:.4f}.txt`` (time, amplitude) and, unless
--no-plot, a two-panel wiggle plot.
"""

from __future__ import annotations

import collections.abc
import logging

import numpy as np
import torch

from . import _abi, engine
from .layer_model import LayerModel
from .source import ricker_spectrum
from .workloads import default_ocean_crust_model

__all__ = ["compute_seismogram", "default_ocean_crust_model"]

logger = logging.getLogger(__name__)


def compute_seismogram(
    model: LayerModel,
    p: float,
    T: float = 64.0,
    nw: int = 2048,
    eps: float = 0.0,
    source_func: collections.abc.Callable | None = None,
    free_surface: bool = False,
) -> tuple[np.ndarray, np.ndarray]:
    """(time (2*nw,), seismogram (2*nw,)) for slowness p; nw a power of two."""
    if nw < 1 or nw & (nw - 1):
        raise ValueError(f"nw={nw} must be a power of two")
    dev = _abi.device()
    dw = 2.0 * np.pi / T
    nwm = nw - 1
    omega = np.arange(1, nwm + 1, dtype=np.float64) * dw
    wmax = nw * dw
    S = ricker_spectrum(omega, wmax) if source_func is None else source_func(omega, wmax)
    S_t = torch.from_numpy(np.ascontiguousarray(np.asarray(S, dtype=np.complex128))).to(dev)
    layers = engine.layers_tensor(model, dev)
    p_t = torch.tensor([float(p)], dtype=torch.float64, device=dev)
    R = engine.forward_grid(layers, p_t, 1.0, dw, 0.0, nwm, free_surface=free_surface)
    trace = engine.hermitian_traces(R * S_t, nw, -float(eps), T)
    nt = 2 * nw
    time = np.arange(nt, dtype=np.float64) * (T / float(nt))
    return time, trace[0].cpu().numpy()


def _plot(time, seismogram, slowness, model, output):
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    fig, axes = plt.subplots(2, 1, figsize=(14, 10))
    t_ocean = 2.0 * model.thickness[0] / model.alpha[0]
    for ax, mask, title in ((axes[0], np.ones_like(time, bool),
                             f"Kennett reflectivity, {model.n_layers}-layer model (p={slowness} s/km)"),
                            (axes[1], time <= 15.0, "Early arrivals (0-15 s)")):
        ax.plot(time[mask], seismogram[mask], "k-", linewidth=0.5)
        ax.fill_between(time[mask], 0, seismogram[mask], where=seismogram[mask] > 0,
                        interpolate=True, color="black", alpha=0.9)
        ax.axhline(0.0, color="k", linewidth=0.3)
        ax.set_xlabel("Time (s)")
        ax.set_ylabel("Amplitude")
        ax.set_title(title)
    axes[1].axvline(t_ocean, color="r", linestyle="--", alpha=0.6,
                    label=r"$2h_{\mathrm{ocean}}/\alpha_{\mathrm{ocean}}$")
    axes[1].legend(fontsize=10)
    fig.tight_layout()
    fig.savefig(output, dpi=150, bbox_inches="tight")
    plt.close(fig)


def main(argv: list[str] | None = None) -> None:
    """CLI: the reference's flags (kennett_seismogram.py:153-198)."""
    import argparse

    ap = argparse.ArgumentParser(
        description="Synthetic seismogram by Kennett reflectivity on the GPU.")
    ap.add_argument("-p", "--slowness", type=float, default=0.2,
                    help="Horizontal slowness / ray parameter in s/km (default: 0.2)")
    ap.add_argument("-T", "--duration", type=float, default=64.0,
                    help="Time window in seconds (default: 64.0)")
    ap.add_argument("-n", "--nw", type=int, default=2048,
                    help="Number of positive frequencies (default: 2048)")
    ap.add_argument("-o", "--output", type=str, default="seismogram_verification.png",
                    help="Output plot filename (default: seismogram_verification.png)")
    ap.add_argument("--no-plot", action="store_true", help="Skip plot generation, save data only")
    ap.add_argument("--free-surface", action="store_true",
                    help="Include free surface reflections (water-column multiples)")
    args = ap.parse_args(argv)
    logging.basicConfig(level=logging.INFO,
                        format="%(asctime)s - %(name)s - %(levelname)s - %(message)s")
    model = default_ocean_crust_model()
    time, seis = compute_seismogram(model, args.slowness, T=args.duration, nw=args.nw,
                                    free_surface=args.free_surface)
    if not args.no_plot:
        try:
            _plot(time, seis, args.slowness, model, args.output)
            logger.info("Plot saved to %s", args.output)
        except ImportError:
            logger.warning("matplotlib not available, skipping plot")
    out = f"seismogram_p{args.slowness:.4f}.txt"
    np.savetxt(out, np.column_stack([time, seis]), fmt="%.8e")
    logger.info("Data saved to %s", out)


if __name__ == "__main__":
    main()
