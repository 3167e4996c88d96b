"""Drop-in for the float64 gather of Kennett_Reflectivity/kennett_gather.py.

``compute_gather`` (kennett_gather.py:127-305) keeps its signature; the
reference spreads slowness batches over a multiprocessing pool, here the
whole sum is one device pass (``n_workers`` is accepted and ignored).  It
differs from ``compute_gather_gpu`` only in its defaults and in tapering with
the ``p_max`` argument instead of the last slowness (kennett_gather.py:228).
"""

from __future__ import annotations

import collections.abc
import logging
from pathlib import Path

import numpy as np

from .kennett_gather_gpu import _gather
from .layer_model import LayerModel
from .workloads import default_ocean_crust_model

__all__ = ["compute_gather", "default_ocean_crust_model", "plot_gather", "main"]


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


logger = logging.getLogger(__name__)


def plot_gather(time: np.ndarray, offsets: np.ndarray, gather: np.ndarray, output,
                title: str = "Synthetic Seismogram Gather", t_max: float | None = None,
                clip: float = 0.8, norm: str = "global", noise_floor_db: float = -60.0) -> None:
    """Variable-area wiggle plot of a gather, one trace per offset (the
    parameters of kennett_gather.py:308-407).  ``norm='global'`` scales every
    trace by the gather's peak (true relative amplitudes) and blanks traces
    whose peak lies ``noise_floor_db`` below it; ``norm='trace'`` scales each
    trace by its own peak.  ``clip`` is the peak excursion as a fraction of
    the offset spacing."""
    if norm not in ("global", "trace"):
        raise ValueError(f"norm must be 'global' or 'trace', got {norm!r}")
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    time = np.asarray(time, dtype=np.float64)
    offsets = np.asarray(offsets, dtype=np.float64)
    gather = np.asarray(gather, dtype=np.float64)
    spacing = offsets[1] - offsets[0] if offsets.size > 1 else 1.0
    keep = time <= t_max if t_max is not None else np.ones(time.shape, bool)
    t = time[keep]
    g = gather[:, keep]
    peaks = np.abs(g).max(axis=1) if g.size else np.zeros(len(offsets))
    if norm == "global":
        top = float(peaks.max()) if peaks.size else 0.0
        scale = np.full(len(offsets), spacing * clip / top if top > 0 else 0.0)
        scale[peaks < top * 10.0 ** (noise_floor_db / 20.0)] = 0.0
    else:
        scale = np.where(peaks > 0, spacing * clip / np.where(peaks > 0, peaks, 1.0), 0.0)
    fig, ax = plt.subplots(1, 1, figsize=(12, 14))
    for x0, tr, sc in zip(offsets, g, scale):
        w = tr * sc
        ax.plot(x0 + w, t, "k-", linewidth=0.4)
        ax.fill_betweenx(t, x0, x0 + w, where=w > 0, interpolate=True, color="black", alpha=0.9)
    ax.set_xlabel("Offset (km)")
    ax.set_ylabel("Time (s)")
    ax.set_title(title)
    ax.invert_yaxis()
    if offsets.size:
        ax.set_xlim(offsets[0] - spacing, offsets[-1] + spacing)
    fig.tight_layout()
    fig.savefig(str(output), dpi=150, bbox_inches="tight")
    plt.close(fig)
    logger.info("Gather plot saved to %s", output)


def gather_cli(argv, *, description: str, r_max: float, np_slow: int, p_max: float,
               t_max: float | None, output: str, workers_flag: bool, compute) -> None:
    """The gather CLIs of kennett_gather.py:410-555 and kennett_gather_gpu.py:286-436:
    the same flags and defaults, the default 5-layer model, a wiggle plot
    unless --no-plot, and <output>.npz with time, offsets and gather."""
    import argparse
    import time as time_module

    ap = argparse.ArgumentParser(description=description)
    ap.add_argument("--r-min", type=float, default=0.5, help="Minimum offset in km (default: 0.5)")
    ap.add_argument("--r-max", type=float, default=r_max,
                    help=f"Maximum offset in km (default: {r_max})")
    ap.add_argument("--dr", type=float, default=0.5, help="Offset spacing in km (default: 0.5)")
    ap.add_argument("-T", "--duration", type=float, default=64.0,
                    help="Time window in seconds (default: 64.0)")
    ap.add_argument("-n", "--nw", type=int, default=2048,
                    help="Number of positive frequencies (default: 2048)")
    ap.add_argument("--np", type=int, default=np_slow, dest="np_slow",
                    help=f"Number of slowness samples (default: {np_slow})")
    ap.add_argument("--p-max", type=float, default=p_max,
                    help=f"Maximum slowness in s/km (default: {p_max})")
    ap.add_argument("--gamma", type=float, default=None,
                    help="Complex frequency damping gamma in rad/s (default: pi/T)")
    ap.add_argument("--t-max", type=float, default=t_max,
                    help=f"Maximum time to display in seconds (default: {t_max})")
    ap.add_argument("-o", "--output", type=str, default=output,
                    help=f"Output plot filename (default: {output})")
    ap.add_argument("--no-plot", action="store_true", help="Skip plot generation, save data only")
    if workers_flag:
        ap.add_argument("-j", "--workers", type=int, default=None,
                        help="Accepted for compatibility; the sum runs on the GPU in one pass")
    else:
        ap.add_argument("--device", type=str, default=None,
                        help="Device (default: the first ROCm GPU)")
    ap.add_argument("--free-surface", action="store_true",
                    help="Include free surface reflections (water-column multiples)")
    args = ap.parse_args(argv)
    logging.basicConfig(level=logging.INFO,
                        format="%(asctime)s - %(name)s - %(levelname)s - %(message)s")
    model = default_ocean_crust_model()
    offsets = np.arange(args.r_min, args.r_max + args.dr / 2, args.dr)
    logger.info("Model: %d layers, %d offsets [%.1f-%.1f km]", model.n_layers, len(offsets),
                offsets[0], offsets[-1])
    t0 = time_module.perf_counter()
    t, off, g = compute(model, offsets, args)
    logger.info("Total computation time: %.3fs", time_module.perf_counter() - t0)
    if not args.no_plot:
        try:
            fs = " + free surface" if args.free_surface else ""
            plot_gather(t, off, g, output=args.output, t_max=args.t_max,
                        title=(f"Kennett reflectivity gather, {model.n_layers}-layer ocean crust{fs}\n"
                               f"discrete wavenumber sum (np={args.np_slow}, p_max={args.p_max} s/km)"))
        except ImportError:
            logger.warning("matplotlib not available, skipping plot")
    out = Path(args.output).with_suffix(".npz")
    np.savez(out, time=t, offsets=off, gather=g)
    logger.info("Data saved to %s", out)


def main(argv: list[str] | None = None) -> None:
    """CLI of kennett_gather.py (defaults: offsets to 20 km, 2048 slownesses up to 0.8 s/km)."""
    gather_cli(argv, description="Seismogram gather via Kennett reflectivity (float64).",
               r_max=20.0, np_slow=2048, p_max=0.8, t_max=None, output="seismogram_gather.png",
               workers_flag=True,
               compute=lambda m, off, a: compute_gather(
                   m, off, T=a.duration, nw=a.nw, np_slow=a.np_slow, p_max=a.p_max,
                   gamma=a.gamma, n_workers=a.workers, free_surface=a.free_surface))


if __name__ == "__main__":
    main()
