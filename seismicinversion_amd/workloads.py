"""Synthetic earth models and the standard benchmark workloads.

The reference ships one hand-written model (``default_ocean_crust_model``,
Kennett_Reflectivity/kennett_seismogram.py:25-54) and no model generator.
The seeded generator below draws layer properties in a fixed order so that
every workload is reproducible from ``(n_layers, seed)``:

* layer 0 is the reference's water layer (vp=1.5 km/s, vs=0, rho=1.0 g/cc,
  h=2.0 km, Qp=2e4, Qs=1e10; kennett_seismogram.py:47-54);
* every lower layer draws, in this order, vp ~ U[1.8, 6.0],
  r ~ U[1.6, 2.0] (vs = vp / r), rho ~ U[1.8, 2.8], h ~ U[0.1, 2.0],
  Qp ~ U[20, 200] (Qs = Qp / 2) from ``numpy.random.default_rng(seed)``;
* the last layer is the elastic half-space (h = inf).

"N-layer" counts the water layer and the half-space, as
``LayerModel.n_layers`` does.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .layer_model import LayerModel

__all__ = [
    "default_ocean_crust_model",
    "seeded_model",
    "uniform_slownesses",
    "omega_grid",
    "GatherConfig",
    "gather_config_c2",
    "uneven_batch_models",
]

WATER = dict(alpha=1.5, beta=0.0, rho=1.0, thickness=2.0, Q_alpha=2.0e4, Q_beta=1.0e10)


def default_ocean_crust_model() -> LayerModel:
    """The reference's 5-layer ocean-crust model (kennett_seismogram.py:25-54)."""
    return LayerModel.from_arrays(
        alpha=[1.5, 1.6, 3.0, 5.0, 2.2],
        beta=[0.0, 0.3, 1.5, 3.0, 1.1],
        rho=[1.0, 2.0, 3.0, 3.0, 1.8],
        thickness=[2.0, 1.0, 1.0, 1.0, np.inf],
        Q_alpha=[20000, 100, 100, 100, 100],
        Q_beta=[1e10, 100, 100, 100, 100],
    )


def seeded_model(n_layers: int, seed: int = 42) -> LayerModel:
    """Water layer over ``n_layers - 1`` random elastic layers (last = half-space)."""
    if n_layers < 2:
        raise ValueError("a model needs the water layer and at least one elastic layer")
    rng = np.random.default_rng(seed)
    cols = {k: [v] for k, v in WATER.items()}
    for i in range(1, n_layers):
        vp = rng.uniform(1.8, 6.0)
        ratio = rng.uniform(1.6, 2.0)
        rho = rng.uniform(1.8, 2.8)
        h = rng.uniform(0.1, 2.0)
        qp = rng.uniform(20.0, 200.0)
        cols["alpha"].append(vp)
        cols["beta"].append(vp / ratio)
        cols["rho"].append(rho)
        cols["thickness"].append(np.inf if i == n_layers - 1 else h)
        cols["Q_alpha"].append(qp)
        cols["Q_beta"].append(qp / 2.0)
    return LayerModel.from_arrays(**cols)


def uniform_slownesses(n_p: int, p_max: float) -> np.ndarray:
    """p_j = j * p_max / n_p for j = 1..n_p (p = 0 excluded), as
    Kennett_Reflectivity/kennett_gather_gpu.py:244-245."""
    dp = p_max / n_p
    return np.arange(1, n_p + 1, dtype=np.float64) * dp


def omega_grid(n_freq: int, T: float = 64.0) -> np.ndarray:
    """omega_k = k * 2*pi/T for k = 1..n_freq (DC excluded),
    as Kennett_Reflectivity/taup_inversion.py:298-301."""
    dw = 2.0 * np.pi / T
    return np.arange(1, n_freq + 1, dtype=np.float64) * dw


@dataclass
class GatherConfig:
    """Arguments of one ``compute_gather_gpu`` call."""

    model: LayerModel
    offsets: np.ndarray
    T: float
    nw: int
    np_slow: int
    p_max: float
    free_surface: bool
    ricker_omega_max: float | None  # None = the reference's default source

    def source_func(self):
        if self.ricker_omega_max is None:
            return None
        from .source import ricker_spectrum

        wmax = float(self.ricker_omega_max)
        return lambda omega, _omega_max: ricker_spectrum(omega, wmax)


def gather_config_c2(nw: int = 1024) -> GatherConfig:
    """20-layer seeded model, free surface on, 8192 slownesses over (0, 1.2] s/km,
    nt = 2*nw samples over T = 64 s, 100 offsets 0.5..50 km, 10 Hz Ricker.

    A 10 Hz Ricker peak with the reference's spectrum (peak at omega = 2 * 0.1 *
    omega_max) needs omega_max = 100*pi rad/s."""
    return GatherConfig(
        model=seeded_model(20, 42),
        offsets=np.arange(0.5, 50.0 + 0.25, 0.5),
        T=64.0,
        nw=nw,
        np_slow=8192,
        p_max=1.2,
        free_surface=True,
        ricker_omega_max=100.0 * np.pi,
    )


def uneven_batch_models(n_models: int = 256, seed: int = 42,
                        nl_range=(3, 60), np_range=(32, 512), p_max: float = 0.8):
    """Models with uneven depth and slowness counts (model i uses seed + i).

    Layer and slowness counts are drawn from ``default_rng(seed)``; returns a
    list of ``(LayerModel, p_samples)``."""
    rng = np.random.default_rng(seed)
    nls = rng.integers(nl_range[0], nl_range[1] + 1, size=n_models)
    nps = rng.integers(np_range[0], np_range[1] + 1, size=n_models)
    return [
        (seeded_model(int(nl), seed + i), uniform_slownesses(int(npp), p_max))
        for i, (nl, npp) in enumerate(zip(nls, nps))
    ]


@dataclass
class StreamItem:
    """One request of the serving stream: arrival offset (s), kind and sizes."""

    t: float
    kind: str
    n_layers: int
    n_p: int
    n_omega: int
    seed: int


def serving_stream(rate: float, n_requests: int, seed: int = 42, lm_fraction: float = 0.2,
                   nl_range=(3, 100), np_range=(16, 2048), omegas=(256, 1024, 4096)):
    """Poisson arrivals at `rate` per second: a fraction `lm_fraction` of
    inversion requests, the rest forward requests; layer and slowness counts
    log-uniform over their ranges, frequency counts uniform over `omegas`."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests)) if rate > 0 else np.zeros(n_requests)
    kinds = np.where(rng.random(n_requests) < lm_fraction, "invert", "forward")
    nls = np.rint(np.exp(rng.uniform(np.log(nl_range[0]), np.log(nl_range[1]), n_requests)))
    nps = np.rint(np.exp(rng.uniform(np.log(np_range[0]), np.log(np_range[1]), n_requests)))
    nws = rng.choice(np.asarray(omegas), n_requests)
    return [StreamItem(float(t[i]), str(kinds[i]), int(nls[i]), int(nps[i]), int(nws[i]),
                       seed + 1000 + i) for i in range(n_requests)]
