"""YAML configuration of a tau-p inversion run (drop-in for
Kennett_Reflectivity/inversion_config.py:1-346).

``load_config`` reads and validates a YAML file into an ``InversionConfig``
(same sections, keys, defaults and validation rules as the reference:
inversion_config.py:174-287); ``save_config`` writes one back so that
loading it reproduces the run (:290-346).  Invalid files raise
``ConfigError`` naming the offending key.

File layout::

    model:
      layers:            # >= 2; layer 0 acoustic (beta 0); last thickness .inf
        - {name: Ocean, alpha: 1.5, beta: 0.0, rho: 1.0, thickness: 2.0,
           Q_alpha: 20000, Q_beta: 1.0e+10}
        - ...
      fixed_layers: [0]  # only the ocean layer can be fixed
    inversion: {p_values: [...], nfreq: 64, perturbation: 0.15, max_iter: 100,
                seed: 42, tol: 1.0e-8}
    output: {directory: figures, formats: [table, profiles, traces, convergence],
             trace_display: {t_max: 15.0, nw: 1024}}
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import yaml

from .layer_model import LayerModel

__all__ = ["InversionConfig", "OutputConfig", "TraceDisplayConfig", "ConfigError",
           "load_config", "save_config", "DEFAULT_CONFIG"]

FORMATS = ("table", "profiles", "traces", "convergence")
LAYER_KEYS = ("alpha", "beta", "rho", "thickness", "Q_alpha", "Q_beta")
DEFAULT_P = [0.05, 0.10, 0.15, 0.20, 0.25, 0.30, 0.35, 0.40, 0.45, 0.50, 0.55, 0.60]
DEFAULT_CONFIG = Path(__file__).resolve().parent / "configs" / "default_ocean_crust.yaml"


@dataclass
class TraceDisplayConfig:
    t_max: float = 15.0
    nw: int = 1024


@dataclass
class OutputConfig:
    directory: Path = field(default_factory=lambda: Path("figures"))
    formats: list[str] = field(default_factory=lambda: list(FORMATS))
    trace_display: TraceDisplayConfig = field(default_factory=TraceDisplayConfig)


@dataclass
class InversionConfig:
    true_model: LayerModel
    layer_names: list[str]
    fixed_layers: list[int]
    p_values: list[float]
    nfreq: int = 64
    perturbation: float = 0.15
    max_iter: int = 100
    seed: int = 42
    tol: float = 1e-8
    output: OutputConfig = field(default_factory=OutputConfig)


class ConfigError(Exception):
    """An invalid inversion config file."""


def _check(ok: bool, msg: str) -> None:
    if not ok:
        raise ConfigError(msg)


def _check_layers(layers) -> None:
    _check(isinstance(layers, list) and len(layers) >= 2,
           f"model.layers must list at least 2 layers (finite layers over a half-space), "
           f"got {len(layers) if isinstance(layers, list) else type(layers).__name__}")
    for i, lay in enumerate(layers):
        where = f"model.layers[{i}]"
        _check(isinstance(lay, dict), f"{where} must be a mapping")
        for key in LAYER_KEYS:
            _check(key in lay, f"{where} is missing '{key}' (every layer needs "
                               f"{', '.join(LAYER_KEYS)})")
        for key, lower_ok in (("alpha", False), ("beta", True), ("rho", False),
                              ("Q_alpha", False), ("Q_beta", False)):
            v = float(lay[key])
            _check(v >= 0 if lower_ok else v > 0,
                   f"{where}.{key} must be {'>= 0' if lower_ok else '> 0'}, got {lay[key]}")
    _check(float(layers[0]["beta"]) == 0.0,
           f"model.layers[0].beta must be 0 (the top layer is the acoustic ocean), "
           f"got {layers[0]['beta']}")
    _check(math.isinf(float(layers[-1]["thickness"])),
           f"model.layers[{len(layers) - 1}].thickness must be .inf (the half-space), "
           f"got {layers[-1]['thickness']}")


def load_config(path) -> InversionConfig:
    """Read and validate an inversion config (inversion_config.py:174-287)."""
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"Config file not found: {path.resolve()}")
    raw = yaml.safe_load(path.read_text(encoding="utf-8"))
    _check(isinstance(raw, dict) and "model" in raw,
           "the config needs a top-level 'model' section with a 'layers' list")
    model_sec = raw["model"]
    _check(isinstance(model_sec, dict) and "layers" in model_sec,
           "the 'model' section needs a 'layers' list")
    layers = model_sec["layers"]
    _check_layers(layers)
    fixed = model_sec.get("fixed_layers", [0])
    _check(fixed == [0], f"model.fixed_layers must be [0] (only the ocean layer can be "
                         f"fixed), got {fixed}")
    cols = {k: [float(lay[k]) for lay in layers] for k in LAYER_KEYS}
    true_model = LayerModel.from_arrays(**cols)
    names = [str(lay.get("name", f"Layer {i}")) for i, lay in enumerate(layers)]

    inv = raw.get("inversion") or {}
    p_values = inv.get("p_values", list(DEFAULT_P))
    _check(isinstance(p_values, list) and len(p_values) > 0,
           "inversion.p_values must be a non-empty list")
    for i, pv in enumerate(p_values):
        _check(float(pv) > 0, f"inversion.p_values[{i}] must be > 0, got {pv}")

    out = raw.get("output") or {}
    formats = out.get("formats", list(FORMATS))
    for fmt in formats:
        _check(fmt in FORMATS, f"output.formats has unknown format '{fmt}' "
                               f"(known: {', '.join(sorted(FORMATS))})")
    td = out.get("trace_display") or {}
    return InversionConfig(
        true_model=true_model, layer_names=names, fixed_layers=list(fixed),
        p_values=[float(pv) for pv in p_values], nfreq=int(inv.get("nfreq", 64)),
        perturbation=float(inv.get("perturbation", 0.15)),
        max_iter=int(inv.get("max_iter", 100)), seed=int(inv.get("seed", 42)),
        tol=float(inv.get("tol", 1e-8)),
        output=OutputConfig(directory=Path(out.get("directory", "figures")),
                            formats=list(formats),
                            trace_display=TraceDisplayConfig(t_max=float(td.get("t_max", 15.0)),
                                                             nw=int(td.get("nw", 1024)))))


class _Dumper(yaml.SafeDumper):
    pass


def _float(dumper, value: float):
    if math.isinf(value):
        return dumper.represent_scalar("tag:yaml.org,2002:float", ".inf" if value > 0 else "-.inf")
    return dumper.represent_scalar("tag:yaml.org,2002:float", repr(float(value)))


_Dumper.add_representer(float, _float)


def save_config(config: InversionConfig, path) -> None:
    """Write ``config`` as YAML that ``load_config`` reads back to the same run
    (inversion_config.py:290-346); the half-space thickness is written .inf."""
    m = config.true_model
    layers = []
    for i in range(m.n_layers):
        lay = {"name": config.layer_names[i]}
        for key in LAYER_KEYS:
            lay[key] = float(getattr(m, key)[i])
        layers.append(lay)
    doc = {
        "model": {"layers": layers, "fixed_layers": list(config.fixed_layers)},
        "inversion": {"p_values": [float(pv) for pv in config.p_values],
                      "nfreq": int(config.nfreq), "perturbation": float(config.perturbation),
                      "max_iter": int(config.max_iter), "seed": int(config.seed),
                      "tol": float(config.tol)},
        "output": {"directory": str(config.output.directory),
                   "formats": list(config.output.formats),
                   "trace_display": {"t_max": float(config.output.trace_display.t_max),
                                     "nw": int(config.output.trace_display.nw)}},
    }
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(yaml.dump(doc, Dumper=_Dumper, default_flow_style=False, sort_keys=False),
                    encoding="utf-8")
