"""Inversion config I/O (inversion_config.py:174-346) and the report writers
of the inversion CLI, on the CPU.  The reference's config schema
(configs/default_ocean_crust.yaml: model.layers with per-layer name and the
six parameters, model.fixed_layers, the inversion section, output with
trace_display) is exercised with a file the test writes in that schema."""

from __future__ import annotations

import os

import numpy as np
import pytest

from seismicinversion_amd.inversion_config import (DEFAULT_CONFIG, ConfigError, load_config,
                                                   save_config)
from seismicinversion_amd.taup_inversion import InversionResult
from seismicinversion_amd.taup_report import (depth_profile, plot_convergence_curves,
                                              write_model_profiles_tikz, write_model_table_latex)
from seismicinversion_amd.workloads import default_ocean_crust_model



def same_model(a, b):
    for k in ("alpha", "beta", "rho", "thickness", "Q_alpha", "Q_beta"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k))


def _reference_schema_yaml(path):
    """The 5-layer ocean-crust run in the reference's config schema, with
    numbers written the way a hand-edited file writes them (ints for Q,
    exponent forms, .inf), as flow-style YAML."""
    import yaml

    layers = [("Ocean", 1.5, 0.0, 1.0, 2.0, 20000, "1.0e+10"),
              ("Sediment", 1.6, 0.3, 2.0, 1.0, 100, 100),
              ("Crust", 3.0, 1.5, 3.0, 1.0, 100, 100),
              ("Upper mantle", 5.0, 3.0, 3.0, 1.0, 100, 100),
              ("Half-space", 2.2, 1.1, 1.8, ".inf", 100, 100)]
    text = "model:\n  layers:\n"
    for name, a, b, r, h, qa, qb in layers:
        text += (f"    - {{name: {name}, alpha: {a}, beta: {b}, rho: {r}, thickness: {h}, "
                 f"Q_alpha: {qa}, Q_beta: {qb}}}\n")
    text += ("  fixed_layers: [0]\n"
             "inversion: {p_values: [0.05, 0.10, 0.15, 0.20, 0.25, 0.30, 0.35, 0.40, 0.45, "
             "0.50, 0.55, 0.60], nfreq: 64, perturbation: 0.15, max_iter: 50, seed: 42, "
             "tol: 1.0e-8}\n"
             "output: {directory: figures, formats: [table, profiles, traces, convergence], "
             "trace_display: {t_max: 15.0, nw: 1024}}\n")
    yaml.safe_load(text)  # well-formed
    path.write_text(text)
    return path


def test_default_and_reference_configs_load_the_same_run(tmp_path):
    ours = load_config(DEFAULT_CONFIG)
    ref = load_config(_reference_schema_yaml(tmp_path / "reference_schema.yaml"))
    for cfg in (ours, ref):
        same_model(cfg.true_model, default_ocean_crust_model())
        assert cfg.layer_names == ["Ocean", "Sediment", "Crust", "Upper mantle", "Half-space"]
        assert cfg.p_values == pytest.approx([0.05 * k for k in range(1, 13)])
        assert (cfg.nfreq, cfg.perturbation, cfg.max_iter, cfg.seed, cfg.tol) == \
            (64, 0.15, 50, 42, 1e-8)
        assert cfg.output.formats == ["table", "profiles", "traces", "convergence"]
        assert (cfg.output.trace_display.t_max, cfg.output.trace_display.nw) == (15.0, 1024)


def test_save_load_round_trip(tmp_path):
    cfg = load_config(DEFAULT_CONFIG)
    cfg.max_iter, cfg.seed, cfg.output.formats = 7, 3, ["table"]
    save_config(cfg, tmp_path / "sub" / "c.yaml")
    assert ".inf" in (tmp_path / "sub" / "c.yaml").read_text()
    back = load_config(tmp_path / "sub" / "c.yaml")
    same_model(back.true_model, cfg.true_model)
    assert back.layer_names == cfg.layer_names and back.p_values == cfg.p_values
    assert (back.nfreq, back.perturbation, back.max_iter, back.seed, back.tol) == \
        (cfg.nfreq, cfg.perturbation, 7, 3, cfg.tol)
    assert back.output == cfg.output


BAD = {
    "no_model": "inversion: {nfreq: 8}\n",
    "one_layer": "model: {layers: [{alpha: 1.5, beta: 0, rho: 1, thickness: .inf, Q_alpha: 1, Q_beta: 1}]}\n",
    "missing_key": "model: {layers: [{alpha: 1.5, beta: 0, rho: 1, thickness: 1, Q_alpha: 1},"
                   " {alpha: 2, beta: 1, rho: 2, thickness: .inf, Q_alpha: 1, Q_beta: 1}]}\n",
    "elastic_top": "model: {layers: [{alpha: 1.5, beta: 0.5, rho: 1, thickness: 1, Q_alpha: 1, Q_beta: 1},"
                   " {alpha: 2, beta: 1, rho: 2, thickness: .inf, Q_alpha: 1, Q_beta: 1}]}\n",
    "finite_bottom": "model: {layers: [{alpha: 1.5, beta: 0, rho: 1, thickness: 1, Q_alpha: 1, Q_beta: 1},"
                     " {alpha: 2, beta: 1, rho: 2, thickness: 3, Q_alpha: 1, Q_beta: 1}]}\n",
    "negative_rho": "model: {layers: [{alpha: 1.5, beta: 0, rho: -1, thickness: 1, Q_alpha: 1, Q_beta: 1},"
                    " {alpha: 2, beta: 1, rho: 2, thickness: .inf, Q_alpha: 1, Q_beta: 1}]}\n",
}
GOOD_LAYERS = ("model: {layers: [{alpha: 1.5, beta: 0, rho: 1, thickness: 1, Q_alpha: 1, Q_beta: 1},"
               " {alpha: 2, beta: 1, rho: 2, thickness: .inf, Q_alpha: 1, Q_beta: 1}]")
BAD["fixed"] = GOOD_LAYERS + ", fixed_layers: [0, 1]}\n"
BAD["p_zero"] = GOOD_LAYERS + "}\ninversion: {p_values: [0.1, 0.0]}\n"
BAD["format"] = GOOD_LAYERS + "}\noutput: {formats: [table, movie]}\n"


@pytest.mark.parametrize("case", sorted(BAD))
def test_invalid_configs_are_rejected(tmp_path, case):
    f = tmp_path / f"{case}.yaml"
    f.write_text(BAD[case])
    with pytest.raises(ConfigError):
        load_config(f)


def test_missing_file():
    with pytest.raises(FileNotFoundError):
        load_config("/nonexistent/config.yaml")


def fake_result():
    true = default_ocean_crust_model()
    from seismicinversion_amd.layer_model import LayerModel

    def scaled(f):
        return LayerModel.from_arrays(alpha=true.alpha * f, beta=true.beta * f, rho=true.rho,
                                      thickness=true.thickness, Q_alpha=true.Q_alpha,
                                      Q_beta=true.Q_beta)
    return InversionResult(true_model=true, initial_model=scaled(1.1),
                           recovered_model=scaled(1.001), misfit_history=[1.0, 0.1, 0.01],
                           grad_norm_history=[5.0, 0.5, 0.05],
                           param_error_history=[0.1, 0.01, 0.001], n_iterations=3,
                           converged=True)


def test_report_writers(tmp_path):
    res = fake_result()
    write_model_table_latex(res, tmp_path / "t.tex", layer_names=["W", "S", "C", "M", "H"])
    tex = (tmp_path / "t.tex").read_text()
    assert tex.count(r"\multicolumn{6}") == 3 and tex.count(r"\infty") == 3
    assert r"\quad H & 2.202 & 1.101 & 1.800 & \infty & 1.0e-01" in tex
    write_model_profiles_tikz(res, tmp_path / "p.tex")
    tikz = (tmp_path / "p.tex").read_text()
    assert tikz.count(r"\begin{tikzpicture}") == 1 and tikz.count("dashed") >= 3
    d, v = depth_profile(res.true_model, "alpha")
    np.testing.assert_array_equal(d, [0, 2, 2, 3, 3, 4, 4, 5, 5, 7])
    np.testing.assert_array_equal(v, np.repeat(res.true_model.alpha, 2))
    pytest.importorskip("matplotlib")
    plot_convergence_curves(res, tmp_path / "c.pdf")
    assert (tmp_path / "c.pdf").stat().st_size > 1000
