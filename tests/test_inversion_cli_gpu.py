"""The inversion CLI (taup_inversion.py:1082-1198) end to end on the device:
config in, LM on the GPU, the saved config and the four report files out."""

from __future__ import annotations

import pytest

from seismicinversion_amd import taup_inversion
from seismicinversion_amd.inversion_config import DEFAULT_CONFIG, load_config

pytestmark = pytest.mark.gpu


def test_inversion_cli_writes_outputs(tmp_path):
    taup_inversion.main(["--config", str(DEFAULT_CONFIG), "--max-iter", "4", "--output-dir",
                         str(tmp_path), "--seed", "5", "--graphs"])
    for name in ("inversion_config.yaml", "taup_model_parameters.tex", "taup_model_profiles.tex",
                 "taup_trace_comparison.pdf", "taup_convergence.pdf"):
        assert (tmp_path / name).stat().st_size > 100, name
    saved = load_config(tmp_path / "inversion_config.yaml")
    assert saved.max_iter == 4 and saved.seed == 5 and str(saved.output.directory) == str(tmp_path)
    # the saved config reproduces the run
    a = taup_inversion.invert_taup(saved.true_model, saved.p_values, nfreq=saved.nfreq,
                                   perturbation=saved.perturbation, max_iter=saved.max_iter,
                                   seed=saved.seed, tol=saved.tol)
    h = a.misfit_history
    assert len(h) == 4 and all(h[i + 1] < h[i] for i in range(3))  # LM accepts only decreases
