"""The oracle LM (oracle/taup_lm_ref.py) against the reference's own
3-iteration invert_taup trajectory (tests/golden/invert_taup_2pct.npz,
written by the reference): misfit, gradient norm and parameter error per
iteration, start and recovered models.  This pins the oracle that checks the
device LM at the 20-layer, 128 x 512, P = 75 workload."""

from __future__ import annotations

import os

import numpy as np

from oracle import taup_lm_ref as LM

GOLD = os.path.join(os.path.dirname(__file__), "golden")
COLS = ("alpha", "beta", "rho", "thickness", "Q_alpha", "Q_beta")


def test_oracle_lm_follows_the_reference_trajectory():
    d = np.load(os.path.join(GOLD, "invert_taup_2pct.npz"))
    tab = np.stack([d[c] for c in COLS], axis=1)
    out = LM.invert(tab, list(d["p_values"]), nfreq=int(d["nfreq"]),
                    perturbation=float(d["perturbation"]), max_iter=int(d["max_iter"]), tol=0.0)
    for key, gkey in (("misfit", "misfit"), ("grad_norm", "grad_norm"), ("param_err", "param_err")):
        got, ref = np.asarray(out[key]), d[gkey]
        assert got.shape == ref.shape
        err = float(np.max(np.abs(got - ref) / np.abs(ref)))
        print(key, err)
        assert err < 1e-9, (key, got, ref)
    n = len(d["alpha"]) - 1
    start = np.exp(out["log_x"][0])
    np.testing.assert_allclose(start[:n], d["init_alpha"][1:], rtol=1e-14)
    np.testing.assert_allclose(start[3 * n:], d["init_thickness"][1:-1], rtol=1e-14)
    rec = np.exp(out["log_x"][-1])
    for k, c in enumerate(("alpha", "beta", "rho")):
        np.testing.assert_allclose(rec[k * n:(k + 1) * n], d["rec_" + c][1:], rtol=1e-9)
    np.testing.assert_allclose(rec[3 * n:], d["rec_thickness"][1:-1], rtol=1e-9)


def test_oracle_lm_start_rule():
    """The start model draws its factors in the reference's order: velocities
    (+-p), densities (+-0.67 p), then thicknesses (+-1.33 p)."""
    true = np.arange(1.0, 8.0)
    s = LM.start_params(true, n_sub=2, perturbation=0.1, seed=3)
    rng = np.random.default_rng(3)
    f = np.concatenate([rng.uniform(0.9, 1.1, 4), rng.uniform(1 - 0.067, 1 + 0.067, 2),
                        rng.uniform(1 - 0.133, 1 + 0.133, 1)])
    np.testing.assert_array_equal(s, true * f)
