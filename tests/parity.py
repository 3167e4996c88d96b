"""Parity helpers shared by the tests.

The reference's own forward test (Kennett_Reflectivity/test_torch_gradients.py:79)
measures max|R - R_ref| / max|R_ref| over the frequencies of one slowness and
asks for <= 1e-12.  That bar is met wherever the reflectivity map is well
conditioned.  Where it is not (large slowness, near interface-wave poles,
deep models), every float64 evaluation carries more rounding error than
that, the reference's own included: on its default model at p = 1.2 its
golden vector is 4.9e-10 from the exact value of its own algorithm.

So each row is judged against the exact value: the oracle run in x87
extended precision on the float64 inputs (``exact``; 2048 times finer than
float64, pinned against 40-digit mpmath in tests/test_oracle.py).  A row
passes when its error against the exact value is at most

    max(1e-12, FACTOR * max(e_ref, e_ulp))

with FACTOR = 2, e_ref the error of the reference's own float64 evaluation
of that row (its golden vector where the reference wrote one, and the
float64 oracle, which restates the reference's operation order) and e_ulp
the largest change of the row under random 1-ulp perturbations of the layer
parameters (the float64 conditioning of the row).  A kernel that passes is
at most twice as far from the exact result as the reference itself, or as
a 1-ulp change of the inputs moves it.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle import kennett_numpy as K
from seismicinversion_amd.layer_model import LayerModel

COLS = ("alpha", "beta", "rho", "thickness", "Q_alpha", "Q_beta")
REL_TOL = 1e-12
FACTOR = 2.0
EXT = np.clongdouble


def model_from(d, prefix="") -> LayerModel:
    return LayerModel.from_arrays(**{c: np.asarray(d[prefix + c]) for c in COLS})


def row_rel_err(R, R_ref):
    """max_w |R - R_ref| / max_w |R_ref| for every slowness row (in the wider
    of the two precisions)."""
    R = np.atleast_2d(R)
    R_ref = np.atleast_2d(R_ref)
    dt = np.result_type(R, R_ref)
    R, R_ref = R.astype(dt), R_ref.astype(dt)
    return np.asarray(np.max(np.abs(R - R_ref), axis=-1) /
                      (np.max(np.abs(R_ref), axis=-1) + 1e-30), dtype=np.float64)


def exact(model, p, omega, free_surface):
    """The reference algorithm's value on these inputs, to extended precision."""
    return K.reflectivity_batch(model, p, omega, free_surface, ctype=EXT)


def ulp_sensitivity(model, p, omega, free_surface, trials=6, seed=1234):
    base = K.reflectivity_batch(model, p, omega, free_surface)
    rng = np.random.default_rng(seed)
    worst = np.zeros(base.shape[0])
    for _ in range(trials):
        cols = {c: np.array(getattr(model, c), dtype=np.float64) for c in COLS}
        for c in ("alpha", "beta", "rho", "Q_alpha"):
            cols[c] = cols[c] * (1.0 + rng.choice([-1.0, 1.0], size=cols[c].shape) * 2.0 ** -52)
        pert = K.reflectivity_batch(LayerModel.from_arrays(**cols), p, omega, free_surface)
        worst = np.maximum(worst, row_rel_err(pert, base))
    return worst


@dataclass
class Accuracy:
    """Exact rows and the per-row bar, with what the bar is made of."""

    X: np.ndarray        # exact values (extended precision)
    bar: np.ndarray      # per-row bar on the error against X
    e_oracle: np.ndarray  # float64 oracle's error against X
    e_golden: np.ndarray | None  # the reference's golden vectors' error against X
    e_ulp: np.ndarray    # 1-ulp sensitivity of the row


def accuracy(model, p, omega, free_surface, golden=None) -> Accuracy:
    p = np.atleast_1d(np.asarray(p, np.float64))
    X = exact(model, p, omega, free_surface)
    e_o = row_rel_err(K.reflectivity_batch(model, p, omega, free_surface), X)
    e_g = row_rel_err(golden, X) if golden is not None else None
    e_u = ulp_sensitivity(model, p, omega, free_surface)
    scale = np.maximum(e_o, e_u) if e_g is None else np.maximum(np.maximum(e_o, e_g), e_u)
    return Accuracy(X, np.maximum(REL_TOL, FACTOR * scale), e_o, e_g, e_u)


def assert_accurate(R, acc: Accuracy, what="", table=False):
    """Asserts every row of R is within its bar of the exact value; returns
    the per-row errors against the exact value."""
    err = row_rel_err(R, acc.X)
    if table:
        print(f"{what}: row, HIP err, oracle err" + (", golden err" if acc.e_golden is not None
                                                      else "") + ", 1-ulp sensitivity, bar")
        for k in range(len(err)):
            g = f" {acc.e_golden[k]:.2e}" if acc.e_golden is not None else ""
            print(f"  {k:3d} {err[k]:.2e} {acc.e_oracle[k]:.2e}{g} {acc.e_ulp[k]:.2e} "
                  f"{acc.bar[k]:.2e}")
    bad = np.nonzero(err > acc.bar)[0]
    assert bad.size == 0, (
        f"{what}: {bad.size} rows over the bar; worst row {bad[np.argmax(err[bad] / acc.bar[bad])]}"
        f" err {err[bad].max():.3e} bar {acc.bar[bad].min():.3e}")
    return err


def ulp_tolerance(model, p, omega, free_surface, factor=FACTOR):
    """max(1e-12, factor x the 1-ulp sensitivity) per row, for results that are
    not the reference algorithm's R itself (other formulations, traces)."""
    return np.maximum(REL_TOL, factor * ulp_sensitivity(model, p, omega, free_surface))


def assert_rows_close(R, R_ref, tol, what=""):
    err = row_rel_err(R, R_ref)
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, (
        f"{what}: {bad.size} rows over tolerance; worst row {bad[np.argmax(err[bad] / tol[bad])]}"
        f" err {err[bad].max():.3e} tol {tol[bad].min():.3e}")
    return err
