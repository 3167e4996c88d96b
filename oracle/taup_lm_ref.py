"""Oracle Newton-Levenberg-Marquardt tau-p inversion (TEST INFRASTRUCTURE ONLY).

A CPU restatement of Kennett_Reflectivity/taup_inversion.py on the oracle's
torch forward (kennett_torch_ref):

* misfit, gradient and exact Hessian summed over slownesses, with the
  log-parameter transform g_log = m g, H_log = m m^T H + diag(g m)
  (taup_inversion.py:112-187); the Hessian by torch.func.hessian, as there;
* the misfit alone for trial models (taup_inversion.py:190-230);
* the damped Newton step (H + lam I) dx = -g by np.linalg.solve
  (taup_inversion.py:238-254);
* the start model (perturbation rule), damping schedule (lam_0 = 0.01
  max|diag H|, x5 on rejection, /2 on acceptance floored at 1e-15, at most
  10 trials) and convergence tests (taup_inversion.py:262-470).

Slownesses are evaluated in vmapped groups of GROUP for speed; the sums over
slowness are taken in slowness order.  Pinned by tests/test_oracle_lm.py
against the reference's own 3-iteration trajectory
(tests/golden/invert_taup_2pct.npz); it is the checker of the device LM at
the 20-layer, 128 x 512, P = 75 workload (tests/golden/c4_lm_oracle.npz,
written by tests/golden/make_c4_lm.py).
"""

from __future__ import annotations

import numpy as np
import torch

from . import kennett_torch_ref as T

F = torch.float64
GROUP = 8


def model_columns(tab: np.ndarray) -> list[torch.Tensor]:
    """The six (nl,) float64 tensors of a (nl, 6) table, h = inf -> 1e30
    (the torch path's half-space marker, kennett_torch.py:374)."""
    t = torch.tensor(np.asarray(tab, dtype=np.float64).copy())
    t[:, 3][torch.isinf(t[:, 3])] = 1e30
    return [t[:, k].clone() for k in range(6)]


def _chi_p(m, p, robs, cols, omega):
    a, b, r, h = T.unpack_params(m, *cols[:4])
    R = T.reflectivity(a, b, r, h, cols[4], cols[5], p, omega)
    d = torch.view_as_real(R) - robs
    return (d * d).sum()


def observed(cols, p_values, omega) -> torch.Tensor:
    """(np, nw, 2) real view of R at the model (taup_inversion.py:302-315)."""
    with torch.no_grad():
        return torch.stack([torch.view_as_real(T.reflectivity(*cols, float(p), omega))
                            for p in p_values])


def misfit_grad_hessian_phys(m: np.ndarray, cols, p_values, omega, robs):
    """(chi, g, H) in physical parameters, summed over slownesses."""
    mt = torch.tensor(np.asarray(m, dtype=np.float64))
    ps = torch.tensor([float(p) for p in p_values], dtype=F)
    fn = lambda mm, pp, ro: _chi_p(mm, pp, ro, cols, omega)  # noqa: E731
    chi, g, H = 0.0, np.zeros(len(m)), np.zeros((len(m), len(m)))
    for s in range(0, len(ps), GROUP):
        sl = slice(s, s + GROUP)
        c = torch.func.vmap(fn, in_dims=(None, 0, 0))(mt, ps[sl], robs[sl])
        gg = torch.func.vmap(torch.func.grad(fn), in_dims=(None, 0, 0))(mt, ps[sl], robs[sl])
        hh = torch.func.vmap(torch.func.hessian(fn), in_dims=(None, 0, 0))(mt, ps[sl], robs[sl])
        for k in range(c.shape[0]):  # slowness order
            chi += float(c[k])
            g += gg[k].numpy()
            H += hh[k].numpy()
    return chi, g, H


def misfit_grad_hessian(log_x, cols, p_values, omega, robs):
    """(chi, g_log, H_log) (taup_inversion.py:112-187)."""
    m = np.exp(np.asarray(log_x, dtype=np.float64))
    chi, g, H = misfit_grad_hessian_phys(m, cols, p_values, omega, robs)
    return chi, m * g, np.outer(m, m) * H + np.diag(g * m)


def misfit(log_x, cols, p_values, omega, robs) -> float:
    """taup_inversion.py:190-230."""
    mt = torch.tensor(np.exp(np.asarray(log_x, dtype=np.float64)))
    with torch.no_grad():
        return float(sum(float(_chi_p(mt, float(p), robs[k], cols, omega))
                         for k, p in enumerate(p_values)))


def start_params(true_params: np.ndarray, n_sub: int, perturbation: float, seed: int):
    """The perturbed start model (taup_inversion.py:326-348)."""
    n = len(true_params)
    rng = np.random.default_rng(seed)
    f = np.ones(n)
    n_vel = 2 * n_sub
    f[:n_vel] = rng.uniform(1.0 - perturbation, 1.0 + perturbation, n_vel)
    f[n_vel:n_vel + n_sub] = rng.uniform(1.0 - perturbation * 0.67, 1.0 + perturbation * 0.67,
                                         n_sub)
    f[n_vel + n_sub:] = rng.uniform(1.0 - perturbation * 1.33, 1.0 + perturbation * 1.33,
                                    n_sub - 1)
    return true_params * f


def invert(tab: np.ndarray, p_values, nfreq: int = 64, perturbation: float = 0.15,
           max_iter: int = 100, seed: int = 42, tol: float = 1e-8, max_trials: int = 10,
           log=None) -> dict:
    """The reference's invert_taup loop (taup_inversion.py:262-470) on the
    oracle.  Returns the histories (misfit, grad_norm, param_err, lam before
    each step, trials per step) and the iterates (log-space, one row per
    iteration, plus the final one)."""
    cols = model_columns(tab)
    omega = torch.arange(1, nfreq + 1, dtype=F) * (2.0 * np.pi / 64.0)
    robs = observed(cols, p_values, omega)
    true_params = T.pack_params(*cols[:4]).numpy()
    n_sub = len(cols[0]) - 1
    log_x = np.log(start_params(true_params, n_sub, perturbation, seed))
    out = {"misfit": [], "grad_norm": [], "param_err": [], "lam": [], "trials": [],
           "log_x": [log_x.copy()]}
    lam = None
    for it in range(max_iter):
        chi, g, H = misfit_grad_hessian(log_x, cols, p_values, omega, robs)
        gn = float(np.linalg.norm(g))
        err = float(np.linalg.norm(np.exp(log_x) - true_params) / np.linalg.norm(true_params))
        out["misfit"].append(chi)
        out["grad_norm"].append(gn)
        out["param_err"].append(err)
        if log is not None:
            log(f"oracle LM iter {it}: chi={chi:.6e} |g|={gn:.6e} err={err:.6e}")
        if lam is None:
            lam = 0.01 * max(float(np.abs(np.diag(H)).max()), 1e-10)
        if gn < tol:
            break
        if len(out["misfit"]) > 1 and abs(out["misfit"][-2] - chi) / max(abs(chi), 1e-30) < tol:
            break
        out["lam"].append(lam)
        accepted = False
        for trial in range(max_trials):
            x_try = log_x + np.linalg.solve(H + lam * np.eye(len(g)), -g)
            if misfit(x_try, cols, p_values, omega, robs) < chi:
                log_x = x_try
                lam = max(lam / 2.0, 1e-15)
                accepted = True
                out["trials"].append(trial + 1)
                break
            lam *= 5.0
        out["log_x"].append(log_x.copy())
        if not accepted:
            out["trials"].append(-1)
            break
    return out
