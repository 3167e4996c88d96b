"""The differentiation + Newton-LM workload at its stated size on the GPU:
20-layer seeded model (seed 42), 128 slownesses in (0, 0.6], 512 frequencies,
P = 75 packed parameters (taup_inversion.py:112-187, 262-470;
kennett_torch.py:644-684).

* misfit, gradient and the full 75 x 75 exact Hessian summed over all 128
  slownesses, at the first three iterates of the oracle LM, against the
  oracle's torch.func.hessian (tests/golden/c4_lm_oracle.npz) to 1e-8, in
  both Hessian modes (MFMA Gram + second-order term, and full HVPs);
* Hessian symmetry, and H d against central differences of the device
  gradient at step 1e-6 along three directions;
* the 10-iteration device LM against the oracle LM's trajectory.
"""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from seismicinversion_amd import engine
from seismicinversion_amd.kennett_torch import model_to_tensors
from seismicinversion_amd.layer_model import LayerModel
from seismicinversion_amd.taup_inversion import _log_transform, _Problem, invert_taup

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "c4_lm_oracle.npz")
TOL = 1e-8


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def c4():
    d = np.load(GOLD)
    tab = d["table"]
    model = LayerModel.from_arrays(*[tab[:, k] for k in range(6)])
    p = list(d["p"])
    t = model_to_tensors(model)
    refs = [t[k] for k in ("alpha", "beta", "rho", "thickness", "Q_alpha", "Q_beta")]
    w = torch.arange(1, int(d["nfreq"]) + 1, dtype=torch.float64) * (2 * np.pi / 64.0)
    obs = engine.forward(engine.layers_tensor(model), torch.tensor(p), w.to(torch.complex128))
    return d, model, p, _Problem(*refs, p, w, obs)


@pytest.mark.parametrize("mode", ["gram", "hvp"])
def test_full_hessian_and_gradient_vs_oracle(c4, mode):
    d, _, _, prob = c4
    assert prob.P == 75
    for k in range(d["H_log"].shape[0]):
        m = np.exp(d["log_x"][k])
        chi, g, H = prob.grad_hessian(m, mode)
        g_log, H_log = _log_transform(m, g, H)
        e_chi = abs(chi - d["misfit"][k]) / d["misfit"][k]
        e_g, e_H = rel(g_log, d["g_log"][k]), rel(H_log, d["H_log"][k])
        print(f"iterate {k} ({mode}): chi {e_chi:.2e}, g {e_g:.2e}, H {e_H:.2e}, "
              f"H asymmetry {rel(H, H.T):.2e}", flush=True)
        assert e_chi < 1e-10 and e_g < TOL and e_H < TOL
        assert rel(H, H.T) < 1e-10


def test_modes_agree_and_hessian_matches_central_differences(c4):
    d, _, _, prob = c4
    m = np.exp(d["log_x"][1])
    _, g0, H = prob.grad_hessian(m, "gram")
    _, _, H_hvp = prob.grad_hessian(m, "hvp")
    assert rel(H, H_hvp) < 1e-10
    rng = np.random.default_rng(5)
    step = 1e-6
    for j in range(3):
        dm = rng.standard_normal(prob.P) * m  # relative perturbation of every parameter
        _, gp, _ = prob.grad_hessian(m + step * dm, "gram")
        _, gm, _ = prob.grad_hessian(m - step * dm, "gram")
        fd = (gp - gm) / (2 * step)
        e = rel(H @ dm, fd)
        print(f"direction {j}: H d vs central differences {e:.2e}", flush=True)
        assert e < 1e-6


def test_lm_trajectory_vs_oracle(c4):
    d, model, p, _ = c4
    res = invert_taup(model, p, nfreq=int(d["nfreq"]), perturbation=float(d["perturbation"]),
                      max_iter=len(d["misfit"]), seed=int(d["seed"]), tol=0.0)
    n = min(len(res.misfit_history), len(d["misfit"]))
    e_chi = np.abs(np.array(res.misfit_history[:n]) - d["misfit"][:n]) / d["misfit"][:n]
    e_err = np.abs(np.array(res.param_error_history[:n]) - d["param_err"][:n])
    print("misfit rel diff per iteration", e_chi, "\nparam error diff", e_err, flush=True)
    assert n == len(d["misfit"])
    assert np.all(e_chi[:4] < 1e-8)  # the first iterations follow the oracle closely
    assert np.all(e_chi < 1e-5) and np.all(e_err < 1e-7)
    assert res.converged == bool(d["param_err"][-1] < 0.01)
    # misfit decreases monotonically (every accepted step lowers it)
    assert np.all(np.diff(res.misfit_history) < 0)
