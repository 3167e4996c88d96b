"""The differentiable drop-in (kennett_torch), the Frechet and
scattering-matrix drop-ins, the MFMA Gram kernel and the Levenberg-Marquardt
inversion, against the reference's golden vectors (tests/golden/*.npz,
written by the reference itself) and the oracle."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from oracle import kennett_numpy as K
from oracle import kennett_torch_ref as T
from parity import model_from
from seismicinversion_amd import engine
from seismicinversion_amd import kennett_torch as kt
from seismicinversion_amd.frechet_analytical import frechet_kennett
from seismicinversion_amd.scattering_matrices import (ocean_bottom_interface,
                                                      solid_solid_interface)
from seismicinversion_amd.taup_inversion import (compute_taup_traces, invert_taup,
                                                 taup_misfit_grad_hessian)
from seismicinversion_amd.workloads import default_ocean_crust_model, seeded_model

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name))


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def ref():
    d = gold("torch_derivs.npz")
    tens = kt.model_to_tensors(model_from(d))
    return d, tens, torch.tensor(d["omega"])


def test_forward_jacobian_hessian_api_vs_golden(ref):
    d, tens, w = ref
    for pv, key in ((0.2, "02"), (0.4, "04")):
        R = kt.kennett_reflectivity_torch(**tens, p=pv, omega=w)
        assert R.device.type == "cpu" and R.dtype == torch.complex128
        assert rel(R.numpy(), d[f"R_p{key}"]) < 1e-12
        J = kt.jacobian(**tens, p=pv, omega=w)
        assert J.shape == (32, 15)
        assert rel(J.numpy(), d[f"J_p{key}"]) < 1e-10
        Jfs = kt.jacobian(**tens, p=pv, omega=w, free_surface=True)
        assert rel(Jfs.numpy(), d[f"Jfs_p{key}"]) < 1e-10
    H = kt.hessian(**tens, p=0.2, omega=w, R_obs=torch.tensor(d["R_obs_hess"]))
    assert rel(H.numpy(), d["H_p02"]) < 1e-10
    H0 = kt.hessian(**tens, p=0.4, omega=w)
    assert rel(H0.numpy(), d["H0_p04"]) < 1e-10


def test_torch_func_transforms_on_user_closures(ref):
    """The reference's own jacobian/hessian recipes (torch.func.jacrev on Re/Im,
    torch.func.hessian of the misfit) applied to a closure over the drop-in."""
    d, tens, w = ref
    a, b, r, h, qa, qb = (tens[k] for k in ("alpha", "beta", "rho", "thickness", "Q_alpha",
                                            "Q_beta"))
    params = kt._pack_params(a, b, r, h)

    def fwd(pv):
        return kt.kennett_reflectivity_torch(*kt._unpack_params(pv, a, b, r, h), qa, qb, 0.2, w)

    J = torch.func.jacrev(lambda x: fwd(x).real)(params) + \
        1j * torch.func.jacrev(lambda x: fwd(x).imag)(params)
    assert rel(J.numpy(), d["J_p02"]) < 1e-10
    R_obs = torch.tensor(d["R_obs_hess"])

    def misfit(pv):
        res = fwd(pv) - R_obs
        return (res.real ** 2 + res.imag ** 2).sum()

    H = torch.func.hessian(misfit)(params)
    assert rel(H.numpy(), d["H_p02"]) < 1e-10
    H2 = torch.autograd.functional.hessian(misfit, params)
    assert rel(H2.numpy(), d["H_p02"]) < 1e-10
    # plain backward
    x = params.clone().requires_grad_(True)
    misfit(x).backward()
    xt = params.clone().requires_grad_(True)
    res = T.reflectivity(*T.unpack_params(xt, a, b, r, h), qa, qb, 0.2, w) - R_obs
    (res.real ** 2 + res.imag ** 2).sum().backward()
    assert rel(x.grad.numpy(), xt.grad.numpy()) < 1e-10


def test_frechet_kennett_vs_golden(ref):
    d, _, _ = ref
    m = model_from(d)
    for pv, key in ((0.2, "02"), (0.4, "04")):
        F = frechet_kennett(m.alpha, m.beta, m.rho, m.thickness, m.Q_alpha, m.Q_beta, pv,
                            d["omega"])
        assert rel(F, d[f"F_p{key}"]) < 1e-10


def test_scattering_matrices_vs_golden():
    d = gold("interfaces.npz")
    m = model_from(d)
    bc = m.complex_velocity_s()
    for jp, pv in enumerate(d["p"]):
        eta, neta = d["eta"][:, jp], d["neta"][:, jp]
        c = ocean_bottom_interface(pv, eta[0], m.rho[0], eta[1], neta[1], m.rho[1], bc[1])
        for k, M in enumerate((c.Rd, c.Ru, c.Tu, c.Td)):
            ref_m = d[("Rd", "Ru", "Tu", "Td")[k]][0, jp]
            assert np.max(np.abs(M - ref_m)) <= 1e-11 * np.max(np.abs(ref_m)) + 1e-300
        for il in (1, 7, 18):
            c = solid_solid_interface(pv, eta[il], neta[il], m.rho[il], bc[il], eta[il + 1],
                                      neta[il + 1], m.rho[il + 1], bc[il + 1])
            for k, M in enumerate((c.Rd, c.Ru, c.Tu, c.Td)):
                ref_m = d[("Rd", "Ru", "Tu", "Td")[k]][il, jp]
                assert np.max(np.abs(M - ref_m)) <= 1e-11 * np.max(np.abs(ref_m))


@pytest.mark.parametrize("P,K", [(1, 7), (15, 64), (17, 4099), (75, 131072), (33, 10000),
                                 (63, 8192), (130, 40), (395, 8192)])
def test_mfma_gram_matches_matmul(P, K):
    rng = np.random.default_rng(P * 1000 + K)
    X = rng.standard_normal((P, K))
    G = engine.gram(torch.from_numpy(X).cuda(), 2.0).cpu().numpy()
    expect = 2.0 * X @ X.T
    assert rel(G, expect) < 1e-13
    np.testing.assert_array_equal(G, G.T)


def test_lm_hessian_modes_agree_and_match_torch_ad():
    m = default_ocean_crust_model()
    tens = kt.model_to_tensors(m)
    refs = (tens["alpha"], tens["beta"], tens["rho"], tens["thickness"], tens["Q_alpha"],
            tens["Q_beta"])
    p_values = [0.1, 0.2, 0.3, 0.4]
    w = torch.arange(1, 33, dtype=torch.float64) * (2 * np.pi / 64)
    R_obs = {p: kt.kennett_reflectivity_torch(*refs, p, w).detach() for p in p_values}
    x0 = np.log(kt._pack_params(*refs[:4]).numpy() * 1.03)
    chi_g, g_g, H_g = taup_misfit_grad_hessian(x0, *refs, p_values, w, R_obs, "gram")
    chi_h, g_h, H_h = taup_misfit_grad_hessian(x0, *refs, p_values, w, R_obs, "hvp")
    assert abs(chi_g - chi_h) <= 1e-14 * chi_h
    assert rel(H_g, H_h) < 1e-11
    # the reference's recipe through the oracle's torch restatement
    phys = torch.tensor(np.exp(x0))
    tot_g = np.zeros(len(x0))
    tot_H = np.zeros((len(x0), len(x0)))
    for pv in p_values:
        def mis(pvec):
            R = T.reflectivity(*T.unpack_params(pvec, *refs[:4]), refs[4], refs[5], pv, w)
            res = R - R_obs[pv]
            return (res.real ** 2 + res.imag ** 2).sum()
        tot_g += torch.func.grad(mis)(phys).numpy()
        tot_H += torch.func.hessian(mis)(phys).numpy()
    mv = np.exp(x0)
    assert rel(g_g, mv * tot_g) < 1e-9
    assert rel(H_g, np.outer(mv, mv) * tot_H + np.diag(tot_g * mv)) < 1e-9


def test_invert_taup_trajectory_vs_golden():
    d = gold("invert_taup_2pct.npz")
    res = invert_taup(model_from(d), list(d["p_values"]), nfreq=int(d["nfreq"]),
                      perturbation=float(d["perturbation"]), max_iter=int(d["max_iter"]))
    print("misfit", res.misfit_history, "ref", d["misfit"])
    assert len(res.misfit_history) == len(d["misfit"])
    np.testing.assert_allclose(res.misfit_history[:2], d["misfit"][:2], rtol=1e-8)
    np.testing.assert_allclose(res.param_error_history, d["param_err"], rtol=1e-6)
    init = model_from(d, "init_")
    np.testing.assert_allclose(res.initial_model.alpha, init.alpha, rtol=1e-15)


def test_invert_taup_convergence_like_reference_tests():
    """test_taup_inversion.py:41-65: 5 % -> < 0.1 %; 2 % -> < 1e-6 in <= 20 iterations."""
    m = default_ocean_crust_model()
    r5 = invert_taup(m, [0.1, 0.2, 0.3, 0.4], nfreq=32, perturbation=0.05, max_iter=30)
    assert r5.converged and r5.param_error_history[-1] < 1e-3
    r2 = invert_taup(m, [0.1, 0.2, 0.3, 0.4], nfreq=32, perturbation=0.02, max_iter=20)
    assert r2.converged and r2.n_iterations <= 20 and r2.param_error_history[-1] < 1e-6


def test_compute_taup_traces_vs_oracle():
    m = seeded_model(7, 4)
    p_values = [0.1, 0.35]
    t, tr = compute_taup_traces(m, p_values, nw=512)
    w = np.arange(1, 512) * 2 * np.pi / 64
    S = K.ricker_spectrum(w, 512 * 2 * np.pi / 64)
    for pv in p_values:
        R = K.reflectivity(m, pv, w)
        _, expect = K.hermitian_ifft((R * S)[None], 512, 0.0, 64.0)
        assert rel(tr[pv], expect[0]) < 1e-11


def test_residual_norms_kernel():
    rng = np.random.default_rng(3)
    for n_p, nw in ((1, 1), (7, 300), (33, 1023)):
        R = torch.tensor(rng.standard_normal((n_p, nw)) + 1j * rng.standard_normal((n_p, nw)),
                         device="cuda")
        Ro = torch.tensor(rng.standard_normal((n_p, nw)) + 1j * rng.standard_normal((n_p, nw)),
                          device="cuda")
        chi = engine.residual_norms(R, Ro).cpu().numpy()
        want = (np.abs(R.cpu().numpy() - Ro.cpu().numpy()) ** 2).sum(1)
        assert np.max(np.abs(chi - want) / want) < 1e-13
        assert torch.equal(engine.residual_norms(R, Ro), engine.residual_norms(R, Ro))


@pytest.mark.parametrize("mode", ["gram", "hvp"])
def test_lm_graph_replays_match_eager(mode):
    """Each replay of the captured misfit / gradient / Hessian graphs, with
    eager work and synchronisations in between, equals the eager result bit
    for bit."""
    from seismicinversion_amd.taup_inversion import _Problem

    m = seeded_model(7, 5)
    tens = kt.model_to_tensors(m)
    refs = (tens["alpha"], tens["beta"], tens["rho"], tens["thickness"], tens["Q_alpha"],
            tens["Q_beta"])
    p_values = list(np.linspace(0.05, 0.5, 12))
    w = torch.arange(1, 97, dtype=torch.float64) * (2 * np.pi / 64)
    obs = engine.forward(engine.layers_tensor(m), torch.tensor(p_values), w.to(torch.complex128))
    eager = _Problem(*refs, p_values, w, obs)
    graphed = _Problem(*refs, p_values, w, obs, use_graphs=True)
    base = kt._pack_params(*refs[:4]).numpy()
    rng = np.random.default_rng(11)
    for it in range(4):
        phys = base * rng.uniform(0.95, 1.05, base.size)
        chi_e, g_e, H_e = eager.grad_hessian(phys, mode)
        mis_e = eager.misfit(phys * 1.01)
        torch.cuda.synchronize()
        chi_r, g_r, H_r = graphed.grad_hessian(phys, mode)  # eager, then captured, then replays
        mis_r = graphed.misfit(phys * 1.01)
        assert chi_r == chi_e and mis_r == mis_e, f"replay {it}"
        assert np.array_equal(g_r, g_e) and np.array_equal(H_r, H_e), f"replay {it}"
    assert set(graphed._graphs) == {"gh_" + mode, "misfit"}


@pytest.mark.parametrize("device_loop", [True, False])
def test_invert_taup_with_graphs_equals_eager(device_loop):
    m = seeded_model(6, 8)
    p = list(np.linspace(0.05, 0.45, 10))
    kw = dict(nfreq=48, perturbation=0.05, max_iter=4, tol=0.0, device_loop=device_loop)
    a = invert_taup(m, p, **kw)
    b = invert_taup(m, p, use_graphs=True, **kw)
    assert a.misfit_history == b.misfit_history
    assert a.grad_norm_history == b.grad_norm_history
    assert np.array_equal(a.recovered_model.alpha, b.recovered_model.alpha)


def test_20_layer_jacobian_hessian_frechet_vs_golden():
    """The reference's own Jacobian (jacrev), Hessian (torch.func.hessian of
    ||R - R_obs||^2) and analytical Frechet derivatives on the seeded 20-layer
    model (P = 75), 8 frequencies, p = 0.15 and 0.35
    (tests/golden/torch_derivs_20layer.npz; kennett_torch.py:611-684,
    frechet_analytical.py:583): the drop-ins through the device kernels to
    1e-10 (Jacobian, Frechet) and 1e-8 (Hessian)."""
    d = gold("torch_derivs_20layer.npz")
    m = model_from(d)
    tens = kt.model_to_tensors(m)
    w = torch.tensor(d["omega"])
    for pv, key in ((0.15, "015"), (0.35, "035")):
        R = kt.kennett_reflectivity_torch(**tens, p=pv, omega=w)
        assert rel(R.numpy(), d[f"R_p{key}"]) < 1e-12
        J = kt.jacobian(**tens, p=pv, omega=w)
        assert J.shape == (8, 75)
        eJ = rel(J.numpy(), d[f"J_p{key}"])
        F = frechet_kennett(m.alpha, m.beta, m.rho, m.thickness, m.Q_alpha, m.Q_beta, pv,
                            d["omega"])
        eF = rel(F, d[f"F_p{key}"])
        H = kt.hessian(**tens, p=pv, omega=w, R_obs=torch.tensor(d[f"R_obs_p{key}"]))
        eH = rel(H.numpy(), d[f"H_p{key}"])
        print(f"p={pv}: J {eJ:.2e}, Frechet {eF:.2e}, H {eH:.2e}")
        assert eJ < 1e-10 and eF < 1e-10 and eH < 1e-8


@pytest.mark.parametrize("p,fs,nw", [(0.2, False, 1024), (0.45, True, 512)])
def test_forward_model_torch_vs_oracle(p, fs, nw):
    """forward_model_torch (kennett_torch.py:488-544): reflectivity x Ricker
    through the Hermitian FFT, against the oracle's trace (NumPy) to the
    reflectivity's parity bar, and its gradient (through the device adjoint)
    against torch AD of the same pipeline on the oracle's torch forward."""
    from parity import ulp_tolerance

    for model in (default_ocean_crust_model(), seeded_model(12, 4)):
        tens = kt.model_to_tensors(model)
        t, s = kt.forward_model_torch(**tens, p=p, T=64.0, nw=nw, free_surface=fs)
        assert t.shape == (2 * nw,) and s.shape == (2 * nw,) and s.dtype == torch.float64
        dw = 2 * np.pi / 64.0
        w = np.arange(1, nw) * dw
        R = K.reflectivity_batch(model, np.array([p]), w, fs)
        t_o, s_o = K.hermitian_ifft(R * K.ricker_spectrum(w, nw * dw)[None, :], nw, 0.0, 64.0)
        np.testing.assert_array_equal(t.numpy(), t_o)
        err = np.max(np.abs(s.numpy() - s_o[0])) / np.max(np.abs(s_o[0]))
        tol = ulp_tolerance(model, np.array([p]), w, fs)[0]
        print(f"p={p} nl={model.n_layers}: trace {err:.2e} (bar {tol:.2e})")
        assert err <= tol

        # gradient of a scalar of the trace, device adjoint vs oracle AD
        a, b, r, h, qa, qb = (tens[k].clone() for k in ("alpha", "beta", "rho", "thickness",
                                                       "Q_alpha", "Q_beta"))
        x = kt._pack_params(a, b, r, h).clone().requires_grad_(True)
        weights = torch.linspace(-1.0, 1.0, 2 * nw, dtype=torch.float64)
        _, s1 = kt.forward_model_torch(*kt._unpack_params(x, a, b, r, h), qa, qb, p, 64.0, nw, fs)
        (s1 * weights).sum().backward()
        xt = kt._pack_params(a, b, r, h).clone().requires_grad_(True)
        wt = torch.arange(1, nw, dtype=torch.float64) * dw
        Rt = T.reflectivity(*T.unpack_params(xt, a, b, r, h), qa, qb, p, wt, fs)
        a_src = 0.1 * nw * dw
        Z = wt ** 2 / (4.0 * a_src ** 2)
        St = Z * torch.exp(-Z + 1j * wt * (5.0 / (a_src * 2.0 ** 0.5)))
        Y = Rt * St
        X = torch.cat([torch.zeros(1, dtype=torch.complex128), Y,
                       torch.zeros(1, dtype=torch.complex128), torch.conj(torch.flip(Y, [0]))])
        (torch.fft.fft(X).real * weights).sum().backward()
        eg = rel(x.grad.numpy(), xt.grad.numpy())
        print(f"   gradient vs oracle AD {eg:.2e}")
        assert eg < 1e-8
