"""Gather on the GPU (kh_forward_grid -> kh_bessel_sum -> kh_hermitian_traces)
against the reference's float64 gather (golden), the oracle, scipy's J0 and
NumPy's FFT, plus size-independent properties at the full gather size."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch
from scipy.special import j0 as scipy_j0

from oracle import kennett_numpy as K
from parity import model_from
from seismicinversion_amd import engine
from seismicinversion_amd.kennett_gather import compute_gather
from seismicinversion_amd.kennett_gather_gpu import compute_gather_gpu, gather_device
from seismicinversion_amd.source import ricker_spectrum
from seismicinversion_amd.workloads import gather_config_c2, seeded_model

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_compute_gather_vs_reference_float64_golden():
    d = np.load(os.path.join(GOLD, "gather_cpu.npz"))
    m = model_from(d)
    t, off, g = compute_gather(m, d["offsets"], T=float(d["T"]), nw=int(d["nw"]),
                               np_slow=int(d["np_slow"]), p_max=float(d["p_max"]),
                               free_surface=bool(d["free_surface"]))
    np.testing.assert_array_equal(t, d["time"])
    err = rel(g, d["gather"])
    print("gather vs reference float64 gather:", err)
    assert err <= 1e-12


def test_compute_gather_gpu_vs_oracle_20_layers():
    m = seeded_model(20, 42)
    offsets = np.arange(0.5, 8.0 + 0.25, 0.5)
    kw = dict(T=64.0, nw=256, np_slow=1024, p_max=1.2, free_surface=True)
    src = lambda w, wmax: ricker_spectrum(w, 100.0 * np.pi)  # noqa: E731
    _, _, g = compute_gather_gpu(m, offsets, source_func=src, **kw)
    # oracle: compute_gather_gpu tapers with p_max = last slowness
    _, _, go = K.gather(m, offsets, source_func=src, **kw)
    err = rel(g, go)
    print("20-layer gather vs oracle:", err)
    assert err <= 1e-12


def test_j0_against_scipy():
    """One slowness, R = 1, S = 1: U[r, k] / (taper p w dp) = J0(w p r)."""
    dev = torch.device("cuda")
    nw, dw, p, dp = 1024, 2 * np.pi / 64.0, 0.9, 1e-3
    offsets = np.concatenate([[0.0, 1e-3, 0.37], np.linspace(0.5, 80.0, 61)])
    R = torch.ones((1, nw - 1), dtype=torch.complex128, device=dev)
    U = engine.bessel_sum(R, torch.tensor([p], dtype=torch.float64), dp, 1e300, 1.0, dw,
                          torch.from_numpy(offsets)).cpu().numpy()
    w = np.arange(1, nw) * dw
    J = U.real / (p * w * dp)[None, :]
    x = (w * p)[None, :] * offsets[:, None]
    expect = scipy_j0(x)
    err = np.max(np.abs(J - expect))
    print("max |J0 - scipy j0| over", expect.size, "arguments up to", x.max(), ":", err)
    # scipy's own error (1e-15 below 12, 3e-15 above) plus its rounding of
    # x - pi/4, half an ulp of x in phase (test_host_math.py pins the same
    # J0 against 40-digit mpmath to 4e-16 of the envelope)
    amp = np.sqrt(2.0 / (np.pi * np.maximum(x, 1e-300)))
    env = np.minimum(np.maximum(amp, np.abs(expect)), 1.0)
    excess = (np.abs(J - expect) - amp * 0.5 * np.spacing(x)) / env
    assert excess.max() < 3.5e-15, excess.max()
    assert err < 5e-15
    assert np.all(U.imag == 0)


def test_hermitian_traces_vs_numpy_fft():
    """Half-length complex FFT of the Hermitian vector vs numpy.fft on the full
    one, up to nt = 8192 (and 16384 where the device allows 128 KiB of LDS
    per workgroup)."""
    rng = np.random.default_rng(3)
    sizes = [1, 2, 4, 64, 1024, 2048, 4096]
    if torch.cuda.get_device_properties(0).shared_memory_per_block >= 8192 * 16:
        sizes.append(8192)
    for nw in sizes:
        U = rng.standard_normal((5, nw - 1)) + 1j * rng.standard_normal((5, nw - 1))
        tr = engine.hermitian_traces(torch.from_numpy(U).cuda(), nw, 0.049, 64.0).cpu().numpy()
        _, expect = K.hermitian_ifft(U, nw, 0.049, 64.0)
        assert tr.shape == (5, 2 * nw)
        if nw == 1:
            assert np.all(tr == 0.0)
            continue
        assert rel(tr, expect) < 1e-13, nw


def test_gather_at_reference_defaults():
    """compute_gather_gpu's own defaults (nw = 2048 -> nt = 4096, 4096
    slownesses up to 1.2 s/km): traces against the oracle's Bessel sum and
    FFT of the same device reflectivity."""
    m = seeded_model(5, 42)
    offsets = np.array([1.0, 7.5])
    t, off, g = compute_gather_gpu(m, offsets)
    assert g.shape == (2, 4096) and np.all(np.isfinite(g))
    dev = torch.device("cuda")
    gamma = np.pi / 64.0
    w = np.arange(1, 2048) * (2 * np.pi / 64.0)
    S = K.ricker_spectrum(w, 2048 * 2 * np.pi / 64.0)  # the reference's default source
    _, R, U = gather_device(engine.layers_tensor(m, dev), torch.from_numpy(offsets).to(dev), 64.0,
                            2048, 4096, 1.2, gamma, torch.from_numpy(S).to(dev), False, keep=True)
    pn = np.arange(1, 4097) * (1.2 / 4096)
    Uo = K.bessel_sum(R.cpu().numpy(), pn, w, offsets, 1.2 / 4096, p_max=pn[-1]) * S[None, :]
    assert rel(U.cpu().numpy(), Uo) < 1e-12
    _, tro = K.hermitian_ifft(Uo, 2048, gamma, 64.0)
    assert rel(g, tro) < 1e-12


def test_bessel_sum_linearity_and_zero_offset():
    rng = np.random.default_rng(9)
    np_, nw = 300, 129
    p = torch.arange(1, np_ + 1, dtype=torch.float64) * (1.2 / np_)
    offsets = torch.tensor([0.0, 2.5, 11.0], dtype=torch.float64)
    R1 = torch.from_numpy(rng.standard_normal((np_, nw - 1)) + 1j * rng.standard_normal((np_, nw - 1))).cuda()
    R2 = torch.from_numpy(rng.standard_normal((np_, nw - 1)) + 1j * rng.standard_normal((np_, nw - 1))).cuda()
    args = (p, 1.2 / np_, 1.2, 1.0, 2 * np.pi / 64.0, offsets)
    U1 = engine.bessel_sum(R1, *args).cpu().numpy()
    U2 = engine.bessel_sum(R2, *args).cpu().numpy()
    U12 = engine.bessel_sum(R1 + R2, *args).cpu().numpy()
    assert rel(U12, U1 + U2) < 1e-13
    # r = 0: J0 = 1, U = sum_j R taper p w dp
    pn = p.numpy()
    w = np.arange(1, nw) * 2 * np.pi / 64.0
    taper = 0.5 * (1 + np.cos(np.pi * pn / 1.2))
    expect = ((R1.cpu().numpy() * ((taper * pn)[:, None] * w[None, :] * (1.2 / np_))).sum(0))
    assert rel(U1[0], expect) < 1e-13
    U_o = K.bessel_sum(R1.cpu().numpy(), pn, w, offsets.numpy(), 1.2 / np_, p_max=1.2)
    assert rel(U1, U_o) < 1e-12


@pytest.mark.parametrize("kind", ["offset_grid", "irregular"])
def test_bessel_sum_uniform_and_irregular_slownesses(kind):
    """The sum against the oracle for slownesses p0 + j dp with p0 != dp
    (chunks of uniform slownesses rotate Hankel's phase by w r dp from one
    slowness to the next) and for sorted random slownesses (no fixed step:
    every J0 is evaluated directly), over the small and large regimes."""
    rng = np.random.default_rng(21)
    np_, nw, dp = 300, 97, 0.9 / 300
    if kind == "offset_grid":
        pn = 0.05 + np.arange(np_) * dp
    else:
        pn = np.sort(rng.uniform(0.01, 0.95, np_))
    offsets = np.array([0.0, 0.7, 3.0, 12.5, 40.0])
    dw = 2 * np.pi / 16.0
    Rn = rng.standard_normal((np_, nw - 1)) + 1j * rng.standard_normal((np_, nw - 1))
    U = engine.bessel_sum(torch.from_numpy(Rn).cuda(), torch.from_numpy(pn), dp, pn[-1], 1.0, dw,
                          torch.from_numpy(offsets)).cpu().numpy()
    w = np.arange(1, nw) * dw
    U_o = K.bessel_sum(Rn, pn, w, offsets, dp, p_max=pn[-1])
    assert rel(U, U_o) < 1e-12


def test_full_c2_gather_properties():
    """Full gather workload: 20 layers, 8192 p, nt = 2048, 100 offsets."""
    cfg = gather_config_c2()
    dev = torch.device("cuda")
    layers = engine.layers_tensor(cfg.model, dev)
    off = torch.from_numpy(cfg.offsets).to(dev)
    w = np.arange(1, cfg.nw) * (2 * np.pi / cfg.T)
    S = torch.from_numpy(ricker_spectrum(w, cfg.ricker_omega_max)).to(dev)
    gamma = np.pi / cfg.T
    t1, R, U = gather_device(layers, off, cfg.T, cfg.nw, cfg.np_slow, cfg.p_max, gamma, S,
                             True, keep=True)
    t2 = gather_device(layers, off, cfg.T, cfg.nw, cfg.np_slow, cfg.p_max, gamma, S, True)
    t3 = gather_device(layers, off, cfg.T, cfg.nw, cfg.np_slow, cfg.p_max, gamma, 2 * S, True)
    torch.cuda.synchronize()
    assert torch.equal(t1, t2), "gather is not bitwise reproducible"
    assert torch.equal(t3, 2 * t1), "gather is not linear in the source spectrum"
    tr = t1.cpu().numpy()
    assert tr.shape == (100, 2048) and np.all(np.isfinite(tr))
    # one offset against the oracle's Bessel sum + FFT of the device R
    Rn = R.cpu().numpy()
    pn = np.arange(1, cfg.np_slow + 1) * (cfg.p_max / cfg.np_slow)
    Uo = K.bessel_sum(Rn, pn, w, cfg.offsets[[0, 57]], cfg.p_max / cfg.np_slow,
                      p_max=pn[-1]) * S.cpu().numpy()[None, :]
    assert rel(U.cpu().numpy()[[0, 57]], Uo) < 1e-12


def test_gather_edge_cases():
    m = seeded_model(4, 2)
    t, off, g = compute_gather_gpu(m, np.array([0.0, 3.0]), nw=2, np_slow=3, p_max=0.5)
    assert g.shape == (2, 4) and np.all(np.isfinite(g))
    with pytest.raises(ValueError):
        compute_gather_gpu(m, np.array([1.0]), nw=24, np_slow=8)  # nw not a power of two
