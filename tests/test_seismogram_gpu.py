"""compute_seismogram (kennett_seismogram.py:57-150) on the device against the
oracle's reflectivity times the source through numpy's FFT, and the CLIs."""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import kennett_numpy as K
from parity import ulp_tolerance
from seismicinversion_amd import kennett_gather_gpu, kennett_seismogram
from seismicinversion_amd.kennett_gather_gpu import compute_gather_gpu
from seismicinversion_amd.kennett_seismogram import compute_seismogram
from seismicinversion_amd.workloads import default_ocean_crust_model, seeded_model

pytestmark = pytest.mark.gpu


def oracle_seismogram(model, p, T, nw, eps, fs):
    dw = 2 * np.pi / T
    w = np.arange(1, nw) * dw
    R = K.reflectivity_batch(model, np.array([p]), w, fs)
    Y = R * K.ricker_spectrum(w, nw * dw)[None, :]
    t, tr = K.hermitian_ifft(Y, nw, -eps, T)
    return t, tr[0], ulp_tolerance(model, np.array([p]), w, fs)[0]


@pytest.mark.parametrize("p,eps,fs,nw", [(0.2, 0.0, False, 2048), (0.1, 0.05, True, 1024),
                                          (0.45, 0.0, True, 512), (0.7, 0.02, False, 256)])
def test_seismogram_vs_oracle(p, eps, fs, nw):
    for m in (default_ocean_crust_model(), seeded_model(12, 4)):
        t, s = compute_seismogram(m, p, T=64.0, nw=nw, eps=eps, free_surface=fs)
        to, so, tol = oracle_seismogram(m, p, 64.0, nw, eps, fs)
        np.testing.assert_array_equal(t, to)
        err = np.max(np.abs(s - so)) / np.max(np.abs(so))
        assert err <= max(1e-12, tol), (err, tol)


def test_seismogram_errors_and_cli(tmp_path):
    with pytest.raises(ValueError):
        compute_seismogram(default_ocean_crust_model(), 0.2, nw=1000)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        kennett_seismogram.main(["-p", "0.25", "-n", "256", "--no-plot"])
    finally:
        os.chdir(cwd)
    d = np.loadtxt(tmp_path / "seismogram_p0.2500.txt")
    t, s = compute_seismogram(default_ocean_crust_model(), 0.25, nw=256)
    assert d.shape == (512, 2)
    np.testing.assert_allclose(d[:, 1], s, rtol=1e-8, atol=1e-8 * np.abs(s).max())


def test_gather_gpu_cli_npz(tmp_path):
    out = tmp_path / "g.png"
    kennett_gather_gpu.main(["--r-max", "2.0", "-n", "64", "--np", "128", "--no-plot",
                             "-o", str(out), "--free-surface"])
    d = np.load(tmp_path / "g.npz")
    t, off, g = compute_gather_gpu(default_ocean_crust_model(), np.arange(0.5, 2.25, 0.5), nw=64,
                                   np_slow=128, free_surface=True)
    np.testing.assert_array_equal(d["offsets"], off)
    np.testing.assert_array_equal(d["gather"], g)
    np.testing.assert_array_equal(d["time"], t)
