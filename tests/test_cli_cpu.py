"""The gather CLIs' flags, defaults and .npz output (kennett_gather.py:410-555,
kennett_gather_gpu.py:286-436) and the wiggle plot, on the CPU: the device
compute is replaced by the oracle's gather for the CLI plumbing."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import kennett_numpy as K
from seismicinversion_amd.kennett_gather import gather_cli, plot_gather


def test_gather_cli_flags_defaults_and_npz(tmp_path):
    seen = {}

    def compute(model, offsets, a):
        seen.update(vars(a))
        seen["model_layers"] = model.n_layers
        return K.gather(model, offsets, T=a.duration, nw=a.nw, np_slow=a.np_slow, p_max=a.p_max,
                        gamma=a.gamma, free_surface=a.free_surface)

    out = tmp_path / "g.png"
    gather_cli(["--r-min", "1.0", "--r-max", "2.0", "--dr", "0.5", "-n", "32", "--np", "64",
                "--no-plot", "-o", str(out), "--free-surface"],
               description="t", r_max=50.0, np_slow=4096, p_max=1.2, t_max=30.0,
               output="gather_gpu.png", workers_flag=False, compute=compute)
    d = np.load(tmp_path / "g.npz")
    assert set(d.files) == {"time", "offsets", "gather"}
    np.testing.assert_array_equal(d["offsets"], [1.0, 1.5, 2.0])
    assert d["gather"].shape == (3, 64) and d["time"].shape == (64,)
    assert seen["p_max"] == 1.2 and seen["duration"] == 64.0 and seen["gamma"] is None
    assert seen["free_surface"] and seen["model_layers"] == 5 and seen["device"] is None
    t, off, g = K.gather(K_model(), np.array([1.0, 1.5, 2.0]), T=64.0, nw=32, np_slow=64,
                         p_max=1.2, free_surface=True)
    np.testing.assert_array_equal(d["gather"], g)


def K_model():
    from seismicinversion_amd.workloads import default_ocean_crust_model

    return default_ocean_crust_model()


def test_cpu_cli_accepts_workers_flag(tmp_path):
    got = {}

    def compute(model, offsets, a):
        got["workers"] = a.workers
        return np.arange(4.0), offsets, np.zeros((len(offsets), 4))

    gather_cli(["-j", "3", "--no-plot", "-o", str(tmp_path / "x.png"), "--r-max", "1.0"],
               description="t", r_max=20.0, np_slow=2048, p_max=0.8, t_max=None,
               output="seismogram_gather.png", workers_flag=True, compute=compute)
    assert got["workers"] == 3 and (tmp_path / "x.npz").exists()


@pytest.mark.parametrize("norm", ["global", "trace"])
def test_plot_gather_writes_png(tmp_path, norm):
    pytest.importorskip("matplotlib")
    t = np.linspace(0, 4, 256)
    off = np.array([0.5, 1.0, 1.5])
    g = np.sin(8 * t)[None, :] * np.array([1.0, 0.5, 1e-9])[:, None]
    out = tmp_path / f"{norm}.png"
    plot_gather(t, off, g, out, t_max=3.0, norm=norm)
    assert out.stat().st_size > 1000
    with pytest.raises(ValueError):
        plot_gather(t, off, g, out, norm="bad")
