"""The serving stream at its full size distribution on the GPU: Poisson
arrivals of forward requests with layer counts log-uniform in [3, 100],
slownesses log-uniform in [16, 2048] and frequency counts in {256, 1024,
4096} (plus the largest and the smallest corner of that box), and 20 % of
5-iteration Newton-LM inversions, served by one continuous-batching server
with host and device-resident results.  Every forward result equals the
same request computed alone bit for bit, two rows of every request match
the oracle to the parity bar, and every inversion equals its solo run
(the reference's batched call, kennett_reflectivity_gpu.py:88, and its
inversion, taup_inversion.py:262-470)."""

from __future__ import annotations

import time

import numpy as np
import pytest
import torch

from parity import accuracy, assert_accurate
from seismicinversion_amd import batch, engine
from seismicinversion_amd.batch import ModelSpec
from seismicinversion_amd.serve import DeviceBackend, KennettServer, stream_requests
from seismicinversion_amd.taup_inversion import invert_taup
from seismicinversion_amd.workloads import StreamItem, seeded_model, serving_stream

pytestmark = pytest.mark.gpu


def _stream():
    items = serving_stream(500.0, 70, seed=11, lm_fraction=0.2)
    t_end = items[-1].t
    # the two corners of the size box
    items += [StreamItem(t_end + 1e-3, "forward", 100, 2048, 4096, 5001),
              StreamItem(t_end + 2e-3, "forward", 3, 16, 256, 5002)]
    reqs = stream_requests(items)
    for it, (kind, spec) in zip(items, reqs):
        if kind == "forward":  # damped grids and the free surface on some
            spec.gamma = np.pi / 64.0 if it.seed % 3 == 0 else 0.0
            spec.free_surface = bool(it.seed % 2)
    return items, reqs


def test_full_size_stream_served_equals_alone_and_oracle():
    items, reqs = _stream()
    kinds = [k for k, _ in reqs]
    n_inv = kinds.count("invert")
    fwd_sizes = [(s.n_layers, len(s.p), s.nw) for k, s in reqs if k == "forward"]
    print(f"{len(reqs)} requests, {n_inv} inversions; forward nl up to "
          f"{max(a for a, _, _ in fwd_sizes)}, np up to {max(b for _, b, _ in fwd_sizes)}, "
          f"nw {sorted(set(c for _, _, c in fwd_sizes))}", flush=True)
    assert n_inv >= 8 and max(c for _, _, c in fwd_sizes) == 4096
    futs = []
    with KennettServer(backend=DeviceBackend(n_streams=3)) as srv:
        t0 = time.perf_counter()
        for k, (it, (kind, spec)) in enumerate(zip(items, reqs)):
            d = it.t - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            futs.append(srv.submit_forward(spec, on_device=(k % 4 == 1)) if kind == "forward"
                        else srv.submit_invert(spec))
        outs = [f.result(timeout=600) for f in futs]
    st = srv.stats()
    print("batches", st["batches"], "latency", st["latency_forward"], st["latency_invert"],
          flush=True)
    assert st["failed"] == 0 and st["batches"] < len(reqs) - n_inv
    assert all(r["nonfinite"] == 0 for r in srv.records if r["kind"] == "forward")
    for (kind, spec), out in zip(reqs, outs):
        if kind != "forward":
            continue
        s = spec
        if isinstance(out, torch.Tensor):
            assert out.is_cuda
            out = out.cpu().numpy()
        alone = engine.forward_grid(engine.layers_tensor(s.model), torch.from_numpy(s.p), s.k0,
                                    s.dw, s.gamma, s.nw, s.free_surface, block=64,
                                    omega_per_item=batch.omega_per_lane(s.nw, 64, 8))
        np.testing.assert_array_equal(out, alone.cpu().numpy())
        rows = np.unique(np.random.default_rng(len(s.p) * s.nw).integers(0, len(s.p), 2))
        w = (s.k0 + np.arange(s.nw)) * s.dw + 1j * s.gamma
        assert_accurate(out[rows], accuracy(s.model, s.p[rows], w, s.free_surface),
                        f"served nl={s.n_layers} np={len(s.p)} nw={s.nw}")
    for (kind, spec), res in zip(reqs, outs):
        if kind != "invert":
            continue
        solo = invert_taup(spec.true_model, list(spec.p), nfreq=spec.nfreq,
                           perturbation=spec.perturbation, max_iter=spec.iterations,
                           seed=spec.seed, tol=0.0, hessian_mode=spec.hessian_mode)
        assert res.misfit_history == solo.misfit_history
        assert res.param_error_history == solo.param_error_history
        np.testing.assert_array_equal(res.recovered_model.alpha, solo.recovered_model.alpha)
        np.testing.assert_array_equal(res.recovered_model.thickness,
                                      solo.recovered_model.thickness)


def test_device_results_stay_valid_after_the_server_stops():
    """A device-resident result is a view of its batch's output; it stays
    valid after the server has stopped, and host results are ordinary arrays
    the caller may keep and modify."""
    spec = ModelSpec(model=seeded_model(30, 9), p=np.linspace(0.01, 0.7, 300), nw=1024)
    with KennettServer(backend=DeviceBackend(n_streams=2)) as srv:
        f_dev = srv.submit_forward(spec, on_device=True)
        f_host = srv.submit_forward(ModelSpec(model=spec.model, p=spec.p, nw=spec.nw))
        R_dev, R_host = f_dev.result(timeout=60), f_host.result(timeout=60)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(R_dev.cpu().numpy(), R_host)
    keep = R_host.copy()
    R_host[:] = 0  # the caller's memory
    np.testing.assert_array_equal(R_dev.cpu().numpy(), keep)
