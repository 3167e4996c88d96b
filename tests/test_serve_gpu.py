"""The continuous-batching server on the GPU: requests submitted from several
threads at staggered times are batched into persistent launches, and every
result is bitwise what the same request gives alone."""

from __future__ import annotations

import threading
import time

import numpy as np
import pytest
import torch

from seismicinversion_amd import batch, engine
from seismicinversion_amd.batch import ModelSpec
from seismicinversion_amd.serve import DeviceBackend, InvertSpec, KennettServer
from seismicinversion_amd.taup_inversion import invert_taup
from seismicinversion_amd.workloads import seeded_model, serving_stream, uniform_slownesses

pytestmark = pytest.mark.gpu


def spec_of(item):
    return ModelSpec(model=seeded_model(item.n_layers, item.seed),
                     p=uniform_slownesses(item.n_p, 0.8), nw=item.n_omega,
                     gamma=np.pi / 64.0 if item.seed % 3 == 0 else 0.0,
                     free_surface=bool(item.seed % 2))


def test_results_identical_alone_and_in_the_mix():
    items = [it for it in serving_stream(2000.0, 60, seed=7, lm_fraction=0.0,
                                         np_range=(16, 256), omegas=(256, 1024))]
    specs = [spec_of(it) for it in items]
    backend = DeviceBackend(n_streams=2, shape=(64, 8))
    futs = [None] * len(specs)
    with KennettServer(backend=backend) as srv:
        inv = srv.submit_invert(InvertSpec(true_model=seeded_model(5, 3),
                                           p=np.array([0.1, 0.2, 0.3]), nfreq=32, iterations=3))

        def client(idx):
            for i in idx:
                time.sleep(items[i].t / 4)
                futs[i] = srv.submit_forward(specs[i], on_device=(i % 4 == 0))

        ths = [threading.Thread(target=client, args=(list(range(k, len(specs) - 10, 3)),))
               for k in range(3)]
        futs[len(specs) - 10:] = srv.submit_forward_many(specs[len(specs) - 10:])
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        outs = [f.result(timeout=120) for f in futs]
        inv_res = inv.result(timeout=300)
    st = srv.stats()
    print("batches", st["batches"], "for", len(specs), "forward requests; latency",
          st["latency_forward"])
    assert st["batches"] < len(specs)
    assert all(r["nonfinite"] == 0 for r in srv.records if r["kind"] == "forward")
    for s, R in zip(specs, outs):
        alone = engine.forward_grid(engine.layers_tensor(s.model), torch.from_numpy(s.p), s.k0,
                                    s.dw, s.gamma, s.nw, s.free_surface, block=64,
                                    omega_per_item=batch.omega_per_lane(s.nw, 64, 8)
                                    ).cpu().numpy()
        if isinstance(R, torch.Tensor):  # device-resident result
            assert R.is_cuda
            R = R.cpu().numpy()
        np.testing.assert_array_equal(R, alone)
    solo = invert_taup(seeded_model(5, 3), [0.1, 0.2, 0.3], nfreq=32, max_iter=3, tol=0.0,
                       hessian_mode="hvp", perturbation=0.05)
    assert inv_res.misfit_history == solo.misfit_history
