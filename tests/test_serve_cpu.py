"""Scheduling logic of the continuous-batching server on the CPU: a host
backend computes forward requests with the oracle and counts launches, so
admission by memory, batching, deadlines and per-request failures are
exercised without a GPU."""

from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from oracle import kennett_numpy as K
from seismicinversion_amd.batch import ModelSpec
from seismicinversion_amd.layer_model import LayerModel
from seismicinversion_amd.serve import InvertSpec, KennettServer
from seismicinversion_amd.workloads import seeded_model


class HostBackend:
    """Runs each batch on a worker thread with the NumPy oracle; a batch
    becomes 'done' when its thread ends (like an event on a stream)."""

    slots = 1

    def __init__(self, delay=0.0, poison=None):
        self.launches = []
        self.delay = delay
        self.poison = poison or set()

    def free_bytes(self):
        return 1 << 30

    @staticmethod
    def forward_bytes(spec):
        return 16 * len(spec.p) * spec.nw

    @staticmethod
    def invert_bytes(spec):
        return 1000

    def launch_forward(self, specs, to_host=None):
        out = {}

        def run():
            time.sleep(self.delay)
            res = []
            for s in specs:
                w = (s.k0 + np.arange(s.nw)) * s.dw + 1j * s.gamma
                R = K.reflectivity_batch(s.model, s.p, w, s.free_surface)
                if id(s) in self.poison:
                    R[0, 0] = np.nan
                res.append(R)
            out["res"] = res

        t = threading.Thread(target=run)
        t.start()
        self.launches.append(len(specs))
        return {"t": t, "out": out}

    def done(self, h):
        return not h["t"].is_alive()

    def results(self, h):
        h["t"].join()
        return h["out"]["res"], None

    def invert(self, spec):
        return {"iterations": spec.iterations}


def fwd(nl, n_p, nw, seed=1, fs=False):
    return ModelSpec(model=seeded_model(nl, seed), p=np.linspace(0.05, 0.6, n_p), nw=nw,
                     free_surface=fs)


def test_requests_are_batched_and_results_match_alone():
    be = HostBackend(delay=0.05)
    specs = [fwd(3 + i % 7, 4 + i % 5, 16 + i, seed=i, fs=bool(i % 2)) for i in range(24)]
    with KennettServer(backend=be, hbm_budget=1 << 30) as srv:
        futs = []
        for s in specs:
            futs.append(srv.submit_forward(s))
            time.sleep(0.002)
        outs = [f.result(timeout=60) for f in futs]
    assert len(be.launches) < len(specs), "requests were not batched"
    assert sum(be.launches) == len(specs)
    for s, R in zip(specs, outs):
        w = (s.k0 + np.arange(s.nw)) * s.dw
        np.testing.assert_array_equal(R, K.reflectivity_batch(s.model, s.p, w, s.free_surface))
    st = srv.stats()
    assert st["completed"] == len(specs) and st["latency_forward"]["count"] == len(specs)
    assert st["latency_forward"]["p99_ms"] >= st["latency_forward"]["p50_ms"] > 0


def test_memory_budget_queues_and_rejects():
    be = HostBackend(delay=0.02)
    small = [fwd(4, 8, 32, seed=i) for i in range(6)]
    need = HostBackend.forward_bytes(small[0])
    with KennettServer(backend=be, hbm_budget=2 * need) as srv:
        futs = [srv.submit_forward(s) for s in small]
        big = srv.submit_forward(fwd(4, 64, 512))
        with pytest.raises(MemoryError):
            big.result(timeout=30)
        for f in futs:
            f.result(timeout=60)
        assert srv.in_use <= 2 * need
    assert max(be.launches) <= 2, "admission exceeded the memory budget"


def test_invalid_model_fails_alone_and_others_proceed():
    be = HostBackend()
    bad = ModelSpec(model=LayerModel.from_arrays(alpha=[1.5], beta=[0.0], rho=[1.0],
                                                 thickness=[np.inf], Q_alpha=[1e4],
                                                 Q_beta=[1e4]), p=np.array([0.1]), nw=8)
    with KennettServer(backend=be, hbm_budget=1 << 30) as srv:
        f_bad = srv.submit_forward(bad)
        f_ok = srv.submit_forward(fwd(5, 3, 8))
        with pytest.raises(ValueError):
            f_bad.result(timeout=30)
        assert f_ok.result(timeout=30).shape == (3, 8)
    assert srv.stats()["failed"] == 1


def test_nonfinite_flag_and_deadline():
    spec = fwd(5, 3, 8)
    be = HostBackend(poison={id(spec)})
    with KennettServer(backend=be, hbm_budget=1 << 30) as srv:
        srv.submit_forward(spec).result(timeout=30)
    assert srv.records[-1]["nonfinite"] == 1
    # a request that cannot be admitted before its deadline is cancelled
    be = HostBackend(delay=0.3)
    first = fwd(4, 8, 32)
    need = HostBackend.forward_bytes(first)
    with KennettServer(backend=be, hbm_budget=need) as srv:
        f1 = srv.submit_forward(first)
        time.sleep(0.05)
        f2 = srv.submit_forward(fwd(4, 8, 32, seed=3), deadline_s=0.05)
        with pytest.raises(TimeoutError):
            f2.result(timeout=30)
        f1.result(timeout=30)


def test_inversions_run_beside_forward_batches():
    be = HostBackend()
    with KennettServer(backend=be, hbm_budget=1 << 30) as srv:
        fi = srv.submit_invert(InvertSpec(true_model=seeded_model(4, 2), p=np.array([0.1, 0.2]),
                                          nfreq=16, iterations=3))
        ff = srv.submit_forward(fwd(4, 2, 8))
        assert fi.result(timeout=30) == {"iterations": 3}
        ff.result(timeout=30)
    assert srv.stats()["latency_invert"]["count"] == 1


def test_submit_many_batches_and_fails_bad_requests_alone():
    be = HostBackend(delay=0.02)
    bad = ModelSpec(model=LayerModel.from_arrays(alpha=[1.5], beta=[0.0], rho=[1.0],
                                                 thickness=[np.inf], Q_alpha=[1e4],
                                                 Q_beta=[1e4]), p=np.array([0.1]), nw=8)
    specs = [fwd(3 + i % 4, 3, 16, seed=i) for i in range(12)]
    with KennettServer(backend=be, hbm_budget=1 << 30) as srv:
        futs = srv.submit_forward_many(specs[:6] + [bad] + specs[6:])
        outs = [f.result(timeout=60) if i != 6 else None for i, f in enumerate(futs)]
        with pytest.raises(ValueError):
            futs[6].result(timeout=30)
    assert len(futs) == 13 and len(be.launches) <= 2 and sum(be.launches) == 12
    for s, R in zip(specs, [o for i, o in enumerate(outs) if i != 6]):
        w = (s.k0 + np.arange(s.nw)) * s.dw
        np.testing.assert_array_equal(R, K.reflectivity_batch(s.model, s.p, w, s.free_surface))
    assert srv.outstanding == 0


def test_oversized_and_malformed_models_fail_alone():
    """A 241-layer model (above the kernels' 240-layer cap), an (nl, 5)
    table and an empty frequency grid each fail with ValueError at
    admission; the good requests submitted around them are batched and
    served."""
    deep = seeded_model(241, 4)
    bad_specs = [ModelSpec(model=deep, p=np.array([0.1]), nw=8),
                 ModelSpec(model=np.ones((4, 5)), p=np.array([0.1]), nw=8),
                 ModelSpec(model=seeded_model(5, 2), p=np.array([0.1]), nw=0)]
    good = [fwd(4 + i, 3, 8, seed=i) for i in range(4)]
    be = HostBackend(delay=0.02)
    with KennettServer(backend=be, hbm_budget=1 << 30) as srv:
        futs = srv.submit_forward_many([good[0], bad_specs[0], good[1], bad_specs[1],
                                        good[2], bad_specs[2], good[3]])
        bad_futs = [futs[1], futs[3], futs[5]]
        for f in bad_futs:
            with pytest.raises(ValueError):
                f.result(timeout=30)
        for f, s in zip([futs[0], futs[2], futs[4], futs[6]], good):
            R = f.result(timeout=30)
            w = (s.k0 + np.arange(s.nw)) * s.dw
            np.testing.assert_array_equal(R, K.reflectivity_batch(s.model, s.p, w, False))
    assert srv.stats()["failed"] == 3
    assert sum(be.launches) == 4


def test_inversion_above_the_derivative_cap_fails_alone():
    be = HostBackend()
    with KennettServer(backend=be, hbm_budget=1 << 30) as srv:
        f_bad = srv.submit_invert(InvertSpec(true_model=seeded_model(121, 1), p=np.array([0.1])))
        f_ok = srv.submit_invert(InvertSpec(true_model=seeded_model(5, 1), p=np.array([0.1])))
        with pytest.raises(ValueError):
            f_bad.result(timeout=30)
        assert f_ok.result(timeout=30) == {"iterations": 5}


def test_result_pool_blocks_return_when_the_last_view_dies():
    """The page-locked result pool's bookkeeping on a plain host buffer: a
    batch's block stays taken while any result view (or a NumPy view of
    one) is alive, returns when the last one is garbage, and freed
    neighbours merge back into one block."""
    import gc

    import torch

    from seismicinversion_amd.serve import PinnedResultPool, _Lease

    g = PinnedResultPool.GRANULE
    pool = PinnedResultPool(8 * g, buf=torch.zeros(8 * g, dtype=torch.uint8))
    b1, b2 = pool.take(3 * g - 5), pool.take(2 * g)
    assert b1 == (0, 3 * g) and b2 == (3 * g, 2 * g) and pool.free_bytes() == 3 * g
    assert pool.take(4 * g) is None  # no block that large is free
    l1, l2 = _Lease(pool, *b1), _Lease(pool, *b2)
    x = l1.array(0, 4, (2, 2))
    y = l1.array(4, 10, (3, 2))
    x[:] = 1 + 2j
    assert np.all(l1.c128[:4].numpy() == 1 + 2j)  # the view is the block's memory
    z = l2.array(0, 8, (8,))
    l1.drop()
    l2.drop()  # the batches let go; their views still hold the blocks
    assert pool.free_bytes() == 3 * g
    keep = y[1:]  # a NumPy view of a result keeps the block too
    del x, y
    gc.collect()
    assert pool.free_bytes() == 3 * g
    del keep
    gc.collect()
    assert pool.free_bytes() == 6 * g
    del z
    gc.collect()
    assert pool.free == [(0, 8 * g)]  # merged back into one block
