"""Launch-shape autotuner for the reflectivity kernels.

A shape is (work-items per workgroup, frequencies per work-item).  For a
workload signature — the kernel, the largest layer count rounded up to a
power of two and the frequency count — the tuner times every candidate shape
on the device once (HIP events on the current stream, median of a few
launches) and caches the fastest; later calls with the same signature reuse
it.  The cache is per process, guarded by a lock, and can be saved to and
loaded from JSON so a service can start warm.
"""

from __future__ import annotations

import json
import threading

import numpy as np
import torch

__all__ = ["CANDIDATES", "multi_shape", "forward_shape", "cache", "save", "load", "clear"]

CANDIDATES = [(64, 8), (128, 8), (64, 4), (128, 4), (256, 4), (256, 2), (64, 2), (64, 1)]
_LOCK = threading.Lock()
_CACHE: dict[str, tuple[int, int]] = {}


def _bucket(n: int) -> int:
    return 1 << max(1, int(np.ceil(np.log2(max(n, 2)))))


def _time(fn, reps: int = 3) -> float:
    fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        out.append(a.elapsed_time(b))
    return float(np.median(out))


def multi_shape(specs, dev=None) -> tuple[int, int]:
    """Best (block, omega_per_item) for a kh_forward_multi batch like `specs`."""
    from . import batch

    key = "multi:%d:%d:%d" % (_bucket(max(s.n_layers for s in specs)),
                              _bucket(max(s.nw for s in specs)), _bucket(len(specs)))
    with _LOCK:
        if key in _CACHE:
            return _CACHE[key]
    best, best_t = None, float("inf")
    for block, wpt in CANDIDATES:
        plan = batch.plan_multi(specs, block, wpt, dev)
        t = _time(lambda: batch.launch(plan))
        if t < best_t:
            best, best_t = (block, wpt), t
    with _LOCK:
        _CACHE[key] = best
    return best


def forward_shape(layers, p, k0, dw, gamma, nw, free_surface=False) -> tuple[int, int]:
    """Best (block, omega_per_item) for kh_forward_grid on this workload shape."""
    from . import engine

    n_p = p.numel() if isinstance(p, torch.Tensor) else int(np.asarray(p).size)
    key = "grid:%d:%d:%d" % (_bucket(layers.shape[0]), _bucket(n_p), _bucket(nw))
    with _LOCK:
        if key in _CACHE:
            return _CACHE[key]
    out = torch.empty((n_p, nw), dtype=torch.complex128, device=layers.device)
    best, best_t = None, float("inf")
    for block, wpt in CANDIDATES:
        t = _time(lambda: engine.forward_grid(layers, p, k0, dw, gamma, nw, free_surface, out,
                                              block, wpt))
        if t < best_t:
            best, best_t = (block, wpt), t
    with _LOCK:
        _CACHE[key] = best
    return best


def cache() -> dict[str, tuple[int, int]]:
    with _LOCK:
        return dict(_CACHE)


def clear() -> None:
    with _LOCK:
        _CACHE.clear()


def save(path: str) -> None:
    with open(path, "w") as f:
        json.dump({k: list(v) for k, v in cache().items()}, f, indent=1)


def load(path: str) -> None:
    with open(path) as f:
        data = json.load(f)
    with _LOCK:
        _CACHE.update({k: (int(v[0]), int(v[1])) for k, v in data.items()})
