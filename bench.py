"""Benchmark: Kennett P-SV reflectivity served vs plain on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline workload (one "step"): N = 400 forward requests drawn from the
mixed-size serving distribution (seed 42: layer counts round(exp(U[ln 3,
ln 100])), slownesses round(exp(U[ln 16, ln 2048])) uniform in (0, 0.8] s/km,
frequency counts uniform over {256, 1024, 4096}, omega_k = k*2*pi/64),
submitted at once (``submit_forward_many``) to the continuous-batching
server and served with device-resident results.  Each step submits new
request objects, so validation and packing of every layer table are timed.
``value`` = p*omega*layer updates per second (one update = one interface
step of one (p, omega) lane).

Beside it, on the same requests:
* plain: the one-model-per-launch synchronous GPU path (one kh_forward_grid
  launch per request, synchronised), with device results and, through the
  kennett_reflectivity_batch drop-in, with host results;
* served with host results (pinned staging + copy-out);
* served / plain per size class (small: nl <= 20 and np*nw <= 2.6e5, where
  one request leaves most of the device idle; medium; large);
* an open-loop sweep over arrival rates (default 100, 200, 500, 1000, 2000,
  5000 req/s): Poisson arrivals, 20 % of them 5-iteration LM inversions
  (slownesses <= 32, frequencies <= 128), results to the host, admission
  deadlines; per rate p50/p99 latency from arrival to result for all
  forward requests, for 20-layer forward requests and for inversions,
  forward updates/s, completed and rejected counts and queue lengths;
* kernels: the gather workload's forward kernel (20 layers, 8192 p x 1023
  damped omega, free surface), the full gather (forward + J0 wavenumber sum +
  inverse FFT, 100 offsets, nt 2048, and the reference CLI's -n 2048 shape,
  nt 4096), the 256-model uneven batch (persistent launch vs one launch per
  model) and the 20-layer Newton-LM inversion (75 parameters, 128 p x 512
  omega, 10 iterations: the device loop eager and from a captured HIP graph,
  and the host loop).

``roofline``: the persistent multi-model kernel of the served steps, FP64
algorithmic FLOPs (340 per update) per launch over the average launch
duration (HIP events on the server's compute stream, where the batch
kernels run back to back).  ``cpu_baseline``: the NumPy oracle
(vectorised over slownesses) on the host cores, on evenly spaced slowness
rows of every request; ``single_core`` is one process on the middle row.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FLOP_PER_UPDATE = 340.0          # 41 cmul (6) + 25 cadd (2) + 4 cdiv (11)
PEAK_FP64_TFLOPS = 78.6          # MI355X FP64 vector
PEAK_HBM_GBS = 8000.0
UNIT = "p*omega*layer updates/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--rates", default="100,200,500,1000,2000,5000",
                    help="open-loop arrival rates (req/s), comma separated")
    ap.add_argument("--open-seconds", type=float, default=2.5,
                    help="arrival window of each open-loop rate")
    ap.add_argument("--forward-deadline", type=float, default=0.5,
                    help="open loop: a forward request not admitted within this many seconds "
                         "is rejected")
    ap.add_argument("--invert-deadline", type=float, default=2.0,
                    help="open loop: an inversion not started within this many seconds is "
                         "rejected")
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="processes of the all-core CPU baseline (the box's share is 16 cores)")
    ap.add_argument("--warmup-seconds", type=float, default=2.0,
                    help="untimed forward launches before anything is timed (clock ramp)")
    ap.add_argument("--section", choices=("all", "lm", "lmloop", "open", "served", "hessian",
                                          "gather"),
                    default="all",
                    help="'lm': only the 20-layer Newton-LM figures; 'hessian': only --steps "
                         "misfit + gradient + exact-Hessian evaluations of that LM after "
                         "--warmup untimed ones (for profiling its kernels); 'lmloop': that LM "
                         "run for --steps iterations on the device loop, after a --warmup "
                         "iteration run (for profiling a whole iteration's kernels); 'open': "
                         "only the open-loop rate sweep; 'served': only the headline served "
                         "steps and their roofline; 'gather': only the 20-layer gather "
                         "(8192 p x 1023 omega, 100 offsets), --steps runs after --warmup")
    ap.add_argument("--traffic-json", default=os.path.join(HERE, "profiles", "traffic_multi_served.json"),
                    help="HBM bytes per launch of the multi-model kernel from rocprofv3 --pmc")
    return ap.parse_args()


def warm_clocks(torch, dev, seconds):
    """Run the forward kernel for a while first: the first launches of a
    process run before the board has raised its clocks."""
    from seismicinversion_amd import engine
    from seismicinversion_amd.workloads import seeded_model, uniform_slownesses

    L = engine.layers_tensor(seeded_model(20, 42), dev)
    p = torch.from_numpy(uniform_slownesses(4096, 1.2)).to(dev)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        engine.forward_grid(L, p, 1.0, 2 * np.pi / 64.0, 0.0, 1024)
        torch.cuda.synchronize()


def keep_busy(fn, torch, seconds=0.4):
    """Repeat fn() for a moment so a kernel timed after host-bound phases
    runs at the clocks a busy device holds."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        fn()
        torch.cuda.synchronize()


def timed(fn, reps, torch):
    """Mean wall seconds of fn() over reps, synchronised on both sides."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def event_ms(fn, reps, torch):
    """Mean HIP-event milliseconds of fn() on the current stream (after a
    moment of untimed repeats)."""
    keep_busy(fn, torch)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in ev]))


# Work shared with forked CPU workers (set before the pool is created).
_CPU = {}


def _cpu_rows(chunk):
    """Worker: the oracle on `rows` evenly spaced slowness rows of requests lo..hi."""
    from oracle import kennett_numpy as K

    lo, hi, rows = chunk
    upd = 0
    for s in _CPU["specs"][lo:hi]:
        idx = np.unique(np.linspace(0, len(s.p) - 1, min(rows, len(s.p))).round().astype(int))
        w = (s.k0 + np.arange(s.nw)) * s.dw + 1j * s.gamma
        K.reflectivity_batch(s.model, s.p[idx], w, s.free_surface)
        upd += len(idx) * s.nw * (s.n_layers - 1)
    return upd


def _cpu_gather_part(rng):
    """Worker: reflectivity and partial wavenumber sum of slownesses j0..j1 of
    the gather workload (the per-chunk work of kennett_gather.py:246-280)."""
    from oracle import kennett_numpy as K

    g = _CPU["gather"]
    j0, j1 = rng
    ps = g["p"][j0:j1]
    R = K.reflectivity_batch(g["model"], ps, g["w"] + 1j * g["gamma"], True)
    return K.bessel_sum(R, ps, g["w"], g["offsets"], g["dp"], p_max=g["p"][-1])


def host_cpu():
    """The host's CPU model and logical core count (from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cores": os.cpu_count()}


def cpu_baselines(specs, workers):
    """The oracle timed on the host before the GPU is touched (workers are
    forked): one process on the middle slowness row of every request, all
    workers on `workers` rows of every request, and the whole gather
    workload split over the workers by slowness chunks."""
    import multiprocessing as mp

    from oracle import kennett_numpy as K
    from seismicinversion_amd.workloads import gather_config_c2

    s0 = specs[0]
    K.reflectivity_batch(s0.model, s0.p[:1], (s0.k0 + np.arange(8)) * s0.dw, False)  # warm-up
    upd1 = 0
    t0 = time.perf_counter()
    for s in specs:
        mid = len(s.p) // 2
        w = (s.k0 + np.arange(s.nw)) * s.dw + 1j * s.gamma
        K.reflectivity_batch(s.model, s.p[mid:mid + 1], w, s.free_surface)
        upd1 += s.nw * (s.n_layers - 1)
    dt1 = time.perf_counter() - t0

    cfg = gather_config_c2()
    dw = 2 * np.pi / cfg.T
    p = np.arange(1, cfg.np_slow + 1) * (cfg.p_max / cfg.np_slow)
    _CPU["specs"] = specs
    _CPU["gather"] = dict(model=cfg.model, p=p, w=np.arange(1, cfg.nw) * dw, gamma=np.pi / cfg.T,
                          offsets=cfg.offsets, dp=cfg.p_max / cfg.np_slow)
    chunks = [(i, min(len(specs), i + 5), workers) for i in range(0, len(specs), 5)]
    ranges = [(j, min(cfg.np_slow, j + 32)) for j in range(0, cfg.np_slow, 32)]
    pool = mp.get_context("fork").Pool(workers)
    try:
        pool.map(_cpu_rows, [(0, 1, 1)] * workers)  # start-up
        t0 = time.perf_counter()
        updn = sum(pool.map(_cpu_rows, chunks, chunksize=1))
        dtn = time.perf_counter() - t0
        t0 = time.perf_counter()
        parts = pool.map(_cpu_gather_part, ranges, chunksize=1)
        U = parts[0].copy()
        for part in parts[1:]:  # chunks added in slowness order
            U += part
        w = np.arange(1, cfg.nw) * dw
        S = K.ricker_spectrum(w, cfg.ricker_omega_max)
        _, traces = K.hermitian_ifft(U * S[None, :], cfg.nw, np.pi / cfg.T, cfg.T)
        dtg = time.perf_counter() - t0
    finally:
        # close and join (not terminate): every worker exits before the GPU
        # part starts and none outlives the run
        pool.close()
        pool.join()
    cpu = {"value": updn / dtn, "unit": UNIT, "cores": workers, "kind": "port",
           "sample": f"{workers} evenly spaced slowness rows of each of the {len(specs)} requests "
                     f"({updn:.3e} updates), NumPy oracle in {workers} forked processes "
                     f"(one thread each), {dtn:.1f} s; the oracle is vectorised over "
                     "slownesses in the reference's (np, nw, 2, 2) layout, while the "
                     "reference loops in Python over slownesses and layers "
                     "(kennett_reflectivity_gpu.py:142-147, 157-191), so the reference "
                     "itself would be slower and GPU/CPU ratios are conservative",
           "single_core": {"value": upd1 / dt1, "cores": 1,
                           "sample": f"middle slowness row of each request ({upd1:.3e} updates), "
                                     f"{dt1:.1f} s"},
           "host": host_cpu()}
    gather_cpu = {"s": dtg, "cores": workers,
                  "what": "whole gather workload: reflectivity + wavenumber sum per 32-slowness "
                          "chunk in the workers, chunks added in order, FFT in the parent"}
    # the reference's CPU case: 5 layers, 64 slownesses in (0, 0.6], 1023
    # frequencies, one process (repeated for at least a second)
    m1, p1, w1 = c1_inputs()
    K.reflectivity_batch(m1, p1, w1, False)
    n1, t0 = 0, time.perf_counter()
    while True:
        R1 = K.reflectivity_batch(m1, p1, w1, False)
        n1 += 1
        if time.perf_counter() - t0 > 1.0:
            break
    c1_cpu = {"ms": (time.perf_counter() - t0) / n1 * 1e3, "cores": 1, "R": R1}
    return cpu, gather_cpu, traces, c1_cpu


def c1_inputs():
    from seismicinversion_amd.workloads import omega_grid, seeded_model, uniform_slownesses

    return seeded_model(5, 42), uniform_slownesses(64, 0.6), omega_grid(1023)


def c1_figure(torch, dev, c1_cpu):
    """The 5-layer, 64 p x 1023 omega case on the device: the
    kennett_reflectivity_batch drop-in (host arrays in and out) and the
    device-resident launch, against the NumPy oracle's time on one core."""
    from seismicinversion_amd import engine
    from seismicinversion_amd.kennett_reflectivity_gpu import kennett_reflectivity_batch

    m1, p1, w1 = c1_inputs()
    upd = len(p1) * len(w1) * (m1.n_layers - 1)
    drop = lambda: kennett_reflectivity_batch(m1, p1, w1, False)  # noqa: E731
    R = drop()
    for _ in range(20):
        drop()
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < 0.5:
        R = drop()
        n += 1
    t_drop = (time.perf_counter() - t0) / n
    L = engine.layers_tensor(m1, dev)
    P = torch.from_numpy(p1).to(dev)
    Rd = torch.empty((len(p1), len(w1)), dtype=torch.complex128, device=dev)
    dev_call = lambda: engine.forward_grid(L, P, 1.0, 2 * np.pi / 64.0, 0.0, len(w1), False, Rd)  # noqa: E731
    dev_call()
    ms = event_ms(dev_call, 20, torch)
    ref = c1_cpu["R"]
    diff = float(np.max(np.abs(R - ref).max(axis=1) / np.abs(ref).max(axis=1)))
    return {"workload": "5 layers (seed 42), 64 p in (0, 0.6], 1023 omega, free surface off",
            "updates": upd, "cpu_oracle_one_core_ms": c1_cpu["ms"],
            "dropin_host_arrays_ms": t_drop * 1e3, "device_resident_ms": ms,
            "dropin_over_cpu": c1_cpu["ms"] / (t_drop * 1e3),
            "max_row_rel_diff_vs_oracle": diff}


def gather_figures(torch, dev, steps, warmup):
    """The 20-layer gather (8192 p x 1023 omega, 100 offsets, nt 2048), inputs
    resident, timed over `steps` runs after `warmup` untimed ones (for
    profiling its kernels)."""
    from seismicinversion_amd import engine
    from seismicinversion_amd.kennett_gather_gpu import gather_device
    from seismicinversion_amd.source import ricker_spectrum
    from seismicinversion_amd.workloads import gather_config_c2

    cfg = gather_config_c2()
    L = engine.layers_tensor(cfg.model, dev)
    dw, gamma = 2 * np.pi / cfg.T, np.pi / cfg.T
    off = torch.from_numpy(cfg.offsets).to(dev)
    S = torch.from_numpy(ricker_spectrum(np.arange(1, cfg.nw) * dw, cfg.ricker_omega_max)).to(dev)
    gat = lambda: gather_device(L, off, cfg.T, cfg.nw, cfg.np_slow, cfg.p_max, gamma, S, True)  # noqa: E731
    for _ in range(max(1, warmup)):
        gat()
    return {"workload": "20 layers, 8192 p x 1023 omega, 100 offsets, nt 2048",
            "steps": steps, "ms": event_ms(gat, max(1, steps), torch),
            "bessel_terms": len(cfg.offsets) * cfg.np_slow * (cfg.nw - 1)}


def kernel_figures(torch, dev):
    from seismicinversion_amd import autotune, batch, engine
    from seismicinversion_amd.kennett_gather_gpu import gather_device
    from seismicinversion_amd.source import ricker_spectrum
    from seismicinversion_amd.workloads import gather_config_c2, uneven_batch_models

    out = {}
    # forward kernel on the gather workload
    cfg = gather_config_c2()
    L = engine.layers_tensor(cfg.model, dev)
    p = torch.arange(1, cfg.np_slow + 1, dtype=torch.float64, device=dev) * (cfg.p_max / cfg.np_slow)
    dw, gamma, nw = 2 * np.pi / cfg.T, np.pi / cfg.T, cfg.nw - 1
    R = torch.empty((cfg.np_slow, nw), dtype=torch.complex128, device=dev)
    fwd = lambda: engine.forward_grid(L, p, 1.0, dw, gamma, nw, True, R)  # noqa: E731
    fwd()
    ms = event_ms(fwd, 10, torch)
    upd = cfg.np_slow * nw * (cfg.model.n_layers - 1)
    out["forward_gather_workload"] = {
        "kernel": "kh::forward_kernel<W, true> (default shape)", "ms": ms, "updates": upd,
        "updates_per_s": upd / (ms * 1e-3),
        "tflops": FLOP_PER_UPDATE * upd / (ms * 1e-3) / 1e12,
        "frac_fp64_peak": FLOP_PER_UPDATE * upd / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
    # the Global Matrix Method's block-Riccati kernel on the same grid
    gmm = lambda: engine.gmm_forward_grid(L, p, 1.0, dw, gamma, nw, True)  # noqa: E731
    gms = event_ms(gmm, 5, torch)
    Rg = gmm()
    fwd()
    diff = float(((Rg - R).abs().amax(dim=1) / R.abs().amax(dim=1)).max())
    out["gmm_gather_workload"] = {"kernel": "kh::gmm_kernel<1, true>", "ms": gms,
                                  "updates_per_s": upd / (gms * 1e-3),
                                  "max_row_rel_diff_vs_kennett": diff}
    # full gather, inputs resident
    off = torch.from_numpy(cfg.offsets).to(dev)
    w = np.arange(1, cfg.nw) * dw
    S = torch.from_numpy(ricker_spectrum(w, cfg.ricker_omega_max)).to(dev)
    gat = lambda: gather_device(L, off, cfg.T, cfg.nw, cfg.np_slow, cfg.p_max, gamma, S, True)  # noqa: E731
    gat()
    out["gather"] = {"workload": "20 layers, 8192 p x 1023 omega, 100 offsets, nt 2048",
                     "ms": event_ms(gat, 5, torch)}
    out["_gather_traces"] = gat().cpu().numpy()
    # the reference CLI's meaning of -n 2048: nw = 2048 (2047 evaluated
    # frequencies), nt = 4096, same model, slownesses and offsets
    cfg2 = gather_config_c2(2048)
    w2 = np.arange(1, cfg2.nw) * dw
    S2 = torch.from_numpy(ricker_spectrum(w2, cfg2.ricker_omega_max)).to(dev)
    gat2 = lambda: gather_device(L, off, cfg2.T, cfg2.nw, cfg2.np_slow, cfg2.p_max, gamma, S2, True)  # noqa: E731
    gat2()
    ms2 = event_ms(gat2, 5, torch)
    upd2 = cfg2.np_slow * (cfg2.nw - 1) * (cfg2.model.n_layers - 1)
    out["gather_nw2048"] = {"workload": "20 layers, 8192 p x 2047 omega, 100 offsets, nt 4096 "
                                        "(kennett_gather_gpu CLI -n 2048)",
                            "ms": ms2, "forward_updates": upd2,
                            "bessel_terms": len(cfg2.offsets) * cfg2.np_slow * (cfg2.nw - 1)}
    # 256 uneven models: persistent launch vs one launch per model
    specs = [batch.ModelSpec(model=m, p=pp, nw=1024) for m, pp in uneven_batch_models(256, 42)]
    upd = sum(s.updates() for s in specs)
    Ls = [engine.layers_tensor(s.model, dev) for s in specs]
    Ps = [torch.from_numpy(s.p).to(dev) for s in specs]
    Os = [torch.empty((len(s.p), s.nw), dtype=torch.complex128, device=dev) for s in specs]

    def plain():
        for s, Lm, Pm, Om in zip(specs, Ls, Ps, Os):
            engine.forward_grid(Lm, Pm, s.k0, s.dw, s.gamma, s.nw, False, Om)

    keep_busy(plain, torch)
    t_plain = timed(plain, 3, torch)
    shape = autotune.multi_shape(specs, dev=dev)
    plan = batch.plan_multi(specs, *shape, dev=dev)
    keep_busy(lambda: batch.launch(plan), torch)
    t_multi = timed(lambda: batch.launch(plan), 3, torch)
    out["uneven_256_models"] = {"updates": upd, "shape": list(shape),
                                "plain_ms": t_plain * 1e3, "multi_ms": t_multi * 1e3,
                                "multi_updates_per_s": upd / t_multi,
                                "speedup": t_plain / t_multi}
    out["lm_20_layers"] = lm_figures(torch, dev)
    return out


def lm_figures(torch, dev):
    """20-layer Newton-LM: P = 75 packed parameters, 128 p in (0, 0.6] x 512
    omega, start model perturbed 5 %, 10 iterations, three ways: the device
    loop eager (the default: the whole iteration, trials and decision
    included, enqueued with no host round trip), the device loop replayed
    from one captured HIP graph per iteration, and the host loop that solves
    and decides on the host one trial at a time as the reference does.  One
    misfit/gradient/exact-Hessian evaluation is timed on its own.  Beside it
    the served inversion size (17 layers, 32 p x 128 omega, 5 iterations)
    and the reference CLI's default inversion (5 layers, 12 p x 64 omega),
    where the kernels are short and launch overhead counts."""
    from seismicinversion_amd.taup_inversion import _Problem, invert_taup
    from seismicinversion_amd.workloads import seeded_model, uniform_slownesses

    variants = {"eager": {}, "graphs": {"use_graphs": True}, "host_loop": {"device_loop": False}}
    m = seeded_model(20, 42)
    p = list(uniform_slownesses(128, 0.6))
    for kw in variants.values():  # warm-up
        invert_taup(m, p[:4], nfreq=16, max_iter=3, tol=0.0, **kw)
    runs = {}
    for _ in range(2):
        for name, kw in variants.items():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = invert_taup(m, p, nfreq=512, perturbation=0.05, max_iter=10, tol=0.0, **kw)
            torch.cuda.synchronize()
            runs.setdefault(name, []).append((time.perf_counter() - t0, res))
    best = {k: min(t for t, _ in v) for k, v in runs.items()}
    res = {k: v[0][1] for k, v in runs.items()}
    refs, p_, w, obs, phys = lm_problem(torch, dev)
    prob = _Problem(*refs, p_, w, obs)
    prob.grad_hessian(phys)
    t_gh = timed(lambda: prob.grad_hessian(phys), 3, torch)
    prob_g = _Problem(*refs, p_, w, obs, use_graphs=True)
    prob_g.grad_hessian(phys)  # eager
    prob_g.grad_hessian(phys)  # captured
    t_gh_graph = timed(lambda: prob_g.grad_hessian(phys), 3, torch)
    # the served inversion size: 5 requests of 5 iterations
    small = [(seeded_model(17, 100 + k), list(uniform_slownesses(32, 0.8))) for k in range(5)]
    t_small = {}
    for _ in range(2):
        for name, kw in variants.items():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            its = 0
            for mm, pp in small:
                r = invert_taup(mm, pp, nfreq=128, perturbation=0.05, max_iter=5, tol=0.0, **kw)
                its += r.n_iterations
            torch.cuda.synchronize()
            t_small.setdefault(name, []).append((time.perf_counter() - t0) * 1e3 / its)
    # the reference CLI's default inversion (configs/default_ocean_crust.yaml:
    # 5 layers, 12 p, 64 omega), 20 iterations: kernels of a few microseconds,
    # where launching them is the cost
    from seismicinversion_amd.inversion_config import DEFAULT_CONFIG, load_config

    cfg = load_config(DEFAULT_CONFIG)
    t_cli = {}
    for _ in range(2):
        for name, kw in variants.items():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = invert_taup(cfg.true_model, cfg.p_values, nfreq=cfg.nfreq,
                            perturbation=cfg.perturbation, max_iter=20, seed=cfg.seed, tol=0.0,
                            **kw)
            torch.cuda.synchronize()
            t_cli.setdefault(name, []).append((time.perf_counter() - t0) * 1e3 / r.n_iterations)
    n_it = res["eager"].n_iterations
    return {"P": prob.P, "n_p": 128, "n_omega": 512, "iterations": n_it,
            "eager_s": best["eager"], "graphs_s": best["graphs"],
            "host_loop_s": best["host_loop"],
            "eager_ms_per_iteration": best["eager"] * 1e3 / n_it,
            "graphs_ms_per_iteration": best["graphs"] * 1e3 / res["graphs"].n_iterations,
            "host_loop_ms_per_iteration": best["host_loop"] * 1e3 / res["host_loop"].n_iterations,
            "loops": "eager and graphs: the device loop (trials and decision on the device); "
                     "host_loop: solve and decision on the host, one trial at a time",
            "same_trajectory": res["graphs"].misfit_history == res["eager"].misfit_history,
            "host_loop_max_rel_misfit_diff": float(np.max(np.abs(
                np.array(res["host_loop"].misfit_history) - res["eager"].misfit_history)
                / np.array(res["eager"].misfit_history))),
            "misfit_first_last": [res["eager"].misfit_history[0],
                                  res["eager"].misfit_history[-1]],
            "param_error_last": res["eager"].param_error_history[-1],
            "grad_hessian_ms": t_gh * 1e3,
            "grad_hessian_graph_replay_ms": t_gh_graph * 1e3,
            "tangent_over_adjoint_lane_updates": prob.P * 128 * 512 * 19,
            "served_size_17_layers_32p_128w": {
                f"{k}_ms_per_iteration": min(v) for k, v in t_small.items()},
            "reference_cli_5_layers_12p_64w": {
                f"{k}_ms_per_iteration": min(v) for k, v in t_cli.items()}}


def lm_problem(torch, dev):
    """The 20-layer Newton-LM's device problem (P = 75, 128 p in (0, 0.6] x
    512 omega, observations of the seeded model) and the start model scaled
    by 1.02, where its evaluations are timed."""
    from seismicinversion_amd import engine, kennett_torch as kt
    from seismicinversion_amd.taup_inversion import _Problem
    from seismicinversion_amd.workloads import seeded_model, uniform_slownesses

    m = seeded_model(20, 42)
    p = list(uniform_slownesses(128, 0.6))
    tens = kt.model_to_tensors(m)
    refs = [tens[k] for k in ("alpha", "beta", "rho", "thickness", "Q_alpha", "Q_beta")]
    w = torch.arange(1, 513, dtype=torch.float64) * (2 * np.pi / 64)
    obs = engine.forward(engine.layers_tensor(m, dev), torch.tensor(p), w.to(torch.complex128))
    return refs, p, w, obs, kt._pack_params(*refs[:4]).numpy() * 1.02


def lm_loop_figures(torch, steps, warmup):
    """The 20-layer Newton-LM (128 p x 512 omega, 5 % start perturbation) on
    the device loop: a `warmup`-iteration run, then one `steps`-iteration
    run timed whole."""
    from seismicinversion_amd.taup_inversion import invert_taup
    from seismicinversion_amd.workloads import seeded_model, uniform_slownesses

    m = seeded_model(20, 42)
    p = list(uniform_slownesses(128, 0.6))
    invert_taup(m, p, nfreq=512, perturbation=0.05, max_iter=max(1, warmup), tol=0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = invert_taup(m, p, nfreq=512, perturbation=0.05, max_iter=steps, tol=0.0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"iterations": res.n_iterations, "s": dt, "ms_per_iteration": dt * 1e3 / res.n_iterations,
            "misfit_first_last": [res.misfit_history[0], res.misfit_history[-1]]}


def hessian_figures(torch, dev, steps, warmup):
    """Misfit + gradient + exact Hessian of the 20-layer LM: `warmup`
    untimed evaluations, then `steps` timed one by one."""
    from seismicinversion_amd.taup_inversion import _Problem

    refs, p, w, obs, phys = lm_problem(torch, dev)
    prob = _Problem(*refs, p, w, obs)
    for _ in range(max(1, warmup)):
        chi, g, H = prob.grad_hessian(phys)
    ms = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        chi, g, H = prob.grad_hessian(phys)
        ms.append((time.perf_counter() - t0) * 1e3)
    return {"P": prob.P, "n_p": len(p), "n_omega": len(w), "evaluations_ms": ms,
            "min_ms": min(ms), "misfit": chi, "grad_norm": float(np.linalg.norm(g)),
            "hessian_trace": float(np.trace(H))}


def open_loop_sweep(torch, rates, seconds, fwd_deadline, inv_deadline):
    """The serving stream in open loop, one arrival rate after another on one
    server: Poisson arrivals for `seconds` at each rate, 80 % forward
    requests and 20 % 5-iteration Newton-LM inversions (drawn sizes; the
    inversions keep nl and cap slownesses at 32 and frequencies at 128),
    results to the host.  20-layer forward requests (slowness and frequency
    counts from the same distributions) arrive beside the mix at
    max(20, 5 % of the rate) per second, so that the 20-layer percentiles rest
    on more than the mix's own few 20-layer draws.  A forward request not
    admitted within `fwd_deadline` s, or an inversion not started within
    `inv_deadline` s, is rejected (TimeoutError) and counted.  Results are
    dropped as they complete (only their records are kept), so host memory
    stays bounded at any rate."""
    from seismicinversion_amd.serve import DeviceBackend, KennettServer, stream_requests
    from seismicinversion_amd.workloads import serving_stream

    def run(osrv, stream, reqs):
        lock = threading.Lock()
        left = [len(reqs)]
        all_done = threading.Event()

        def on_done(_f):
            with lock:
                left[0] -= 1
                if left[0] == 0:
                    all_done.set()

        t0 = time.perf_counter()
        for it, (kind, spec) in zip(stream, reqs):
            d = it.t - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            f = (osrv.submit_forward(spec, deadline_s=fwd_deadline) if kind == "forward"
                 else osrv.submit_invert(spec, deadline_s=inv_deadline))
            f.add_done_callback(on_done)  # the future itself is not kept
        t_submit = time.perf_counter() - t0
        all_done.wait()
        return t_submit, time.perf_counter() - t0

    out = []
    with KennettServer(backend=DeviceBackend(n_streams=3)) as osrv:
        # untimed warm-up at the first rate: populates the allocator's cached
        # blocks of the server's streams (a cold server's first second runs
        # forward p99 several times higher)
        warm = serving_stream(rates[0], max(1, int(rates[0])), seed=41, lm_fraction=0.2)
        run(osrv, warm, stream_requests(warm))
        for k, rate in enumerate(rates):
            n = max(1, int(rate * seconds))
            mix = serving_stream(rate, n, seed=42 + k, lm_fraction=0.2)
            probe_rate = max(20.0, 0.05 * rate)
            probe = serving_stream(probe_rate, max(1, int(probe_rate * seconds)), seed=700 + k,
                                   lm_fraction=0.0, nl_range=(20, 20))
            stream = sorted(mix + probe, key=lambda it: it.t)
            reqs = stream_requests(stream)
            osrv.reset_stats()
            t_submit, t_all = run(osrv, stream, reqs)
            st = osrv.stats()
            print(f"open loop {rate:g} req/s: {len(stream)} requests in {t_submit:.2f} s, "
                  f"drained at {t_all:.2f} s", file=sys.stderr, flush=True)
            out.append({
                "rate_req_per_s": rate, "window_s": seconds, "requests": n,
                "inversions": sum(1 for it in mix if it.kind == "invert"),
                "probe_20_layer_rate_req_per_s": probe_rate, "probe_requests": len(probe),
                "achieved_arrival_rate_req_per_s": len(stream) / t_submit,
                "drain_s": t_all,
                "forward_updates_per_s": st["forward_updates_per_s"],
                "completed": st["completed_by_kind"], "rejected": st["rejected_by_kind"],
                "failed_other": st["failed"] - sum(st["rejected_by_kind"].values()),
                "admission_wait_ms": st["admission_wait"], "queued_max": st["waiting_max"],
                "forward_phases_ms": st["forward_phases"],
                "batches": st["batches"],
                "latency_forward_ms": st["latency_forward"],
                "latency_forward_20_layers_ms": st["latency_forward_20_layers"],
                "latency_invert_ms": st["latency_invert"]})
    return {"deadlines_s": {"forward": fwd_deadline, "invert": inv_deadline}, "rates": out}


def served_roofline(log, traffic_json, wpt):
    """FP64 roofline of the persistent multi-model kernel over the timed
    served launches: algorithmic FLOPs (340 per update) per launch over the
    mean launch duration (HIP events on the server's compute stream)."""
    kern_ms = float(np.mean([m for m, _, _ in log]))
    upd_per_launch = float(np.mean([u for _, u, _ in log]))
    out_per_launch = float(np.mean([e for _, _, e in log]))
    achieved = FLOP_PER_UPDATE * upd_per_launch / (kern_ms * 1e-3) / 1e12
    # HBM bytes per launch: the ratio of measured (rocprofv3 FETCH_SIZE and
    # WRITE_SIZE, calibrated on a device copy) to algorithmic bytes over the
    # served dispatches of the counter runs, times this run's algorithmic
    # bytes per launch
    traffic = None
    if os.path.exists(traffic_json):
        with open(traffic_json) as f:
            ratio = json.load(f).get("traffic_over_algorithmic")
        if ratio:
            traffic = float(ratio) * 16.0 * out_per_launch
    return {"bound": "fp64-vector", "pipe": "FP64 VALU (v_fma_f64 and friends), not MFMA "
            "and not HBM: 340 FLOP per 16 output bytes per interface",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic, "kernel": f"kh::multi_kernel<{wpt}>",
            "kernel_ms_spans": "the whole kh_forward_multi call: counter reset, "
                               "kh::stage_tables_kernel and the persistent sweep",
            "launches": len(log),
            "kernel_ms": kern_ms, "updates_per_launch": upd_per_launch,
            "flop_per_update": FLOP_PER_UPDATE,
            "algorithmic_bytes_per_launch": 16.0 * out_per_launch}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(every box has one MI355X; how the path shards is described in the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.section == "lm":
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        warm_clocks(torch, dev, args.warmup_seconds)
        print(json.dumps({"section": "lm", "lm_20_layers": lm_figures(torch, dev)}))
        return
    if args.section == "lmloop":
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        print(json.dumps({"section": "lmloop", "lm_20_layers_loop": lm_loop_figures(
            torch, args.steps, args.warmup)}))
        return
    if args.section == "hessian":
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        print(json.dumps({"section": "hessian", "lm_20_layers_evaluation": hessian_figures(
            torch, dev, args.steps, args.warmup)}))
        return
    if args.section == "gather":
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        warm_clocks(torch, dev, args.warmup_seconds)
        print(json.dumps({"section": "gather", "gather": gather_figures(
            torch, dev, args.steps, args.warmup)}))
        return
    if args.section == "open":
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        warm_clocks(torch, dev, args.warmup_seconds)
        print(json.dumps({"section": "open", "open_loop": open_loop_sweep(
            torch, [float(r) for r in args.rates.split(",") if r], args.open_seconds,
            args.forward_deadline, args.invert_deadline)}))
        return

    from seismicinversion_amd import engine
    from seismicinversion_amd.kennett_reflectivity_gpu import kennett_reflectivity_batch
    from seismicinversion_amd.batch import ModelSpec
    from seismicinversion_amd.serve import DeviceBackend, KennettServer, stream_requests
    from seismicinversion_amd.workloads import serving_stream

    items = serving_stream(0.0, args.requests, seed=42, lm_fraction=0.0)
    specs = [s for _, s in stream_requests(items)]
    updates = sum(s.updates() for s in specs)
    full = args.section == "all"
    if full:
        cpu, gather_cpu, cpu_traces, c1_cpu = cpu_baselines(specs, args.cpu_workers)

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    warm_clocks(torch, dev, args.warmup_seconds)

    # ---- served, device-resident results (headline)
    be = DeviceBackend(n_streams=3, kernel_timing=True)
    srv = KennettServer(backend=be).start()

    def fresh(batch_specs):
        """New request objects for the same models and grids: the server
        validates and packs every layer table again (ModelSpec caches its
        table on the object), as it would for new requests."""
        return [ModelSpec(model=x.model, p=x.p, nw=x.nw, dw=x.dw, k0=x.k0, gamma=x.gamma,
                          free_surface=x.free_surface) for x in batch_specs]

    def served(on_device=True, batch_specs=specs):
        futs = srv.submit_forward_many(fresh(batch_specs), on_device=on_device)
        for k in range(len(futs)):  # each result is dropped once it has arrived
            futs[k].result()
            futs[k] = None

    for _ in range(args.warmup):
        served()
    torch.cuda.synchronize()
    be.kernel_log.clear()
    b0 = srv.batches
    t_served = timed(served, args.steps, torch)
    batches_per_step = (srv.batches - b0) / args.steps
    log = list(be.kernel_log)
    if not full:
        srv.stop()
        r = served_roofline(log, args.traffic_json, be.shape[1])
        print(json.dumps({"section": "served", "value": updates / t_served,
                          "ms_per_step": t_served * 1e3, "batches_per_step": batches_per_step,
                          "roofline": r}))
        return
    served()  # warm the host-result path's staging and pages
    served(False)
    t_served_host = timed(lambda: served(False), 1, torch)

    # ---- plain: one synchronous launch per request
    def plain(batch_specs=specs):
        for x in batch_specs:
            engine.forward_grid(engine.layers_tensor(x.model, dev), torch.from_numpy(x.p).to(dev),
                                x.k0, x.dw, x.gamma, x.nw, x.free_surface)
            torch.cuda.synchronize()

    def plain_host(batch_specs=specs):
        for x in batch_specs:
            kennett_reflectivity_batch(x.model, x.p, (x.k0 + np.arange(x.nw)) * x.dw,
                                       x.free_surface)

    keep_busy(plain, torch)
    t_plain = timed(plain, 1, torch)
    t_plain_host = timed(plain_host, 1, torch)

    # ---- served / plain per size class of the same 400 requests
    classes = {
        "small: nl <= 20 and np*nw <= 2.6e5": lambda x: x.n_layers <= 20 and len(x.p) * x.nw <= 2.6e5,
        "medium: the rest with updates <= 1e8": lambda x: not (x.n_layers <= 20 and len(x.p) * x.nw <= 2.6e5) and x.updates() <= 1e8,
        "large: updates > 1e8": lambda x: x.updates() > 1e8,
    }
    size_classes = {}
    for name, pred in classes.items():
        sub = [x for x in specs if pred(x)]
        if not sub:
            continue
        upd = sum(x.updates() for x in sub)
        keep_busy(lambda: served(True, sub), torch, 0.2)
        ts = timed(lambda: served(True, sub), 3, torch)
        keep_busy(lambda: plain(sub), torch, 0.2)
        tp = timed(lambda: plain(sub), 2, torch)
        tph = timed(lambda: plain_host(sub), 1, torch)
        tsh = timed(lambda: served(False, sub), 2, torch)
        size_classes[name] = {
            "requests": len(sub), "updates": upd,
            "served_updates_per_s": upd / ts, "plain_updates_per_s": upd / tp,
            "served_over_plain": tp / ts,
            "served_host_updates_per_s": upd / tsh, "plain_host_updates_per_s": upd / tph,
            "served_over_plain_host_results": tph / tsh}
    srv.stop()

    # ---- open loop with inversions, results to the host, over the arrival rates
    open_loop = None
    if args.open_seconds > 0:
        open_loop = open_loop_sweep(torch, [float(r) for r in args.rates.split(",") if r],
                                    args.open_seconds, args.forward_deadline,
                                    args.invert_deadline)

    kernels = kernel_figures(torch, dev)
    kernels["c1_plain_5_layers"] = c1_figure(torch, dev, c1_cpu)

    # ---- roofline of the persistent multi-model kernel over the timed served steps
    roofline = served_roofline(log, args.traffic_json, be.shape[1])

    g = kernels["gather"]
    g["cpu_all_cores"] = gather_cpu
    g["gpu_over_cpu"] = gather_cpu["s"] / (g["ms"] * 1e-3)
    g["max_rel_diff_vs_cpu"] = float(np.max(np.abs(kernels.pop("_gather_traces") - cpu_traces))
                                     / np.max(np.abs(cpu_traces)))
    value = updates / t_served
    line = {
        "metric": "p*omega*layer updates/s (complex128), served vs plain, with p50/p99 latency",
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_served * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded layered earth models)",
        "config": {"workload": f"{len(specs)} forward requests of the mixed-size serving "
                               "stream (seed 42; nl log-uniform 3-100, np log-uniform 16-2048 "
                               "in (0,0.8] s/km, nw in {256,1024,4096}) served by continuous "
                               "batching with device-resident results",
                   "requests": len(specs), "updates_per_step": updates,
                   "batches_per_step": batches_per_step, "streams": 3,
                   "launch_shape": list(be.shape), "parallelism": "single GPU"},
        "plain_updates_per_s": updates / t_plain,
        "served_over_plain": t_plain / t_served,
        "served_host_results_updates_per_s": updates / t_served_host,
        "plain_host_results_updates_per_s": updates / t_plain_host,
        "served_over_plain_host_results": t_plain_host / t_served_host,
        "size_classes": size_classes,
        "open_loop": open_loop,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
