"""Benchmark: Kennett P-SV reflectivity served vs plain on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline workload (one "step"): N = 400 forward requests drawn from the
mixed-size serving distribution (seed 42: layer counts round(exp(U[ln 3,
ln 100])), slownesses round(exp(U[ln 16, ln 2048])) uniform in (0, 0.8] s/km,
frequency counts uniform over {256, 1024, 4096}, omega_k = k*2*pi/64),
submitted at once (``submit_forward_many``) to the continuous-batching
server and served with
device-resident results.  ``value`` = p*omega*layer updates per second
(one update = one interface step of one (p, omega) lane).

Beside it, on the same requests:
* plain: the one-model-per-launch synchronous GPU path (one kh_forward_grid
  launch per request, synchronised), with device results and, through the
  kennett_reflectivity_batch drop-in, with host results;
* served with host results (pinned staging + copy-out);
* an open-loop run: Poisson arrivals at --rate req/s for --open-seconds,
  20 % of them 5-iteration LM inversions (slownesses <= 32, frequencies <= 128),
  results to the host; p50/p99 latency from arrival to result for all forward
  requests, for 20-layer forward requests and for inversions;
* kernels: the gather workload's forward kernel (20 layers, 8192 p x 1023
  damped omega, free surface), the full gather (forward + J0 wavenumber sum +
  inverse FFT, 100 offsets), the 256-model uneven batch (persistent
  launch vs one launch per model) and the 20-layer Newton-LM inversion
  (75 parameters, 128 p x 512 omega, 10 iterations, eager and from HIP
  graphs).

``roofline``: the persistent multi-model kernel of the served steps, FP64
algorithmic FLOPs (340 per update) per launch over the average launch
duration (HIP events on the server's compute stream, where the batch
kernels run back to back).  ``cpu_baseline``: the NumPy
oracle, one process, on the middle slowness row of every request.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
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
    ap.add_argument("--rate", type=float, default=200.0, help="open-loop arrivals per second")
    ap.add_argument("--open-seconds", type=float, default=4.0)
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="processes of the all-core CPU baseline (the box's share is 16 cores)")
    ap.add_argument("--warmup-seconds", type=float, default=2.0,
                    help="untimed forward launches before anything is timed (clock ramp)")
    ap.add_argument("--section", choices=("all", "lm"), default="all",
                    help="'lm': only the 20-layer Newton-LM figures (for profiling its kernels)")
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
    with mp.get_context("fork").Pool(workers) as pool:
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
    cpu = {"value": updn / dtn, "unit": UNIT, "cores": workers, "kind": "port",
           "sample": f"{workers} evenly spaced slowness rows of each of the {len(specs)} requests "
                     f"({updn:.3e} updates), NumPy oracle in {workers} forked processes "
                     f"(one thread each), {dtn:.1f} s",
           "single_core": {"value": upd1 / dt1, "cores": 1,
                           "sample": f"middle slowness row of each request ({upd1:.3e} updates), "
                                     f"{dt1:.1f} s"}}
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
    omega, start model perturbed 5 %, 10 iterations; eager and graph-replayed
    (the graph run includes the captures, which run each computation three
    times), plus one misfit/gradient/exact-Hessian evaluation timed on its
    own, eager and replayed from its captured graph."""
    from seismicinversion_amd import engine, kennett_torch as kt
    from seismicinversion_amd.taup_inversion import _Problem, invert_taup
    from seismicinversion_amd.workloads import seeded_model, uniform_slownesses

    m = seeded_model(20, 42)
    p = list(uniform_slownesses(128, 0.6))
    invert_taup(m, p[:4], nfreq=16, max_iter=1, tol=0.0)  # warm-up
    runs = {}
    for graphs in (False, True):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = invert_taup(m, p, nfreq=512, perturbation=0.05, max_iter=10, tol=0.0,
                          use_graphs=graphs)
        torch.cuda.synchronize()
        runs[graphs] = (time.perf_counter() - t0, res)
    tens = kt.model_to_tensors(m)
    refs = [tens[k] for k in ("alpha", "beta", "rho", "thickness", "Q_alpha", "Q_beta")]
    w = torch.arange(1, 513, dtype=torch.float64) * (2 * np.pi / 64)
    obs = engine.forward(engine.layers_tensor(m, dev), torch.tensor(p), w.to(torch.complex128))
    prob = _Problem(*refs, p, w, obs)
    phys = kt._pack_params(*refs[:4]).numpy() * 1.02
    prob.grad_hessian(phys)
    t_gh = timed(lambda: prob.grad_hessian(phys), 3, torch)
    # the same evaluation replayed from its captured graph (capture excluded)
    prob_g = _Problem(*refs, p, w, obs, use_graphs=True)
    prob_g.grad_hessian(phys)
    t_gh_graph = timed(lambda: prob_g.grad_hessian(phys), 3, torch)
    res = runs[False][1]
    return {"P": prob.P, "n_p": 128, "n_omega": 512, "iterations": res.n_iterations,
            "eager_s": runs[False][0], "graphs_s": runs[True][0],
            "eager_ms_per_iteration": runs[False][0] * 1e3 / res.n_iterations,
            "graphs_ms_per_iteration": runs[True][0] * 1e3 / runs[True][1].n_iterations,
            "same_trajectory": runs[True][1].misfit_history == res.misfit_history,
            "misfit_first_last": [res.misfit_history[0], res.misfit_history[-1]],
            "param_error_last": res.param_error_history[-1],
            "grad_hessian_ms": t_gh * 1e3,
            "grad_hessian_graph_replay_ms": t_gh_graph * 1e3,
            "tangent_over_adjoint_lane_updates": prob.P * 128 * 512 * 19}


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

    from seismicinversion_amd import engine
    from seismicinversion_amd.kennett_reflectivity_gpu import kennett_reflectivity_batch
    from seismicinversion_amd.serve import DeviceBackend, KennettServer, stream_requests
    from seismicinversion_amd.workloads import serving_stream

    items = serving_stream(0.0, args.requests, seed=42, lm_fraction=0.0)
    specs = [s for _, s in stream_requests(items)]
    updates = sum(s.updates() for s in specs)
    cpu, gather_cpu, cpu_traces, c1_cpu = cpu_baselines(specs, args.cpu_workers)

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    warm_clocks(torch, dev, args.warmup_seconds)

    # ---- served, device-resident results (headline)
    be = DeviceBackend(n_streams=3, kernel_timing=True)
    srv = KennettServer(backend=be).start()

    def served(on_device=True):
        for f in srv.submit_forward_many(specs, on_device=on_device):
            f.result()

    for _ in range(args.warmup):
        served()
    torch.cuda.synchronize()
    be.kernel_log.clear()
    b0 = srv.batches
    t_served = timed(served, args.steps, torch)
    batches_per_step = (srv.batches - b0) / args.steps
    log = list(be.kernel_log)
    served()  # warm the host-result path's staging and pages
    served(False)
    t_served_host = timed(lambda: served(False), 1, torch)
    srv.stop()

    # ---- plain: one synchronous launch per request
    def plain():
        for s in specs:
            engine.forward_grid(engine.layers_tensor(s.model, dev), torch.from_numpy(s.p).to(dev),
                                s.k0, s.dw, s.gamma, s.nw, s.free_surface)
            torch.cuda.synchronize()

    def plain_host():
        for s in specs:
            kennett_reflectivity_batch(s.model, s.p, (s.k0 + np.arange(s.nw)) * s.dw,
                                       s.free_surface)

    keep_busy(plain, torch)
    t_plain = timed(plain, 1, torch)
    t_plain_host = timed(plain_host, 1, torch)

    # ---- open loop with inversions, results to the host
    open_loop = None
    if args.open_seconds > 0:
        n_open = max(1, int(args.rate * args.open_seconds))
        stream = serving_stream(args.rate, n_open, seed=42, lm_fraction=0.2)
        # 20-layer forward requests (slowness and frequency counts from the
        # same distributions) arriving beside the mix at a tenth of its rate,
        # so that the 20-layer latency percentiles rest on more than the few
        # 20-layer draws of the mix itself
        probe_rate = 0.1 * args.rate
        probe = serving_stream(probe_rate, max(1, int(probe_rate * args.open_seconds)), seed=7,
                               lm_fraction=0.0, nl_range=(20, 20))
        stream = sorted(stream + probe, key=lambda it: it.t)
        reqs = stream_requests(stream)
        with KennettServer(backend=DeviceBackend(n_streams=3)) as osrv:
            # untimed warm-up on the same server: the stream's first second,
            # replayed at its own pace, populates the allocator's cached
            # blocks of the server's streams, so the timed run sees steady
            # state (a cold server's first second runs forward p99 several
            # times higher)
            t0 = time.perf_counter()
            warm = []
            for it, (kind, spec) in zip(stream, reqs):
                if it.t > 1.0:
                    break
                d = it.t - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                warm.append(osrv.submit_forward(spec) if kind == "forward"
                            else osrv.submit_invert(spec))
            for f in warm:
                f.result()
            osrv.reset_stats()
            t0 = time.perf_counter()
            futs = []
            for it, (kind, spec) in zip(stream, reqs):
                d = it.t - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                futs.append(osrv.submit_forward(spec) if kind == "forward"
                            else osrv.submit_invert(spec))
            for f in futs:
                f.result()
        st = osrv.stats()
        open_loop = {"rate_req_per_s": args.rate, "requests": n_open,
                     "probe_20_layer_rate_req_per_s": probe_rate, "probe_requests": len(probe),
                     "inversions": sum(1 for k, _ in reqs if k == "invert"),
                     "forward_updates_per_s": st["forward_updates_per_s"],
                     "batches": st["batches"], "failed": st["failed"],
                     "latency_forward_ms": st["latency_forward"],
                     "latency_forward_20_layers_ms": st["latency_forward_20_layers"],
                     "latency_invert_ms": st["latency_invert"]}

    kernels = kernel_figures(torch, dev)
    kernels["c1_plain_5_layers"] = c1_figure(torch, dev, c1_cpu)

    # ---- roofline of the persistent multi-model kernel over the timed served steps
    kern_ms = float(np.mean([m for m, _, _ in log]))
    upd_per_launch = float(np.mean([u for _, u, _ in log]))
    out_per_launch = float(np.mean([e for _, _, e in log]))
    achieved = FLOP_PER_UPDATE * upd_per_launch / (kern_ms * 1e-3) / 1e12
    # HBM bytes per launch: the ratio of measured (rocprofv3 FETCH_SIZE and
    # WRITE_SIZE, calibrated on a device copy) to algorithmic bytes over the
    # served dispatches of the counter runs, times this run's algorithmic
    # bytes per launch
    traffic = None
    if os.path.exists(args.traffic_json):
        with open(args.traffic_json) as f:
            ratio = json.load(f).get("traffic_over_algorithmic")
        if ratio:
            traffic = float(ratio) * 16.0 * out_per_launch
    roofline = {"bound": "mfma", "pipe": "fp64-vector", "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                "traffic": traffic, "kernel": f"kh::multi_kernel<{be.shape[1]}>",
                "kernel_ms_spans": "the whole kh_forward_multi call: counter reset, "
                                   "kh::stage_tables_kernel and the persistent sweep",
                "launches": len(log),
                "kernel_ms": kern_ms, "updates_per_launch": upd_per_launch,
                "flop_per_update": FLOP_PER_UPDATE,
                "algorithmic_bytes_per_launch": 16.0 * out_per_launch}

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
        "open_loop": open_loop,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
