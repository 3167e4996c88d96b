"""Benchmark: Kennett P-SV reflectivity throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one "step"): the forward reflectivity of the gather workload —
20-layer seeded model, free surface on, 8192 slownesses uniform in (0, 1.2]
s/km, 1023 damped frequencies omega_k = k*2*pi/64 + i*pi/64 (nt = 2048,
T = 64 s) — with the layer table, slownesses and frequencies resident in HBM.

Metric: p*omega*layer updates per second, one update = one interface step of
one (p, omega) lane (8192 * 1023 * 19 = 1.59e8 per step).

Prints one JSON line with the throughput, a roofline entry for the dominant
kernel (FP64 algorithmic FLOPs = 340 per update, timed with HIP events on the
launch stream) and a CPU baseline (the NumPy oracle, one core, on a bounded
sample of the same slowness grid).
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
PEAK_FP64_TFLOPS = 78.6          # MI355X FP64 (vector = matrix)
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample-rows", type=int, default=512,
                    help="slowness rows of the CPU baseline sample (0 = skip)")
    ap.add_argument("--traffic-json", default=os.path.join(HERE, "profiles", "traffic_forward_c2.json"),
                    help="HBM bytes per launch measured with rocprofv3 --pmc (optional)")
    return ap.parse_args()


def c2_forward_inputs():
    from seismicinversion_amd.workloads import seeded_model, uniform_slownesses, omega_grid

    model = seeded_model(20, 42)
    p = uniform_slownesses(8192, 1.2)
    omega = omega_grid(1023) + 1j * (np.pi / 64.0)
    return model, p, omega


def cpu_baseline(model, p, omega, rows):
    """The oracle (NumPy, one process) on every (len(p)/rows)-th slowness row."""
    from oracle import kennett_numpy as K

    stride = max(1, len(p) // rows)
    ps = p[::stride][:rows]
    K.reflectivity_batch(model, ps[:8], omega, True)  # warm-up
    t0 = time.perf_counter()
    K.reflectivity_batch(model, ps, omega, True)
    dt = time.perf_counter() - t0
    upd = len(ps) * len(omega) * (model.n_layers - 1)
    return {"value": upd / dt, "unit": "p*omega*layer updates/s", "cores": 1, "kind": "port",
            "sample": f"{len(ps)} of {len(p)} slowness rows (every {stride}th) x {len(omega)} "
                      f"omega x {model.n_layers - 1} interfaces, NumPy oracle, 1 process, "
                      f"{dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(every box has one MI355X; sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from seismicinversion_amd import engine

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    model, p, omega = c2_forward_inputs()
    layers = engine.layers_tensor(model, dev)
    p_t = torch.from_numpy(p).to(dev)
    out = torch.empty((len(p), len(omega)), dtype=torch.complex128, device=dev)
    updates = len(p) * len(omega) * (model.n_layers - 1)

    dw = 2.0 * np.pi / 64.0

    def step():
        engine.forward_grid(layers, p_t, 1.0, dw, np.pi / 64.0, len(omega), free_surface=True,
                            out=out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms_per_step = wall * 1e3 / args.steps
    value = updates / (wall / args.steps)

    achieved = FLOP_PER_UPDATE * updates / (kern_ms * 1e-3) / 1e12
    traffic = None
    if os.path.exists(args.traffic_json):
        with open(args.traffic_json) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "mfma", "pipe": "fp64-vector", "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                "traffic": traffic, "kernel": "kh::forward_kernel<W, true>",
                "kernel_ms": kern_ms, "flop_per_update": FLOP_PER_UPDATE,
                "algorithmic_bytes": 16 * len(p) * len(omega),
                "hbm_gbs": 16 * len(p) * len(omega) / (kern_ms * 1e-3) / 1e9}

    cpu = None
    if args.cpu_sample_rows > 0:
        cpu = cpu_baseline(model, p, omega, args.cpu_sample_rows)

    line = {
        "metric": "p*omega*layer updates/s (complex128)",
        "value": value,
        "unit": "p*omega*layer updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded 20-layer model)",
        "config": {"workload": "forward reflectivity of the gather workload: 20-layer seeded "
                               "model (seed 42), free surface, 8192 p in (0,1.2] s/km, 1023 "
                               "damped omega (T=64 s, nt=2048, gamma=pi/64)",
                   "n_layers": 20, "n_p": len(p), "n_omega": len(omega),
                   "updates_per_step": updates, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
