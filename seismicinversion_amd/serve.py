"""Continuous-batching reflectivity service.

Requests (forward models and small Levenberg-Marquardt inversions, each
with its own layer count, slownesses and frequency grid) arrive from any
thread through ``submit_forward`` / ``submit_invert`` and get a
``concurrent.futures.Future``.  One scheduler thread runs the loop:

1. admission by device memory: a request is admitted when its HBM footprint
   (layer table, slownesses, output, and for inversions the derivative
   workspace) fits in what the budget has free; requests larger than the
   whole budget fail with MemoryError, the rest wait in arrival order;
2. batching: every tick, all admitted forward requests not yet launched go
   into ONE persistent kh_forward_multi launch (requests that arrived while
   earlier batches run join the next launch), enqueued on the next of
   several default-priority HIP streams together with its host->device and
   device->host copies, so consecutive batches overlap;
3. completion: a second thread waits for each batch (HIP event), copies its
   outputs from pinned staging into per-request host arrays on a pool of
   copy threads, resolves the futures and returns the memory and the stream
   slot, while the scheduler keeps forming and launching batches;
4. inversions run their LM iterations on a stream of their own, beside the
   batches; by default one inversion at a time (``max_lm_concurrency``):
   an inversion's derivative kernels fill the device by themselves, so a
   second concurrent one adds little throughput (open loop at 200 req/s with
   20 % inversions: the queue drains in 7.3 s with one, 6.8 s with four) but
   delays forward batches behind two streams of long kernels (forward p99
   13 ms with one, 82 ms with two).  An inversion's derivative launches are
   split to at most ``lm_launch_updates`` (direction, p, omega, interface)
   steps each, and each waits for the forward batches already queued, so a
   forward batch never shares the device with more than one such short
   launch.

Each request keeps a record (arrival, admission, launch, completion times,
updates, status) from which ``stats()`` reports p50/p99 latency and
updates/s.  Results do not depend on the batch a request lands in: the
persistent kernel's per-lane arithmetic is fixed by the launch shape, which
the server holds constant.

The device work goes through a backend object so that the scheduling logic
can be exercised without a GPU (the tests substitute a host backend).
"""

from __future__ import annotations

import gc
import queue
import threading
import time
from collections import deque
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .batch import ModelSpec

__all__ = ["InvertSpec", "Request", "KennettServer", "DeviceBackend", "percentile_stats",
           "stream_requests"]


@dataclass
class InvertSpec:
    """A small LM inversion: recover `true_model` (sub-ocean Vp, Vs, rho, h)
    from its own reflectivity at slownesses p and nfreq frequencies, starting
    from a perturbed model, for `iterations` Newton-LM steps."""

    true_model: object
    p: np.ndarray
    nfreq: int = 64
    perturbation: float = 0.05
    iterations: int = 5
    seed: int = 42
    hessian_mode: str = "hvp"
    max_trials: int = 10
    use_graphs: bool = False  # replay each iteration's device work from HIP graphs

    @property
    def n_layers(self) -> int:
        return len(self.true_model.alpha)

    def updates(self) -> int:
        return len(self.p) * self.nfreq * (self.n_layers - 1)


@dataclass
class Request:
    kind: str
    spec: object
    future: Future
    arrival: float
    deadline: float | None = None
    on_device: bool = False
    mem: int | None = None  # device bytes, set when the request is validated
    record: dict = field(default_factory=dict)


def stream_requests(items, p_max: float = 0.8, lm_np_max: int = 32, lm_nfreq_max: int = 128,
                    lm_hessian_mode: str = "gram"):
    """(kind, spec) for each StreamItem of ``workloads.serving_stream``.

    Forward requests take the drawn sizes: the seeded model with
    ``n_layers`` layers (seed = the item's seed), ``n_p`` slownesses uniform
    in (0, p_max] and ``n_omega`` frequencies omega_k = k * 2*pi/64.
    Inversions (5 Newton-LM iterations with the exact Hessian) keep the drawn
    layer count and cap slownesses and frequencies at ``lm_np_max`` /
    ``lm_nfreq_max``: one exact-Hessian iteration costs about
    4 (nl - 1) tangent-over-adjoint sweeps of the whole (p, omega) grid, so an
    uncapped draw (up to 2048 x 4096 at nl = 100) would take minutes per
    request."""
    from .workloads import seeded_model, uniform_slownesses

    out = []
    for it in items:
        model = seeded_model(it.n_layers, it.seed)
        if it.kind == "forward":
            out.append(("forward", ModelSpec(model=model, p=uniform_slownesses(it.n_p, p_max),
                                             nw=it.n_omega)))
        else:
            out.append(("invert", InvertSpec(
                true_model=model, p=uniform_slownesses(min(it.n_p, lm_np_max), p_max),
                nfreq=min(it.n_omega, lm_nfreq_max), iterations=5, seed=it.seed,
                hessian_mode=lm_hessian_mode)))
    return out


def percentile_stats(records, kind="forward", n_layers=None) -> dict:
    lat = [r["end"] - r["arrival"] for r in records
           if r.get("status") == "ok" and r["kind"] == kind and
           (n_layers is None or r["n_layers"] == n_layers)]
    if not lat:
        return {"count": 0}
    lat = np.asarray(lat) * 1e3
    return {"count": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean())}


# Host results of a batch are copied down in contiguous groups: the first
# of at most this many bytes, each further one this factor larger.
COPY_GROUP_FIRST = 4 << 20
COPY_GROUP_GROWTH = 4


class PinnedResultPool:
    """Page-locked host memory handed out to callers as their results.

    A batch leases one contiguous block for all of its host results; the
    device copies the rows straight into it and each request gets a NumPy
    view of its own part, so there is no second copy on the host and no
    page faulting of fresh result arrays.  The block returns to the pool when
    the last of its views is garbage (a finalizer on each view's tensor);
    a batch that finds no free block of its size falls back to the staging
    buffer and a copy into ordinary host arrays.  First fit over a free list
    kept sorted by offset, with neighbours merged on release."""

    GRANULE = 1 << 16

    def __init__(self, nbytes: int, buf=None):
        import torch

        self.nbytes = int(nbytes) // self.GRANULE * self.GRANULE
        # ``buf``: a uint8 tensor to carve up instead of a new page-locked one
        self.buf = (buf[:self.nbytes] if buf is not None
                    else torch.empty(self.nbytes, dtype=torch.uint8, pin_memory=True))
        self.free = [(0, self.nbytes)] if self.nbytes > 0 else []
        self.lock = threading.Lock()

    def take(self, nbytes: int):
        """(offset, size) of a free block of at least nbytes, or None."""
        size = -(-int(nbytes) // self.GRANULE) * self.GRANULE
        with self.lock:
            for k, (off, n) in enumerate(self.free):
                if n >= size:
                    if n == size:
                        del self.free[k]
                    else:
                        self.free[k] = (off + size, n - size)
                    return off, size
        return None

    def give(self, off: int, size: int) -> None:
        import bisect

        with self.lock:
            k = bisect.bisect_left(self.free, (off, 0))
            self.free.insert(k, (off, size))
            if k + 1 < len(self.free) and off + size == self.free[k + 1][0]:
                self.free[k] = (off, size + self.free[k + 1][1])
                del self.free[k + 1]
            if k > 0 and self.free[k - 1][0] + self.free[k - 1][1] == off:
                self.free[k - 1] = (self.free[k - 1][0], self.free[k - 1][1] + self.free[k][1])
                del self.free[k]

    def free_bytes(self) -> int:
        with self.lock:
            return sum(n for _, n in self.free)


class _Lease:
    """One batch's block of a PinnedResultPool, returned when its count of
    holders (the batch itself plus every view handed out) reaches zero."""

    def __init__(self, pool: PinnedResultPool, off: int, size: int):
        self.pool, self.off, self.size = pool, off, size
        self.holders = 1  # the batch, until all of its results are delivered
        self.lock = threading.Lock()
        import torch

        self.c128 = pool.buf[off:off + size].view(torch.complex128)

    def hold(self) -> None:
        with self.lock:
            self.holders += 1

    def drop(self) -> None:
        with self.lock:
            self.holders -= 1
            last = self.holders == 0
        if last:
            self.pool.give(self.off, self.size)

    def array(self, a: int, b: int, shape):
        """NumPy view of complex elements [a, b) of the block, holding it."""
        import weakref

        base = self.c128[a:b].numpy()
        self.hold()
        # every NumPy view of the result (the reshape below, a caller's
        # slices) refers to `base`, so it dies with the last of them
        weakref.finalize(base, self.drop)
        return base.reshape(shape)


class DeviceBackend:
    """HIP backend: persistent multi-model launches and device LM inversions.

    Batch kernels run back to back on one compute stream (a persistent launch
    fills the device by itself, so kernels gain nothing from overlapping one
    another).  Each of the ``n_streams`` slots has its own copy stream and a
    pinned host staging buffer of ``staging_bytes``: a batch's inputs go up
    through pinned memory on its slot's copy stream, the compute stream waits
    for them, and its outputs come back into the staging buffer on the copy
    stream once the kernel is done (pinned copies run at the link rate;
    pageable memory would be staged by the runtime at a fraction of it).
    Host results go straight into a block of the ``result_pool_bytes``
    page-locked result pool that the caller's arrays then view
    (``PinnedResultPool``); without a free block they come back into the
    slot's staging buffer and ``results_each`` copies them into host arrays
    on ``copy_threads`` threads while later batches run.  The server caps
    batches at ``max_batch_bytes`` (the staging size); a single request larger
    than that gets a one-off pinned buffer."""

    def __init__(self, n_streams: int = 3, shape=(64, 8), device=None,
                 staging_bytes: int = 512 << 20, copy_threads: int = 8,
                 kernel_timing: bool = False, lm_streams: int = 1,
                 lm_launch_updates: float = 5e7, result_pool_bytes: int = 2 << 30):
        import torch

        from . import _abi

        self.torch = torch
        self.dev = _abi.device(device)
        self.compute = torch.cuda.Stream(device=self.dev)
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(n_streams)]
        # one stream per concurrent inversion (match the server's
        # max_lm_concurrency), so that one inversion's host synchronisations do
        # not wait for another's kernels; each stream keeps its own cached
        # workspace blocks in torch's allocator
        self.lm_pool: queue.Queue = queue.Queue()
        for _ in range(max(1, lm_streams)):
            self.lm_pool.put(torch.cuda.Stream(device=self.dev))
        self.lm_streams = max(1, lm_streams)
        # the last batch launched on the compute stream (an inversion's
        # derivative launches wait for it: forward batches go first)
        self.last_batch_done = None
        self.lm_launch_updates = lm_launch_updates
        self.slots = n_streams
        self.shape = shape
        self.max_batch_bytes = int(staging_bytes)
        self.staging = [torch.empty(self.max_batch_bytes, dtype=torch.uint8, pin_memory=True)
                        for _ in range(n_streams)]
        self.result_pool = PinnedResultPool(result_pool_bytes) if result_pool_bytes > 0 else None
        self._free = list(range(n_streams))
        self._lock = threading.Lock()
        self.pool = ThreadPoolExecutor(max_workers=copy_threads,
                                       thread_name_prefix="kennett-copy")
        # with kernel_timing, (launch ms from HIP events on the compute
        # stream, updates, output elements) per completed batch
        self.kernel_timing = kernel_timing
        self.kernel_log: list[tuple[float, int, int]] = []

    def free_bytes(self) -> int:
        free, _ = self.torch.cuda.mem_get_info(self.dev)
        return int(free)

    @staticmethod
    def forward_bytes(spec: ModelSpec) -> int:
        # output, slownesses, layer table, the persistent launch's per-(slowness,
        # layer) records (256 B each), descriptor and NaN/Inf count
        n_p, nl = len(spec.p), spec.n_layers
        return 16 * n_p * spec.nw + 8 * n_p + 48 * nl + 256 * n_p * nl + 100

    @staticmethod
    def invert_bytes(spec: InvertSpec) -> int:
        from . import derivatives

        nl, n_p, nw = spec.n_layers, len(spec.p), spec.nfreq
        return min(derivatives.WORKSPACE_LIMIT, 1 << 34) + 64 * n_p * nw * 4 * nl

    def launch_forward(self, specs: list[ModelSpec], to_host: list[bool] | None = None):
        """Queue one batch: inputs up on a slot's copy stream, the persistent
        launch on the compute stream, then the device->host copies of the
        outputs that go to the host (``to_host[i]``, all by default) and of the
        NaN/Inf counts on the copy stream.  The host does not wait."""
        from . import batch

        torch = self.torch
        if to_host is None:
            to_host = [True] * len(specs)
        with self._lock:
            if not self._free:
                raise RuntimeError("no free stream slot: more batches in flight than streams")
            slot = self._free.pop()
        cp = self.streams[slot]
        try:
            with torch.cuda.stream(cp):
                plan = batch.plan_multi(specs, *self.shape, dev=self.dev, non_blocking=True)
                inputs_ready = torch.cuda.Event()
                inputs_ready.record(cp)
            self.compute.wait_event(inputs_ready)
            with torch.cuda.stream(self.compute):
                t_ev = None
                if self.kernel_timing:
                    t_ev = (torch.cuda.Event(enable_timing=True),
                            torch.cuda.Event(enable_timing=True))
                    t_ev[0].record(self.compute)
                batch.launch(plan)
                # the plan's buffers were allocated on the copy stream and are
                # written here: if anything below raises, the allocator must
                # not hand them out again before this kernel has finished
                for t in (plan.R, plan.desc, plan.layers, plan.p, plan.counter, plan.nonfinite,
                          plan.tables):
                    if t is not None:
                        t.record_stream(self.compute)
                if t_ev is not None:
                    t_ev[1].record(self.compute)
                kernel_done = torch.cuda.Event()
                kernel_done.record(self.compute)
                self.last_batch_done = kernel_done
            cp.wait_event(kernel_done)
            with torch.cuda.stream(cp):
                total = int(plan.r_off[-1])
                counts = torch.empty(len(specs), dtype=torch.int32, pin_memory=True)
                counts.copy_(plan.nonfinite[:len(specs)], non_blocking=True)
                counts_ev = torch.cuda.Event()
                counts_ev.record(cp)
                host = None
                lease = None
                h0 = 0  # first complex element of R that the host buffer holds
                groups = []  # (request indices, event) in copy order
                if any(to_host):
                    hosts = [i for i, h in enumerate(to_host) if h]
                    h0 = int(plan.r_off[min(hosts)])
                    span = int(plan.r_off[max(hosts) + 1]) - h0
                    blk = self.result_pool.take(16 * span) if self.result_pool else None
                    if blk is not None:
                        lease = _Lease(self.result_pool, *blk)
                        host = lease.c128
                    elif 16 * total <= self.max_batch_bytes:
                        h0 = 0
                        host = self.staging[slot][:16 * total].view(torch.complex128)
                    else:
                        h0 = 0
                        host = torch.empty(total, dtype=torch.complex128, pin_memory=True)
                    # host rows go down in a few contiguous copies of growing
                    # size, each with its own event, so the requests of the
                    # first copies are complete before the whole batch has
                    # arrived (the server lists a batch smallest output first)
                    cap = COPY_GROUP_FIRST
                    run: list[int] = []
                    a = b = 0
                    for i in [i for i, h in enumerate(to_host) if h] + [None]:
                        lo = int(plan.r_off[i]) if i is not None else -1
                        if run and (i is None or lo != b or 16 * (b - a) >= cap):
                            host[a - h0:b - h0].copy_(plan.R[a:b], non_blocking=True)
                            e = torch.cuda.Event()
                            e.record(cp)
                            groups.append((run, e))
                            cap *= COPY_GROUP_GROWTH
                            run = []
                        if i is None:
                            break
                        if not run:
                            a = lo
                        run.append(i)
                        b = int(plan.r_off[i + 1])
                ev = torch.cuda.Event()
                ev.record(cp)
        except BaseException:
            self._release_slot(slot)
            raise
        return {"plan": plan, "host": host, "lease": lease, "h0": h0, "to_host": list(to_host),
                "counts": counts, "counts_event": counts_ev, "groups": groups, "event": ev,
                "slot": slot,
                "timing": t_ev,
                "updates": sum(x.updates() for x in specs)}

    def _release_slot(self, slot: int) -> None:
        with self._lock:
            self._free.append(slot)

    def done(self, handle) -> bool:
        return handle["event"].query()

    def results(self, handle):
        """(per-request results, per-request NaN/Inf counts); frees the slot."""
        outs: list = [None] * len(handle["plan"].shapes)
        counts: list = [0] * len(outs)

        def put(i, out, count):
            outs[i] = out
            counts[i] = count

        self.results_each(handle, put)
        return outs, counts

    def results_each(self, handle, deliver):
        """Hands each request of a batch its result as soon as it is complete:
        ``deliver(i, result, nonfinite_count)`` for every request i, device
        results right after the kernel, then host results in their copy order
        (smallest first), each once its copy group has arrived: as a view of
        the batch's page-locked result block, or (no free block) after a copy
        out of the staging buffer into a host array on the copy pool.  Frees
        the slot at the end."""
        from . import batch

        lease = handle.get("lease")
        try:
            plan, to_host = handle["plan"], handle["to_host"]
            handle["counts_event"].synchronize()
            counts = batch.nonfinite_counts(plan, handle["counts"].numpy())
            dev_views = batch.split(plan)
            for i, h in enumerate(to_host):
                if not h:
                    deliver(i, dev_views[i], counts[i])
            h0 = handle["h0"]
            if lease is not None:  # rows copied straight into callers' memory
                for run, e in handle["groups"]:
                    e.synchronize()
                    for i in run:
                        a, b = int(plan.r_off[i]) - h0, int(plan.r_off[i + 1]) - h0
                        deliver(i, lease.array(a, b, plan.shapes[i]), counts[i])
            elif handle["groups"]:
                flat = handle["host"].numpy()
                chunk = 1 << 19  # complex elements (8 MiB) per copy job
                for run, e in handle["groups"]:
                    outs = {i: np.empty(plan.shapes[i], np.complex128) for i in run}
                    jobs = []
                    for i in run:
                        n_p, nw = plan.shapes[i]
                        rows = max(1, chunk // nw)
                        jobs += [(i, r0, min(n_p, r0 + rows)) for r0 in range(0, n_p, rows)]

                    def copy(job, outs=outs):
                        i, r0, r1 = job
                        nw = plan.shapes[i][1]
                        off = int(plan.r_off[i])
                        np.copyto(outs[i][r0:r1], flat[off + r0 * nw:off + r1 * nw].reshape(r1 - r0, nw))

                    e.synchronize()
                    if len(jobs) <= 2:
                        for job in jobs:
                            copy(job)
                    else:
                        for _ in self.pool.map(copy, jobs):
                            pass
                    for i in run:
                        deliver(i, outs[i], counts[i])
            handle["event"].synchronize()
            if lease is not None:
                lease.drop()  # the batch's own hold; the views keep the block
                lease = None
            if handle["timing"] is not None:
                a, b = handle["timing"]
                with self._lock:
                    self.kernel_log.append((a.elapsed_time(b), handle["updates"],
                                            int(plan.r_off[-1])))
        finally:
            if lease is not None:  # an error before the batch let go of its block
                handle["event"].synchronize()  # no copy may still target the block
                lease.drop()
            self._release_slot(handle["slot"])

    def invert(self, spec: InvertSpec):
        from .taup_inversion import invert_taup

        from . import derivatives

        stream = self.lm_pool.get()

        def forward_first():
            ev = self.last_batch_done
            if ev is not None and not self.torch.cuda.is_current_stream_capturing():
                stream.wait_event(ev)

        try:
            with self.torch.cuda.stream(stream), \
                    derivatives.paced(self.lm_launch_updates, forward_first):
                res = invert_taup(spec.true_model, list(spec.p), nfreq=spec.nfreq,
                                  perturbation=spec.perturbation, max_iter=spec.iterations,
                                  seed=spec.seed, tol=0.0, hessian_mode=spec.hessian_mode,
                                  max_trials=spec.max_trials, use_graphs=spec.use_graphs)
                stream.synchronize()
        finally:
            self.lm_pool.put(stream)
        return res


class KennettServer:
    def __init__(self, backend=None, hbm_budget: int | None = None, max_batch_updates: float = 1e9,
                 max_lm_concurrency: int = 1, idle_sleep: float = 2e-4,
                 max_inflight: int | None = None, max_batch_bytes: int | None = None,
                 first_batch_updates: float = 2.5e8):
        self.backend = (backend if backend is not None
                        else DeviceBackend(lm_streams=max(1, max_lm_concurrency)))
        # batches in flight at once (one per stream): further arrivals wait for
        # a free slot and then launch together
        self.max_inflight = max_inflight or getattr(self.backend, "slots", 2)
        if hbm_budget is None:
            hbm_budget = int(0.8 * self.backend.free_bytes())
        self.budget = int(hbm_budget)
        self.in_use = 0
        self.max_batch_updates = max_batch_updates
        # with nothing in flight the device is idle: the first batch of a
        # burst is cut short so that its launch does not wait for a whole
        # batch to be validated (the next batches fill up while it runs)
        self.first_batch_updates = min(first_batch_updates, max_batch_updates)
        # set by arrivals, completions and releases: the scheduler sleeps on it
        self._wake = threading.Event()
        self.max_batch_bytes = (max_batch_bytes if max_batch_bytes is not None
                                else getattr(self.backend, "max_batch_bytes", None))
        self.incoming: queue.Queue = queue.Queue()
        self.waiting: deque[Request] = deque()
        self.completions: queue.Queue = queue.Queue()
        self.n_inflight = 0
        self.outstanding = 0  # submitted and not yet finished
        self.records: list[dict] = []
        self.lock = threading.Lock()
        self.waiting_inv: deque[Request] = deque()
        # longest FIFOs seen since the last reset_stats (forward, inversion)
        self.waiting_max = [0, 0]
        self.lm_threads: list[threading.Thread] = []
        self.max_lm = max_lm_concurrency
        self.lm_running = 0
        self.idle_sleep = idle_sleep
        self._stop = threading.Event()
        self._thread: threading.Thread | None = None
        self._completer: threading.Thread | None = None
        self.batches = 0

    # ------------------------------------------------------------------ client API
    def submit_forward(self, spec: ModelSpec, deadline_s: float | None = None,
                       on_device: bool = False) -> Future:
        """R (np, nw) complex128 as a host array, or with ``on_device`` as a
        device tensor (a view into its batch's output, which stays allocated
        while the caller holds it; no device->host copy is made).  The tensor
        was written on one of the server's streams and is complete when the
        future resolves; a consumer on another stream calls
        ``tensor.record_stream(stream)`` before dropping it, as for any tensor
        shared across streams."""
        return self._submit("forward", spec, deadline_s, on_device)

    def submit_forward_many(self, specs, deadline_s: float | None = None,
                            on_device: bool = False) -> list[Future]:
        """submit_forward for a list of requests, queued as one item: the
        scheduler sees them together instead of one Python call apart."""
        reqs = [self._make("forward", s, deadline_s, on_device) for s in specs]
        self._enqueue([r for r in reqs if not r.future.done()])
        return [r.future for r in reqs]

    def submit_invert(self, spec: InvertSpec, deadline_s: float | None = None) -> Future:
        return self._submit("invert", spec, deadline_s)

    def _make(self, kind, spec, deadline_s, on_device=False) -> Request:
        """A Request for `spec` (validated later, at admission: an unusable
        one fails alone and the service goes on)."""
        fut: Future = Future()
        now = time.perf_counter()
        req = Request(kind=kind, spec=spec, future=fut, arrival=now,
                      deadline=None if deadline_s is None else now + deadline_s,
                      on_device=bool(on_device))
        # layer count and updates are filled in when the scheduler validates
        # the request (one batch at a time, while earlier batches run), not
        # here in the caller's thread
        req.record = {"kind": kind, "arrival": now, "n_layers": 0, "updates": 0,
                      "status": "queued"}
        return req

    def _enqueue(self, reqs: list[Request]) -> None:
        if not reqs:
            return
        with self.lock:
            self.outstanding += len(reqs)
        for r in reqs:
            r.record["counted"] = True
        self.incoming.put(reqs if len(reqs) > 1 else reqs[0])
        self._wake.set()

    def _submit(self, kind, spec, deadline_s, on_device=False) -> Future:
        req = self._make(kind, spec, deadline_s, on_device)
        if not req.future.done():
            self._enqueue([req])
        return req.future

    def start(self) -> "KennettServer":
        # Objects alive now (the caller's data, earlier requests' records) are
        # moved out of the cyclic collector's reach: a full collection over
        # them holds the interpreter, and with it the scheduler thread, for
        # tens of ms (measured: 31-41 ms pauses inside an open-loop run).
        gc.collect()
        gc.freeze()
        self._stop.clear()
        self._thread = threading.Thread(target=self._loop, name="kennett-scheduler", daemon=True)
        self._completer = threading.Thread(target=self._complete_loop, name="kennett-completion",
                                           daemon=True)
        self._thread.start()
        self._completer.start()
        return self

    def stop(self, drain: bool = True) -> None:
        if drain:
            while True:
                with self.lock:
                    if self.outstanding == 0:
                        break
                time.sleep(1e-3)
        self._stop.set()
        self._wake.set()
        if self._thread is not None:
            self._thread.join()
        if self._completer is not None:
            self.completions.put(None)
            self._completer.join()
        for t in self.lm_threads:
            t.join()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    def reset_stats(self) -> None:
        """Forget the records of completed requests (e.g. after a warm-up)."""
        with self.lock:
            self.records = []
            self.batches = 0
            self.waiting_max = [0, 0]

    def stats(self) -> dict:
        with self.lock:
            recs = list(self.records)
        done = [r for r in recs if r.get("status") == "ok"]
        fwd = [r for r in done if r["kind"] == "forward"]
        span = (max(r["end"] for r in done) - min(r["arrival"] for r in done)) if done else 0.0
        rejected = {k: sum(1 for r in recs if r["kind"] == k and
                           r.get("status") in ("TimeoutError", "MemoryError"))
                    for k in ("forward", "invert")}
        def pct(x):
            return ({"p50_ms": float(np.percentile(x, 50) * 1e3),
                     "p99_ms": float(np.percentile(x, 99) * 1e3)} if x else {})

        queue_ms = {k: pct([r["admit"] - r["arrival"] for r in recs
                            if r["kind"] == k and "admit" in r])
                    for k in ("forward", "invert")}
        # forward requests: admission to launch (waiting for a batch and a
        # stream slot) and launch to result (transfer, kernel, copy-out)
        phases = {"admit_to_launch": pct([r["launch"] - r["admit"] for r in fwd
                                          if "launch" in r and "admit" in r]),
                  "launch_to_result": pct([r["end"] - r["launch"] for r in fwd if "launch" in r])}
        return {
            "completed": len(done),
            "failed": sum(1 for r in recs if r.get("status") not in ("ok", "queued")),
            "completed_by_kind": {k: sum(1 for r in done if r["kind"] == k)
                                  for k in ("forward", "invert")},
            "rejected_by_kind": rejected,
            "admission_wait": queue_ms,
            "forward_phases": phases,
            "waiting_max": {"forward": self.waiting_max[0], "invert": self.waiting_max[1]},
            "batches": self.batches,
            "forward_updates": int(sum(r["updates"] for r in fwd)),
            "forward_updates_per_s": (sum(r["updates"] for r in fwd) / span) if span else 0.0,
            "latency_forward": percentile_stats(recs, "forward"),
            "latency_forward_20_layers": percentile_stats(recs, "forward", 20),
            "latency_invert": percentile_stats(recs, "invert"),
        }

    # ------------------------------------------------------------------ scheduler
    def _finish(self, req: Request, result=None, error: BaseException | None = None,
                nonfinite: int | None = None):
        now = time.perf_counter()
        req.record["end"] = now
        if error is not None:
            req.record["status"] = type(error).__name__
            req.future.set_exception(error)
        else:
            req.record["status"] = "ok"
            if nonfinite is not None:
                req.record["nonfinite"] = int(nonfinite)
            elif isinstance(result, np.ndarray):
                req.record["nonfinite"] = int(np.count_nonzero(~np.isfinite(result)))
            req.future.set_result(result)
        with self.lock:
            self.records.append(req.record)
            if req.record.pop("counted", False):
                self.outstanding -= 1

    def _mem_of(self, req: Request) -> int:
        if req.kind == "forward":
            return self.backend.forward_bytes(req.spec)
        return self.backend.invert_bytes(req.spec)

    def _validate(self, req: Request):
        """Raises ValueError for an unusable request.  A forward spec's layer
        table is validated once and cached on the spec (ModelSpec.table), so
        resubmitted specs cost nothing here."""
        from .layer_model import MAX_LAYERS_DERIV, check_table, pack_layers

        spec = req.spec
        if req.kind == "forward":
            spec.check()  # table (shape, values, layer cap), slownesses, frequencies
        else:
            m = spec.true_model
            if isinstance(m, np.ndarray):
                check_table(m, MAX_LAYERS_DERIV)
            else:
                pack_layers(m, MAX_LAYERS_DERIV)
            if len(spec.p) == 0 or spec.nfreq <= 0 or spec.iterations < 0:
                raise ValueError("an inversion needs at least one slowness and one frequency")
        req.record["n_layers"] = spec.n_layers
        req.record["updates"] = spec.updates()

    def _admit(self, fwd_room: float = float("inf")) -> list[Request]:
        """Arrivals join one of two FIFOs; requests leave them admitted.
        Forward requests are admitted in arrival order as device memory allows
        (once one waits, later ones wait too), and only up to `fwd_room`
        updates per call: the scheduler asks for about one batch at a time,
        so a burst of arrivals is validated and reserved while the previous
        batch runs instead of all before the first launch.  An inversion is
        admitted only when an inversion slot is free and its workspace fits,
        so inversions waiting for a slot hold no memory and never hold up
        forward requests.  A request is validated when it reaches the head
        of its FIFO; an unusable one fails alone.  Deadlines are checked at
        the head of each FIFO (requests with equal relative deadlines expire
        in arrival order), so a call costs O(admitted + expired), however
        long the queues are."""
        while True:
            try:
                item = self.incoming.get_nowait()
            except queue.Empty:
                break
            for req in (item if isinstance(item, list) else (item,)):
                (self.waiting if req.kind == "forward" else self.waiting_inv).append(req)
        self.waiting_max[0] = max(self.waiting_max[0], len(self.waiting))
        self.waiting_max[1] = max(self.waiting_max[1], len(self.waiting_inv))
        admitted = []
        now = time.perf_counter()
        for fifo, is_inv in ((self.waiting, False), (self.waiting_inv, True)):
            room = fwd_room
            while fifo:
                req = fifo[0]
                if req.deadline is not None and now > req.deadline:
                    fifo.popleft()
                    self._finish(req, error=TimeoutError("deadline passed before admission"))
                    continue
                if (is_inv and self.lm_running >= self.max_lm) or (not is_inv and room <= 0):
                    break  # FIFO: the rest waits behind the head
                if req.mem is None:
                    try:
                        self._validate(req)
                        req.mem = self._mem_of(req)
                        if req.mem > self.budget:
                            raise MemoryError(f"request needs {req.mem} bytes, "
                                              f"budget is {self.budget}")
                    except (ValueError, MemoryError) as e:
                        fifo.popleft()
                        self._finish(req, error=e)
                        continue
                if not self._reserve(req.mem):
                    break
                fifo.popleft()
                req.record["admit"] = now
                admitted.append(req)
                if is_inv:
                    with self.lock:
                        self.lm_running += 1
                else:
                    room -= req.record["updates"]
        return admitted

    def _reserve(self, nbytes: int) -> bool:
        with self.lock:
            if self.in_use + nbytes > self.budget:
                return False
            self.in_use += nbytes
            return True

    def _release(self, nbytes: int) -> None:
        with self.lock:
            self.in_use -= nbytes
        self._wake.set()

    def _loop(self):
        pending_fwd: list[Request] = []
        while not self._stop.is_set():
            progressed = False
            queued = sum(r.record["updates"] for r in pending_fwd)
            room = self.max_batch_updates - queued
            if self.n_inflight == 0 and not pending_fwd:
                room = self.first_batch_updates
            for req in self._admit(room):
                progressed = True
                if req.kind == "forward":
                    pending_fwd.append(req)
                else:
                    t = threading.Thread(target=self._run_invert, args=(req,), daemon=True)
                    self.lm_threads.append(t)
                    t.start()
            if pending_fwd and self.n_inflight < self.max_inflight:
                first = self.n_inflight == 0
                limit = self.first_batch_updates if first else self.max_batch_updates
                batch, total, nbytes = [], 0.0, 0
                cap = self.max_batch_bytes
                while pending_fwd:
                    nxt = pending_fwd[0].spec
                    b = 0 if pending_fwd[0].on_device else 16 * len(nxt.p) * nxt.nw
                    if batch and (total + nxt.updates() > limit or
                                  (cap is not None and nbytes + b > cap)):
                        break
                    r = pending_fwd.pop(0)
                    batch.append(r)
                    total += nxt.updates()
                    nbytes += b
                # device-resident results first, then host results smallest
                # first: the backend's copies to the host complete in that order
                batch.sort(key=lambda r: (not r.on_device, len(r.spec.p) * r.spec.nw))
                launched = time.perf_counter()
                try:
                    handle = self.backend.launch_forward([r.spec for r in batch],
                                                         [not r.on_device for r in batch])
                except Exception as e:  # launch refused: fail the batch, free its memory
                    for r in batch:
                        self._release(r.mem)
                        self._finish(r, error=e)
                else:
                    for r in batch:
                        r.record["launch"] = launched
                        r.record["batch"] = self.batches
                    self.batches += 1
                    with self.lock:
                        self.n_inflight += 1
                    self.completions.put((handle, batch))
                progressed = True
            self.lm_threads = [t for t in self.lm_threads if t.is_alive()]
            if not progressed:
                # an arrival, completion or release wakes the loop at once;
                # a wake-up that comes between the wait and the clear is not
                # lost: the next pass runs right after the clear
                self._wake.wait(self.idle_sleep)
                self._wake.clear()

    def _complete_loop(self):
        """Completion thread: waits for each launched batch in launch order,
        hands every request its host result and returns the batch's memory and
        stream slot, so the scheduler keeps launching meanwhile."""
        while True:
            item = self.completions.get()
            if item is None:
                return
            handle, batch = item
            each = getattr(self.backend, "results_each", None)
            delivered = [False] * len(batch)
            try:
                if each is not None:
                    def deliver(i, out, count):
                        delivered[i] = True
                        self._release(batch[i].mem)
                        self._finish(batch[i], result=out, nonfinite=count)

                    each(handle, deliver)
                else:
                    outs, counts = self.backend.results(handle)
                    for i, (r, out) in enumerate(zip(batch, outs)):
                        delivered[i] = True
                        self._release(r.mem)
                        self._finish(r, result=out,
                                     nonfinite=None if counts is None else counts[i])
            except Exception as e:
                for i, r in enumerate(batch):
                    if not delivered[i]:
                        self._release(r.mem)
                        self._finish(r, error=e)
            with self.lock:
                self.n_inflight -= 1
            self._wake.set()

    def _run_invert(self, req: Request):
        req.record["launch"] = time.perf_counter()
        try:
            res = self.backend.invert(req.spec)
        except Exception as e:
            res, err = None, e
        else:
            err = None
        self._release(req.mem)
        with self.lock:
            self.lm_running -= 1
        self._finish(req, result=res, error=err)
        self._wake.set()
