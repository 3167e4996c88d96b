"""Many earth models in one persistent launch (kh_forward_multi).

``reflectivity_multi`` takes models of different depths with their own
slownesses and uniform frequency grids and returns each model's (np, nw)
complex128 reflectivity.  A work item is one slowness of one model (all of
its frequency tiles, or a run of them for the few deep, wide models whose
whole rows would leave the device idle at the end of the launch).  Models
are queued by the cost of one item (interfaces x tiles), longest first, so
the persistent workgroups take the long items early.  Results are bitwise
identical to one ``kh_forward_grid_ex`` launch per model with the same block
and frequencies per lane (``omega_per_lane``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi, autotune, engine
from .layer_model import LayerModel, check_table, pack_layers

__all__ = ["ModelSpec", "MultiPlan", "plan_multi", "describe", "reflectivity_multi",
           "omega_per_lane", "DESC_DTYPE"]

# numpy mirror of kh_model_desc (include/kennett_hip.h)
DESC_DTYPE = np.dtype([("layer_off", "<i8"), ("p_off", "<i8"), ("r_off", "<i8"),
                       ("item_off", "<i8"), ("k0", "<f8"), ("dw", "<f8"), ("gamma", "<f8"),
                       ("nl", "<i4"), ("np", "<i4"), ("nw", "<i4"), ("ntiles", "<i4"),
                       ("free_surface", "<i4"), ("half_tile", "<i4"), ("tiles_per_item", "<i4"),
                       ("reserved", "<i4"), ("table_off", "<i8")])
assert DESC_DTYPE.itemsize == 96 and DESC_DTYPE.fields["table_off"][1] == 88


@dataclass
class ModelSpec:
    """One forward request: a model, its slownesses and its frequency grid
    omega_j = (k0 + j)*dw + i*gamma, j < nw."""

    model: LayerModel | np.ndarray
    p: np.ndarray
    nw: int
    dw: float = 2.0 * np.pi / 64.0
    k0: float = 1.0
    gamma: float = 0.0
    free_surface: bool = False

    def table(self) -> np.ndarray:
        """The validated (nl, 6) layer table (computed once; ``ValueError``
        for a model the kernels cannot take, see ``layer_model.check_table``)."""
        tab = self.__dict__.get("_table")
        if tab is None:
            if isinstance(self.model, LayerModel):
                tab = pack_layers(self.model)
            else:
                tab = check_table(self.model)
            self.__dict__["_table"] = tab
        return tab

    def check(self) -> None:
        """Raises ValueError unless the spec can be launched: a valid layer
        table, at least one slowness (a 1-D array) and one frequency."""
        self.table()
        if np.ndim(self.p) != 1 or len(self.p) == 0:
            raise ValueError("a forward request needs a 1-D array of at least one slowness")
        if int(self.nw) <= 0:
            raise ValueError(f"a forward request needs at least one frequency, got nw={self.nw}")

    @property
    def n_layers(self) -> int:
        return self.table().shape[0]

    def updates(self) -> int:
        return len(self.p) * self.nw * (self.n_layers - 1)


@dataclass
class MultiPlan:
    desc: torch.Tensor          # device uint8 view of nmodels kh_model_desc
    layers: torch.Tensor        # (rows, 6) float64 on the device
    p: torch.Tensor             # all slownesses
    R: torch.Tensor             # all outputs, complex128, flat
    r_off: np.ndarray           # per model (input order)
    shapes: list[tuple[int, int]]
    total_items: int
    max_nl: int
    block: int
    wpt: int
    counter: torch.Tensor
    nonfinite: torch.Tensor     # int32 (nmodels,) NaN/Inf output counts, queue order
    order: list[int]            # queue slot -> input index
    records: int                # table records: sum of np * nl
    tables: torch.Tensor | None  # uint8 table workspace, or None: tables staged in LDS


def _to_device(arr: np.ndarray, dev, non_blocking: bool) -> torch.Tensor:
    t = torch.from_numpy(arr)
    if non_blocking:  # pinned source: the copy is queued on the current stream
        t = t.pin_memory()
    return t.to(dev, non_blocking=non_blocking)


# Cost of a tile at half the frequencies per lane relative to a full tile:
# the per-layer phase of a lane's first frequency and the loop overhead are
# spread over half as many updates (64 x 4 runs about 10 % more VALU
# instructions per update than 64 x 8).
HALF_TILE_COST = 0.55

# Largest item, as a fraction of the launch's work per resident workgroup:
# rows of models whose whole-row items would be costlier are split into runs
# of tiles.  With whole rows the drain at the end of a launch of a few deep,
# wide models leaves most of the device idle (a batch of 8 such models of
# the mixed stream: 25 % over the ideal time in a list-scheduling model of
# the queue).  The tables are staged once per slowness before the sweep, so
# a finer split costs no restaging; over the ten batches of the 400-request
# stream launched one by one on MI355X the total is flat between 0.02 and
# 0.15 (53.1-53.4 ms) and 1 % higher at 0.25.
TAIL_FRACTION = 0.1

_RESIDENT: dict = {}


def table_bytes(records: int) -> int:
    """Bytes of the table workspace a launch of `records` (sum of np * nl)
    table records needs."""
    return int(_abi.lib().kh_multi_table_bytes(int(records))) if records > 0 else 0


def resident_workgroups(block: int, wpt: int, max_nl: int, tables: bool = True,
                        dev=None) -> int:
    """Workgroups a persistent launch of this shape starts on the
    staged-table or in-LDS path (cached)."""
    dev = _abi.device(dev)
    key = (dev.index, block, wpt, max_nl if not tables else 0, bool(tables))
    n = _RESIDENT.get(key)
    if n is None:
        out = ctypes.c_int(0)
        with torch.cuda.device(dev):
            _abi.check(_abi.lib().kh_multi_workgroups(block, wpt, max_nl, int(bool(tables)),
                                                      ctypes.byref(out)),
                       "kh_multi_workgroups")
        n = _RESIDENT[key] = max(int(out.value), 1)
    return n


def omega_per_lane(nw: int, block: int, wpt: int) -> int:
    """Frequencies per lane a model of nw frequencies gets in a launch of
    shape (block, wpt): wpt, or wpt // 2 when half-width tiles cost less
    than full ones (a model that fills at most half a tile, e.g. 256
    frequencies at 64 x 8, would otherwise leave half the lanes idle).
    Either way the output is bitwise that of the same block size."""
    if wpt < 2:
        return wpt
    full = -(-nw // (block * wpt))
    half = -(-nw // (block * (wpt // 2)))
    return wpt // 2 if half * HALF_TILE_COST < full else wpt


# Batches of at least this many updates stage their tables with a kernel of
# their own (a workspace of 256 B per (model, slowness, layer)) and sweep
# without LDS; smaller ones stage them in LDS within the one sweep kernel.
# The served 400-request stream (ten batches of about 9e8 updates) runs
# 3.7 % faster on staged tables; a small batch launched while other work
# occupies the device waits for free compute units once instead of twice
# (the open-loop stream's forward p50 went from 2.2 to 4.4 ms with staged
# tables for every batch).  With the threshold at 5e8 one open-loop run gave
# forward p50 2.8 ms and inversion p50 1.21 s, against 2.3 ms and 0.83-0.90 s
# in two runs at 2e8.
TABLES_MIN_UPDATES = 2e8


def describe(specs: list[ModelSpec], block: int, wpt: int, slots: int):
    """Host side of a persistent launch: the descriptors in queue order, the
    queue order (slot -> input index), the work item count and the table
    record count, for `slots` resident workgroups.

    Each model gets its frequencies per lane (omega_per_lane), its tile
    count and, when a whole row would cost more than TAIL_FRACTION of the
    work per workgroup, runs of tiles per item.  Models are queued longest
    item first (then by the model's total), so the smallest items are left
    for the drain at the end of the launch; item_off and table_off are the
    prefix sums of the items and of np * nl in that order."""
    for s in specs:
        s.check()
    tabs = [s.table() for s in specs]
    lanes = [omega_per_lane(s.nw, block, wpt) for s in specs]
    ntiles = [-(-s.nw // (block * f)) for s, f in zip(specs, lanes)]
    # cost of one tile of a row in tile steps (a half-width tile costs
    # HALF_TILE_COST of a full one)
    tile_cost = [(HALF_TILE_COST if f < wpt else 1.0) * (t.shape[0] - 1)
                 for f, t in zip(lanes, tabs)]
    cap = TAIL_FRACTION * sum(len(s.p) * n * c for s, n, c in zip(specs, ntiles, tile_cost)) / slots
    tpi = [n if n * c <= cap else max(1, int(cap // c)) for n, c in zip(ntiles, tile_cost)]
    item_cost = [k * c for k, c in zip(tpi, tile_cost)]
    total_cost = [len(s.p) * n * c for s, n, c in zip(specs, ntiles, tile_cost)]
    order = sorted(range(len(specs)), key=lambda i: (-item_cost[i], -total_cost[i]))
    desc = np.zeros(len(specs), DESC_DTYPE)
    layer_off = np.cumsum([0] + [t.shape[0] for t in tabs])
    p_off = np.cumsum([0] + [len(s.p) for s in specs])
    r_off = np.cumsum([0] + [len(s.p) * s.nw for s in specs])
    item = 0
    records = 0
    for slot, i in enumerate(order):
        s = specs[i]
        whole = tpi[i] >= ntiles[i]
        desc[slot] = (layer_off[i], p_off[i], r_off[i], item, float(s.k0), float(s.dw),
                      float(s.gamma), tabs[i].shape[0], len(s.p), s.nw, ntiles[i],
                      int(bool(s.free_surface)), int(lanes[i] < wpt), 0 if whole else tpi[i],
                      0, records)
        item += len(s.p) * -(-ntiles[i] // tpi[i])
        records += len(s.p) * tabs[i].shape[0]
    return desc, order, item, records


def plan_multi(specs: list[ModelSpec], block: int, wpt: int, dev=None,
               non_blocking: bool = False, tables: bool | None = None) -> MultiPlan:
    """Pack specs into device buffers and descriptors for one launch shape.
    With ``non_blocking`` the host->device copies go through pinned memory and
    are only queued on the current stream (the host does not wait).
    ``tables`` picks the staged-table path (default: batches of at least
    TABLES_MIN_UPDATES updates)."""
    dev = _abi.device(dev)
    tabs = [s.table() for s in specs]
    max_nl = max(t.shape[0] for t in tabs) if tabs else 2
    if tables is None:
        tables = sum(s.updates() for s in specs) >= TABLES_MIN_UPDATES
    slots = resident_workgroups(block, wpt, max_nl, tables, dev) if specs else 1
    desc, order, items, records = describe(specs, block, wpt, slots)
    r_off = np.cumsum([0] + [len(s.p) * s.nw for s in specs])
    layers = _to_device(np.concatenate(tabs, axis=0), dev, non_blocking)
    p = _to_device(np.concatenate([np.asarray(s.p, np.float64) for s in specs])
                   if specs else np.zeros(0), dev, non_blocking)
    R = torch.empty(int(r_off[-1]), dtype=torch.complex128, device=dev)
    return MultiPlan(desc=_to_device(desc.view(np.uint8), dev, non_blocking), layers=layers, p=p, R=R,
                     r_off=r_off, shapes=[(len(s.p), s.nw) for s in specs], total_items=items,
                     max_nl=max_nl, block=block, wpt=wpt,
                     # both are reset by the launch itself (no separate fill kernels)
                     counter=torch.empty(1, dtype=torch.int64, device=dev),
                     nonfinite=torch.empty(max(len(specs), 1), dtype=torch.int32, device=dev),
                     order=order, records=records,
                     tables=(torch.empty(max(int(table_bytes(records)), 1), dtype=torch.uint8,
                                         device=dev) if tables else None))


def launch(plan: MultiPlan, workgroups: int = 0) -> torch.Tensor:
    """Enqueue the persistent launch of a plan on the current stream."""
    dev = plan.R.device
    rc = _abi.lib().kh_forward_multi(_abi.ptr(plan.desc), len(plan.shapes), plan.total_items,
                                     plan.max_nl, _abi.ptr(plan.layers), _abi.ptr(plan.p),
                                     _abi.ptr(plan.R), plan.block, plan.wpt, int(workgroups),
                                     plan.records,
                                     _abi.ptr(plan.tables) if plan.tables is not None else None,
                                     plan.tables.numel() if plan.tables is not None else 0,
                                     _abi.ptr(plan.counter), _abi.ptr(plan.nonfinite),
                                     _abi.stream_handle(dev))
    _abi.check(rc, "kh_forward_multi")
    return plan.R


def nonfinite_counts(plan: MultiPlan, counts_by_slot) -> list[int]:
    """Per input model, the NaN/Inf output count from the plan's queue-order
    counts (``plan.nonfinite`` after the launch, as a host sequence)."""
    out = [0] * len(plan.shapes)
    for slot, i in enumerate(plan.order):
        out[i] = int(counts_by_slot[slot])
    return out


def split(plan: MultiPlan) -> list[torch.Tensor]:
    return [plan.R[int(plan.r_off[i]):int(plan.r_off[i + 1])].view(n_p, nw)
            for i, (n_p, nw) in enumerate(plan.shapes)]


def reflectivity_multi(specs: list[ModelSpec], block: int | None = None,
                       omega_per_item: int | None = None, dev=None,
                       tables: bool | None = None) -> list[torch.Tensor]:
    """Each spec's R (np, nw) on the device, all from one persistent launch.
    The launch shape comes from the autotuner unless given; ``tables`` as in
    plan_multi."""
    if not specs:
        return []
    if block is None or omega_per_item is None:
        block, omega_per_item = autotune.multi_shape(specs, dev=dev)
    plan = plan_multi(specs, block, omega_per_item, dev, tables=tables)
    launch(plan)
    return split(plan)
