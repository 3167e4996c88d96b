"""Host side of the persistent multi-model launch, on the CPU: the
descriptors batch.describe builds are what kh_forward_multi decodes.  Every
(model, slowness, frequency tile) is covered by exactly one work item when
the items are decoded as the kernel decodes them, the offsets are the
prefix sums the kernels index with, and the library's own tile and item
counts agree (host functions of the C-ABI, no GPU)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from seismicinversion_amd import _abi, batch
from seismicinversion_amd.serve import stream_requests
from seismicinversion_amd.workloads import serving_stream


def _stream_batch(n=60, seed=5):
    return [s for _, s in stream_requests(serving_stream(0.0, n, seed=seed, lm_fraction=0.0))]


def _decode(d, local):
    """multi_kernel's decode of a model's local item number."""
    tpi = d["tiles_per_item"] if d["tiles_per_item"] > 0 else d["ntiles"]
    nch = -(-int(d["ntiles"]) // int(tpi))
    ip, ch = divmod(local, nch)
    t0 = ch * tpi
    return ip, t0, min(int(d["ntiles"]), t0 + tpi)


@pytest.mark.parametrize("tail", [None, 1e-9, 10.0])
@pytest.mark.parametrize("block,wpt", [(64, 8), (128, 4), (256, 2), (64, 1)])
def test_items_cover_every_tile_once(monkeypatch, tail, block, wpt):
    if tail is not None:
        monkeypatch.setattr(batch, "TAIL_FRACTION", tail)
    specs = _stream_batch()
    desc, order, total, records = batch.describe(specs, block, wpt, slots=2048)
    lib = _abi.lib()
    lib.kh_multi_items.restype = ctypes.c_longlong
    assert sorted(order) == list(range(len(specs)))
    item = rec = 0
    for slot, i in enumerate(order):
        d, s = desc[slot], specs[i]
        f = batch.omega_per_lane(s.nw, block, wpt)
        assert (d["nl"], d["np"], d["nw"]) == (s.n_layers, len(s.p), s.nw)
        assert d["half_tile"] == int(f < wpt)
        assert d["ntiles"] == -(-s.nw // (block * f)) == lib.kh_multi_tiles(s.nw, block, wpt,
                                                                             int(d["half_tile"]))
        assert d["item_off"] == item and d["table_off"] == rec
        n_items = lib.kh_multi_items(len(s.p), int(d["ntiles"]), int(d["tiles_per_item"]))
        covered = np.zeros((len(s.p), int(d["ntiles"])), int)
        for local in range(n_items):
            ip, t0, t1 = _decode(d, local)
            assert 0 <= ip < len(s.p) and 0 <= t0 < t1 <= d["ntiles"]
            covered[ip, t0:t1] += 1
        assert (covered == 1).all(), f"model {i}: tiles covered {covered.min()}..{covered.max()}"
        item += n_items
        rec += len(s.p) * s.n_layers
    assert total == item and records == rec
    if tail == 1e-9:  # every multi-tile row split to single tiles
        assert all(d["tiles_per_item"] in (0, 1) and (d["tiles_per_item"] == 1 or d["ntiles"] == 1)
                   for d in desc)
    if tail == 10.0:  # whole rows
        assert (desc["tiles_per_item"] == 0).all()


def test_queue_is_longest_item_first():
    specs = _stream_batch(80, seed=9)
    desc, order, _, _ = batch.describe(specs, 64, 8, slots=2048)
    cost = []
    for d in desc:
        tpi = d["tiles_per_item"] if d["tiles_per_item"] > 0 else d["ntiles"]
        cost.append(tpi * (batch.HALF_TILE_COST if d["half_tile"] else 1.0) * (d["nl"] - 1))
    assert all(a >= b for a, b in zip(cost, cost[1:]))


def test_offsets_index_the_concatenated_inputs():
    specs = _stream_batch(20, seed=3)
    desc, order, _, _ = batch.describe(specs, 64, 8, slots=2048)
    layers = np.concatenate([s.table() for s in specs])
    p = np.concatenate([np.asarray(s.p, np.float64) for s in specs])
    for slot, i in enumerate(order):
        d, s = desc[slot], specs[i]
        np.testing.assert_array_equal(layers[d["layer_off"]:d["layer_off"] + d["nl"]], s.table())
        np.testing.assert_array_equal(p[d["p_off"]:d["p_off"] + d["np"]], s.p)
        assert d["r_off"] == sum(len(specs[j].p) * specs[j].nw for j in range(i))


def _good_table(nl=5):
    from seismicinversion_amd.layer_model import pack_layers
    from seismicinversion_amd.workloads import seeded_model

    return pack_layers(seeded_model(nl, 3))


@pytest.mark.parametrize("bad", [
    _good_table(3)[:, :5],                        # five columns
    np.hstack([_good_table(3), np.ones((3, 1))]),  # seven columns
    _good_table(3).ravel(),                       # 1-D
    _good_table(2)[:1],                           # one layer
    np.vstack([_good_table(2)] * 121)[:241],      # 241 layers, above the kernels' cap
])
def test_malformed_tables_are_rejected_before_launch(bad):
    """An (nl, 5), (nl, 7), 1-D, one-layer or 241-layer table never reaches
    the kernels (which index every row as 6 doubles and stage at most 240
    layers): ModelSpec.table() and describe() raise ValueError."""
    spec = batch.ModelSpec(model=bad, p=np.array([0.1, 0.2]), nw=16)
    with pytest.raises(ValueError):
        spec.table()
    with pytest.raises(ValueError):
        batch.describe([batch.ModelSpec(model=bad, p=np.array([0.1]), nw=16)], 64, 8, slots=64)


@pytest.mark.parametrize("column,value", [(0, 0.0), (1, -1.0), (2, 0.0), (4, 0.0), (5, 0.0),
                                          (3, np.nan)])
def test_table_values_are_checked(column, value):
    tab = _good_table(4)
    tab[2, column] = value
    with pytest.raises(ValueError):
        batch.ModelSpec(model=tab, p=np.array([0.1]), nw=8).table()


@pytest.mark.parametrize("column", range(6))
def test_nan_anywhere_is_rejected(column):
    """Also in the water layer, whose Q_beta the kernels never read."""
    tab = _good_table(4)
    tab[0, column] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        batch.ModelSpec(model=tab, p=np.array([0.1]), nw=8).table()


def test_empty_grids_are_rejected():
    """nw = 0 or no slowness: ValueError from describe() (not a division by
    zero in the tile arithmetic)."""
    tab = _good_table(4)
    for spec in (batch.ModelSpec(model=tab, p=np.array([0.1]), nw=0),
                 batch.ModelSpec(model=tab, p=np.zeros(0), nw=8),
                 batch.ModelSpec(model=tab, p=np.array([[0.1]]), nw=8)):
        with pytest.raises(ValueError):
            batch.describe([spec], 64, 8, slots=64)


def test_a_valid_table_passes_unchanged():
    tab = _good_table(6)
    spec = batch.ModelSpec(model=tab, p=np.array([0.1]), nw=8)
    np.testing.assert_array_equal(spec.table(), tab)
    desc, order, total, records = batch.describe([spec], 64, 8, slots=64)
    assert total == 1 and records == 6 and desc[0]["nl"] == 6
