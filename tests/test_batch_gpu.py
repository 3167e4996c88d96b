"""Many uneven models in one persistent launch (kh_forward_multi): every
model's output equals the one-model-per-launch path with the same block and
frequencies per lane bit for bit, and the oracle to the parity bar."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import kennett_numpy as K
from parity import accuracy, assert_accurate
from seismicinversion_amd import autotune, batch, engine
from seismicinversion_amd.workloads import seeded_model, uneven_batch_models

pytestmark = pytest.mark.gpu


def specs_c3(n_models, nw=1024, seed=42):
    return [batch.ModelSpec(model=m, p=p, nw=nw, dw=2 * np.pi / 64.0, k0=1.0)
            for m, p in uneven_batch_models(n_models, seed)]


@pytest.mark.parametrize("tables", [False, True])
@pytest.mark.parametrize("block,wpt", [(256, 2), (128, 1), (256, 4), (64, 2), (64, 8), (128, 8)])
def test_multi_equals_plain_bitwise(block, wpt, tables):
    """Both ways of staging the tables (LDS per item, or a staging kernel
    and scalar loads) give the plain launch's bits."""
    specs = specs_c3(20)
    specs[3].free_surface = True
    specs[5].gamma = np.pi / 64.0
    outs = batch.reflectivity_multi(specs, block=block, omega_per_item=wpt, tables=tables)
    for s, R in zip(specs, outs):
        L = engine.layers_tensor(s.model)
        Rp = engine.forward_grid(L, torch.from_numpy(s.p), s.k0, s.dw, s.gamma, s.nw,
                                 s.free_surface, block=block,
                                 omega_per_item=batch.omega_per_lane(s.nw, block, wpt))
        assert torch.equal(R, Rp), "persistent launch differs from the plain launch"


def test_multi_vs_oracle_and_repeatability():
    specs = specs_c3(12, nw=256)
    outs = batch.reflectivity_multi(specs, block=256, omega_per_item=2)
    outs2 = batch.reflectivity_multi(specs, block=256, omega_per_item=2)
    for s, R, R2 in zip(specs, outs, outs2):
        assert torch.equal(R, R2)
        rows = np.linspace(0, len(s.p) - 1, 4).astype(int)
        w = (s.k0 + np.arange(s.nw)) * s.dw + 1j * s.gamma
        assert_accurate(R.cpu().numpy()[rows], accuracy(s.model, s.p[rows], w, s.free_surface),
                        f"multi nl={s.n_layers}")


def test_full_uneven_batch_and_autotuner():
    """256 models, nl in [3, 60], np in [32, 512], 1024 frequencies."""
    specs = specs_c3(256)
    autotune.clear()
    shape = autotune.multi_shape(specs)
    assert shape in autotune.CANDIDATES
    assert autotune.multi_shape(specs) == shape  # cached
    outs = batch.reflectivity_multi(specs)
    for i in (0, 100, 255):
        s = specs[i]
        Rp = engine.forward_grid(engine.layers_tensor(s.model), torch.from_numpy(s.p), s.k0, s.dw,
                                 s.gamma, s.nw, False, block=shape[0],
                                 omega_per_item=batch.omega_per_lane(s.nw, *shape))
        assert torch.equal(outs[i], Rp)
    assert all(torch.isfinite(torch.view_as_real(R)).all() for R in outs)


@pytest.mark.parametrize("tables", [False, True])
def test_multi_mixed_grids_bitwise(tables):
    """Models with their own frequency counts (one partial tile up to eight
    full ones per slowness at 64 x 8, half-width tiles of 4 frequencies per
    lane for the counts that waste less that way) and start frequencies in
    one launch: each equals its own plain launch bit for bit."""
    specs = specs_c3(10)
    for s, nw, k0 in zip(specs, (100, 700, 4096, 512, 513, 1, 2048, 3000, 64, 256),
                         (1.0, 3.0, 1.0, 7.0, 1.0, 2.0, 1.0, 5.0, 1.0, 1.0)):
        s.nw, s.k0 = nw, k0
    lanes = [batch.omega_per_lane(s.nw, 64, 8) for s in specs]
    assert lanes == [4, 4, 8, 8, 4, 4, 8, 8, 4, 4]
    outs = batch.reflectivity_multi(specs, block=64, omega_per_item=8, tables=tables)
    for s, R, f in zip(specs, outs, lanes):
        L = engine.layers_tensor(s.model)
        Rp = engine.forward_grid(L, torch.from_numpy(s.p), s.k0, s.dw, s.gamma, s.nw,
                                 s.free_surface, block=64, omega_per_item=f)
        assert R.shape == (len(s.p), s.nw)
        assert torch.equal(R, Rp), f"nw {s.nw}: persistent launch differs from the plain launch"


@pytest.mark.parametrize("tables", [False, True])
@pytest.mark.parametrize("block,wpt", [(256, 2), (64, 2)])
def test_nonfinite_counts_per_model(block, wpt, tables):
    """A NaN slowness makes its whole row NaN; the launch counts the non-finite
    outputs of each model (over all frequency tiles of the row: three at
    64 x 2) and the others stay at zero."""
    specs = specs_c3(6, nw=300)
    specs[2].p = specs[2].p.copy()
    specs[2].p[1] = np.nan
    plan = batch.plan_multi(specs, block, wpt, tables=tables)
    batch.launch(plan)
    counts = batch.nonfinite_counts(plan, plan.nonfinite.cpu().numpy())
    assert counts == [300 if i == 2 else 0 for i in range(6)]
    R = batch.split(plan)[2]
    assert not torch.isfinite(torch.view_as_real(R[1])).any()
    # a second launch of the same plan resets the counts before counting again
    batch.launch(plan)
    assert batch.nonfinite_counts(plan, plan.nonfinite.cpu().numpy())[2] == 300


@pytest.mark.parametrize("tables", [False, True])
@pytest.mark.parametrize("tail", [1e-9, None])
def test_tile_runs_for_a_thin_batch(monkeypatch, tail, tables):
    """Three deep models with many frequency tiles and few slownesses: whole
    rows would leave most workgroups idle, so their rows are split into runs
    of tiles (one tile per item with a tiny tail fraction).  Outputs equal
    the plain launches bit for bit, and NaN counts add up over the runs."""
    if tail is not None:
        monkeypatch.setattr(batch, "TAIL_FRACTION", tail)
    specs = [batch.ModelSpec(model=seeded_model(nl, 40 + nl), p=np.linspace(0.01, 0.6, npp),
                             nw=nw, dw=2 * np.pi / 64.0, k0=1.0, free_surface=(nl == 60))
             for nl, npp, nw in ((90, 300, 4096), (60, 200, 3000), (20, 50, 700))]
    specs[1].p[7] = np.nan
    plan = batch.plan_multi(specs, 64, 8, tables=tables)
    desc = plan.desc.cpu().numpy().view(batch.DESC_DTYPE)
    tpi = {int(d["nl"]): int(d["tiles_per_item"]) for d in desc}
    assert tpi[90] > 0 and tpi[60] > 0, tpi  # rows split
    if tail is not None:
        assert tpi == {90: 1, 60: 1, 20: 1}
    batch.launch(plan)
    assert batch.nonfinite_counts(plan, plan.nonfinite.cpu().numpy()) == [0, 3000, 0]
    for s, R in zip(specs, batch.split(plan)):
        Rp = engine.forward_grid(engine.layers_tensor(s.model), torch.from_numpy(s.p), s.k0,
                                 s.dw, s.gamma, s.nw, s.free_surface, block=64,
                                 omega_per_item=batch.omega_per_lane(s.nw, 64, 8))
        assert torch.equal(torch.nan_to_num(R), torch.nan_to_num(Rp)), f"nl={s.n_layers}"


@pytest.mark.parametrize("tables", [False, True])
def test_deepest_and_shallowest_models_together(tables):
    """A model at the layer limit next to a two-interface one, both ways of
    staging the tables: each equals its plain launch bit for bit."""
    nl_max = 240  # KH_MAX_LAYERS
    specs = [batch.ModelSpec(model=seeded_model(nl, 7 + nl), p=np.linspace(0.02, 0.7, npp),
                             nw=nw, dw=2 * np.pi / 64.0, k0=1.0)
             for nl, npp, nw in ((nl_max, 40, 1024), (3, 300, 256))]
    outs = batch.reflectivity_multi(specs, block=64, omega_per_item=8, tables=tables)
    for s, R in zip(specs, outs):
        Rp = engine.forward_grid(engine.layers_tensor(s.model), torch.from_numpy(s.p), s.k0, s.dw,
                                 s.gamma, s.nw, False, block=64,
                                 omega_per_item=batch.omega_per_lane(s.nw, 64, 8))
        assert torch.equal(R, Rp), f"nl={s.n_layers}"
        rows = np.array([0, len(s.p) - 1])
        w = (s.k0 + np.arange(s.nw)) * s.dw
        assert_accurate(R.cpu().numpy()[rows], accuracy(s.model, s.p[rows], w, False),
                        f"nl={s.n_layers}")
