"""The autotuner (jetshow_amd/tune.py): sweep logic and cache on the CPU with a
stand-in context whose launch time is a known function of the options; on the
GPU, a two-point sweep over a small batch, whose FP64 planes do not depend on
the launch shape."""
import json

import numpy as np
import pytest

from jetshow_amd import abi, models, tune


class FakeCtx:
    """Options dict; observe_batch 'takes' a time set by the options."""

    def __init__(self):
        self.opts = {"blocks_per_cu": 4, "queue_block": 64, "piece_budget": 64}
        self.calls = 0

    def set_option(self, k, v):
        self.opts[k] = v

    def get_option(self, k):
        return self.opts.get(k, 0)

    def observe_batch(self, items):
        import time
        self.calls += 1
        # fastest at blocks_per_cu 2, queue_block 256
        t = 0.002 * abs(self.opts["blocks_per_cu"] - 2) + 0.001 * (self.opts["queue_block"] != 256)
        time.sleep(t)
        return []


def _items():
    j, im, p = models.conical_toroidal_model(n_side=16, pixel_pc=0.2)
    return [(j, im, p)]


def test_sweep_picks_fastest_and_restores():
    ctx = FakeCtx()
    best, table, complete = tune.sweep(ctx, _items(), {"blocks_per_cu": [1, 2, 4], "queue_block": [64, 256]},
                                       reps=2)
    assert complete
    assert best == {"blocks_per_cu": 2, "queue_block": 256}
    assert len(table) == 6 and table[0][1] == best
    # options are restored after the sweep
    assert ctx.opts["blocks_per_cu"] == 4 and ctx.opts["queue_block"] == 64


def test_tune_caches_and_applies(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "_device_name", lambda: "test-device")
    cache = str(tmp_path / "tune.json")
    ctx = FakeCtx()
    space = {"blocks_per_cu": [2, 4], "queue_block": [64, 256]}
    best = tune.tune(ctx, _items(), space, reps=1, cache=cache)
    assert best == {"blocks_per_cu": 2, "queue_block": 256}
    assert ctx.opts["blocks_per_cu"] == 2 and ctx.opts["queue_block"] == 256
    data = json.load(open(cache))
    key = "test-device|" + tune.signature(_items())
    assert data[key]["options"] == best and data[key]["tried"] == 4
    # a second context applies the cached shape without timing anything
    ctx2 = FakeCtx()
    assert tune.apply_cached(ctx2, _items(), cache=cache) == best
    assert ctx2.calls == 0 and ctx2.opts["blocks_per_cu"] == 2


def test_truncated_search_is_not_trusted(tmp_path, monkeypatch):
    """A search cut short by its budget starts from the current settings
    (every option near the defaults first), reports itself incomplete, and
    tune() searches again next time instead of reusing it."""
    monkeypatch.setattr(tune, "_device_name", lambda: "test-device")
    cache = str(tmp_path / "tune.json")
    ctx = FakeCtx()
    best, table, complete = tune.sweep(ctx, _items(), {"blocks_per_cu": [1, 2, 4], "queue_block": [64, 256]},
                                       reps=1, budget_s=0.0)
    assert not complete
    assert table[-1][1] == {"blocks_per_cu": 4, "queue_block": 64}   # the defaults come first
    space = {"blocks_per_cu": [2, 4], "queue_block": [64, 256]}
    tune.tune(ctx, _items(), space, reps=1, cache=cache, budget_s=0.0)
    key = "test-device|" + tune.signature(_items())
    assert json.load(open(cache))[key]["complete"] is False
    calls = ctx.calls
    best = tune.tune(ctx, _items(), space, reps=1, cache=cache)
    assert ctx.calls > calls and best == {"blocks_per_cu": 2, "queue_block": 256}
    assert json.load(open(cache))[key]["complete"] is True


def test_signature_ignores_values_but_not_shapes():
    (j, im, p), = _items()
    a = tune.signature([(j, im, p)])
    p2 = models.copy_struct(p)
    p2.nu = p.nu * 2.0
    assert tune.signature([(j, im, p2)]) == a
    im2 = abi.ImageDesc.from_buffer_copy(im)
    im2.n_along = im.n_along * 2
    assert tune.signature([(j, im2, p)]) != a


@pytest.mark.gpu
def test_tune_on_gpu_shapes_agree(ctx, tmp_path):
    """Two launch shapes timed on a small FP64 batch; the planes are identical
    whichever shape runs (FP64 is never split and every ray's arithmetic is
    independent of which lane runs it)."""
    items = []
    for j, im, p in (models.conical_toroidal_model(n_side=32, pixel_pc=0.05),
                     models.paraboloid_helical_model(n_side=16, pixel_pc=0.3)):
        q = models.copy_struct(p)
        q.flags |= 0x100
        items.append((j, im, q))
    saved = {k: ctx.get_option(k) for k in ("blocks_per_cu", "queue_block")}
    try:
        outs = {}
        for bpc in (1, 4):
            ctx.set_option("blocks_per_cu", bpc)
            outs[bpc] = ctx.observe_batch(items)
        for a, b in zip(outs[1], outs[4]):
            for k in ("I", "tau", "flags", "n_rhs"):
                assert np.array_equal(a[k], b[k]), k
        for k, v in saved.items():
            ctx.set_option(k, v)
        best = tune.tune(ctx, items, {"blocks_per_cu": [1, 4], "queue_block": [64]}, reps=1,
                         cache=str(tmp_path / "t.json"))
        assert best["blocks_per_cu"] in (1, 4)
        assert ctx.get_option("blocks_per_cu") == best["blocks_per_cu"]
    finally:
        for k, v in saved.items():
            ctx.set_option(k, v)
