"""The batched launch replayed as a HIP graph: the second launch of a shape
is captured and later ones replay it.  Every replay, with other work of the
program between them, equals the launch without graphs: bit for bit in FP64
and in mixed-precision Stokes I (nothing there depends on launch timing),
within the mixed bar against FP64 in full Stokes (the split of long rays
depends on when the launch's tail sets in)."""
import numpy as np
import pytest

from jetshow_amd import abi, models
from tests.test_gpu_paths import _mixed_errors

pytestmark = pytest.mark.gpu


def _items(precision, stokes):
    out = []
    for k, side in enumerate((32, 48, 40)):
        j, im, p = models.conical_toroidal_model(n_side=side, pixel_pc=0.05 + 0.01 * k)
        p.precision = precision
        p.stokes = stokes
        if precision == abi.PREC_MIXED:
            p.stokes_abs_tol, p.stokes_rel_tol = 1e-20, 1e-7
        out.append((j, im, p))
    return out


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_graph_replays_equal_eager(ctx, precision):
    items = _items(precision, abi.STOKES_I)
    ctx.set_option("graph", 0)
    ref = ctx.observe_batch(items)
    ctx.set_option("graph", 1)
    c0 = ctx.get_option("stat_graph_captures")
    l0 = ctx.get_option("stat_graph_launches")
    other = models.conical_toroidal_model(n_side=16, pixel_pc=0.1)
    for rep in range(4):
        print("graph replay %d, precision %d" % (rep, precision), flush=True)
        got = ctx.observe_batch(items)
        ctx.observe(*other)                      # eager work between replays
        for g, r in zip(got, ref):
            for k in ("tau", "I", "l", "flags", "n_rhs"):
                assert np.array_equal(g[k], r[k]), (rep, k)
    assert ctx.get_option("stat_graph_captures") - c0 == 1
    assert ctx.get_option("stat_graph_launches") - l0 == 3


def test_graph_replays_full_stokes_mixed(ctx):
    j, im, p = models.paraboloid_helical_model(n_side=64, pixel_pc=0.08)
    p.stokes_abs_tol = 1e-20
    g64 = ctx.observe(j, im, p)
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    pm.flags |= 0x100 | 0x200
    ctx.set_option("graph", 1)
    c0 = ctx.get_option("stat_graph_captures")
    for rep in range(4):
        print("graph replay full Stokes %d" % rep, flush=True)
        gm = ctx.observe(j, im, pm)
        rel, quv, _ = _mixed_errors(g64, gm)
        assert rel <= 1e-5 and quv <= 5e-5, (rep, rel, quv)
    assert ctx.get_option("stat_graph_captures") - c0 == 1
