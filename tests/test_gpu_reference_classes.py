"""GPU parity of the further reference classes (ParabaloidCone and the B, N, V
fields of include/BField.h:56-164, NField.h:20-60, VField.h:17-176): field,
frame and coefficient values at points (jsx_fields_host), Geometry::hit,
FP64 images, the tau stripe and batched launches bit for bit against the
oracle; the mixed-precision images against FP64."""
import math

import numpy as np
import pytest

from jetshow_amd import abi, models
from oracle import oracle
from tests.test_gpu_paths import _mixed_errors

pytestmark = pytest.mark.gpu

PC = abi.PC
CASES = models.reference_class_models()
NAMES = [c[0] for c in CASES]


def case(name):
    return {n: (j, im, p) for n, j, im, p in CASES}[name]


def points_and_los(n=600, seed=3):
    rng = np.random.default_rng(seed)
    z = rng.uniform(0.05, 10.0, n) * PC
    rho = z * np.tan(rng.uniform(0.0, 0.06, n))
    phi = rng.uniform(-np.pi, np.pi, n)
    pts = np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)
    th = rng.uniform(0.02, 1.2, n)
    nl = np.stack([np.sin(th), np.zeros(n), np.cos(th)], axis=1)
    return pts, nl


@pytest.mark.parametrize("name", NAMES)
def test_point_values_bit_exact(ctx, name):
    j, im, p = case(name)
    pts, nl = points_and_los()
    g = ctx.fields(j, pts, nl, p.nu)
    o = oracle.point_values(j, p.nu, pts, nl, math=oracle.MATH_JSX)
    assert np.array_equal(g, o, equal_nan=True), name
    assert np.isfinite(g[:, 18:29]).all()
    if name == "ShearedCentralVField":
        q = models.copy_struct(p)
        q.flags = abi.OPT_SHEARED_CENTRAL_REF_BUG
        gb = ctx.fields(j, pts, nl, p.nu, q)
        ob = oracle.point_values(j, p.nu, pts, nl, opt=abi.OPT_SHEARED_CENTRAL_REF_BUG, math=oracle.MATH_JSX)
        assert np.array_equal(gb, ob, equal_nan=True)
        assert np.isnan(gb[:, 3]).all()


def test_paraboloid_cone_hit_bit_exact(ctx):
    j, im, p = case("ParabaloidCone")
    rng = np.random.default_rng(11)
    n = 4000
    o = np.stack([rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n), rng.uniform(-1.0, 10.0, n)], 1) * PC
    th = rng.uniform(0.02, 1.5, n)
    d = np.stack([-np.sin(th), rng.uniform(-0.1, 0.1, n), -np.cos(th)], 1)
    rays = np.concatenate([o - 30 * PC * d, d], axis=1)
    ns_g, seg_g = ctx.hit(j, rays, p)
    ns_o, seg_o = oracle.hit(j, rays, p)
    assert np.array_equal(ns_g, ns_o)
    assert np.array_equal(seg_g, seg_o)
    assert ns_g.sum() > 300


@pytest.mark.parametrize("name", NAMES)
def test_fp64_image_bit_exact(ctx, name):
    j, im, p = case(name)
    print("case", name, flush=True)
    g = ctx.observe(j, im, p)
    o = oracle.observe(j, im, p, None, math=oracle.MATH_JSX, threads=8)
    for k in ("tau", "I", "Q", "U", "V", "l", "flags", "n_rhs", "n_steps"):
        assert np.array_equal(g[k].reshape(-1), o[k]), (name, k)
    assert ((o["flags"] & abi.PIX_HIT) != 0).sum() > 100


DISCONTINUOUS = {"ForceFreeCylindricalBField", "CompositeRandomScalarBFieldZ", "CompositeBKNField",
                 "SheathFlatVField", "SheathCentralVField", "AccParabolicConstConeVField",
                 "ShearedAccParabolicConstConeVField"}


@pytest.mark.parametrize("name", NAMES)
def test_mixed_image_vs_fp64(ctx, name):
    """the mixed solver takes these fields from the FP64 field code and runs
    its FP32 coefficients and stages on them.  Smooth fields: the workload
    bar (I 1e-5 relative, Q/U/V 5e-5 of peak I) against FP64 at 1e-8.  Fields
    with a jump (a sheath's Gamma step, a law switching at z0, a direction
    flipping at x = 0): there the Dormand-Prince error estimate undershoots
    and FP64 at 1e-8 is itself unconverged, so both run tighter (mixed 1e-9,
    FP64 1e-10) and the bar is the larger of 1e-5 and 3x the distance of
    FP64 at 1e-8 from FP64 at 1e-10"""
    j, im, p = case(name)
    jump = name in DISCONTINUOUS
    p64 = models.copy_struct(p)
    if jump:
        p64.stokes_rel_tol = 1e-10
    g64 = ctx.observe(j, im, p64)
    gm = ctx.observe(j, im, models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-9 if jump else 1e-6))
    assert np.array_equal(g64["flags"] & abi.PIX_HIT, gm["flags"] & abi.PIX_HIT)
    rel, quv, nhit = _mixed_errors(g64, gm)
    bar_i, bar_quv = 1e-5, 5e-5
    if jump:
        e64_i, e64_quv, _ = _mixed_errors(g64, ctx.observe(j, im, p))
        bar_i, bar_quv = max(bar_i, 3 * e64_i), max(bar_quv, 3 * e64_quv)
        print("fp64 1e-8 vs 1e-10 %s: max rel I %.3g, QUV/peak %.3g" % (name, e64_i, e64_quv), flush=True)
    print("mixed %s: %d hits, max rel I err %.3g, QUV/peak %.3g" % (name, nhit, rel, quv), flush=True)
    assert rel <= bar_i, (name, rel, bar_i)
    assert quv <= bar_quv, (name, quv, bar_quv)


def test_batch_with_further_classes_bit_exact(ctx):
    """a batch mixing these models with workload models: the former run
    through the plain kernels, the latter through the persistent launch;
    every image equals its single-image FP64 result"""
    items = [(j, im, p) for n, j, im, p in CASES[:4]]
    items.insert(1, models.conical_toroidal_model(n_side=32, pixel_pc=0.05))
    items.append(models.paraboloid_helical_model(n_side=32, pixel_pc=0.1, precision=abi.PREC_FP64))
    outs = ctx.observe_batch(items)
    for (j, im, p), b in zip(items, outs):
        s = ctx.observe(j, im, p)
        for k in ("tau", "I", "Q", "U", "V", "flags", "n_rhs"):
            assert np.array_equal(s[k], b[k]), k


@pytest.mark.parametrize("name", ["ParabaloidCone", "AccParabolicConstConeVField", "SheathCentralVField"])
def test_tau_stripe_bit_exact(ctx, name):
    j, im, p = case(name)
    g = ctx.tau_stripe(j, im, p)
    o = oracle.tau_stripe(j, im, p, math=oracle.MATH_JSX)
    assert np.array_equal(g, o)
    assert (g > 0).any()
