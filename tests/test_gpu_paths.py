"""GPU checks of the other paths: Geometry::hit on the GPU, the persistent
work-queue kernel against the plain kernel, the mixed-precision solver against
FP64, and the pyjetshow / C++ class binding against the oracle."""
import math
import os
import sys

import numpy as np
import pytest

from jetshow_amd import abi, models
from oracle import oracle
from tests import workloads as W
from tests.test_oracle_geometry import check_known_answers, cone, paraboloid

pytestmark = pytest.mark.gpu

OPT_PERSISTENT = 0x100
OPT_APEX_FIRST = 0x200


def test_cone_known_answers_on_gpu(ctx):
    check_known_answers(lambda j, r: ctx.hit(j, r))


@pytest.mark.parametrize("geom", ["cone", "paraboloid", "cylinder", "cylinder_bug"])
def test_hit_matches_oracle(ctx, geom):
    rng = np.random.default_rng(4)
    n = 4000
    los = rng.uniform(0.01, 1.5, n)
    o = np.stack([rng.uniform(-0.2, 0.2, n), rng.uniform(-1, 1, n), rng.uniform(-2, 20, n)], 1) * abi.PC
    d = np.stack([-np.sin(los), rng.uniform(-0.1, 0.1, n), -np.cos(los)], 1)
    rays = np.concatenate([o, d], 1)
    rays[:8] = [[0, 0.1 * abi.PC, 1 * abi.PC, 0, 0, 1]] * 8       # parallel to the axis
    p = abi.default_params(1e10)
    if geom == "cone":
        j = cone(0.05, 100 * abi.PC)
        j.z_clip_min, j.z_clip_max = 0.0, 10 * abi.PC
    elif geom == "paraboloid":
        j = paraboloid()
    else:
        j = cone()
        j.geometry = abi.GEOM_CYLINDER
        j.geo_param = 0.3 * abi.PC
        if geom == "cylinder_bug":
            p.flags = abi.OPT_CYLINDER_REF_BUG
    ns_g, seg_g = ctx.hit(j, rays, p)
    ns_o, seg_o = oracle.hit(j, rays, p)
    assert np.array_equal(ns_g, ns_o)
    assert np.array_equal(seg_g, seg_o)
    assert ns_g.sum() > 100


@pytest.mark.parametrize("order", [0, 1])
def test_persistent_equals_plain(ctx, order):
    for name, (j, im, p) in [("paraboloid", models.paraboloid_helical_model(n_side=128, pixel_pc=0.04)),
                             ("reference_main", W.reference_main_model())]:
        g0 = ctx.observe(j, im, p)
        q = models.copy_struct(p)
        q.flags |= OPT_PERSISTENT | (OPT_APEX_FIRST if order else 0)
        g1 = ctx.observe(j, im, q)
        for k in ("tau", "I", "Q", "U", "V", "l", "flags", "n_rhs", "n_steps"):
            assert np.array_equal(g0[k], g1[k]), (name, k)


def _mixed_errors(g64, gm):
    hit = (g64["flags"] & abi.PIX_HIT) != 0
    assert np.array_equal(hit, (gm["flags"] & abi.PIX_HIT) != 0), "hit masks differ"
    I64, Im = g64["I"][hit], gm["I"][hit]
    pos = I64 > 0
    rel = np.abs(Im[pos] - I64[pos]) / I64[pos]
    peak = I64.max()
    quv = max(np.abs(gm[k][hit] - g64[k][hit]).max() for k in ("Q", "U", "V")) / peak
    return rel.max(initial=0.0), quv, hit.sum()


@pytest.mark.parametrize("persistent", [False, True])
def test_mixed_matches_fp64(ctx, persistent):
    """FP32 coefficients/stages with FP64 roots and step control: I within
    1e-5 relative of FP64 on every hit pixel, Q/U/V within 1e-5 of peak I."""
    cases = [("paraboloid_full", models.paraboloid_helical_model(n_side=128, pixel_pc=0.04)),
             ("conical_I", models.conical_toroidal_model(n_side=64, pixel_pc=0.02)),
             ("reference_main", W.reference_main_model())]
    for name, (j, im, p) in cases:
        g64 = ctx.observe(j, im, p)
        pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
        if persistent:
            pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
        gm = ctx.observe(j, im, pm)
        rel, quv, nhit = _mixed_errors(g64, gm)
        print("mixed %s: %d hits, max rel I err %.3g, QUV/peak %.3g" % (name, nhit, rel, quv), flush=True)
        assert rel <= 1e-5, (name, rel)
        assert quv <= 1e-5, (name, quv)


def test_pyjetshow_get_i_image(ctx):
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jetshow_amd"))
    import pyjetshow
    args = (0.663225, 0.0165, 100, 20, 0.001, 0.1, 0.06832789, 0.1, 1.0, 1e4, 2.0, 5.0, 15.35, 1e5, False)
    tau, I = pyjetshow.get_i_image(*args)
    tau, I = np.array(tau), np.array(I)
    j, im, p = models.get_i_image_model(*args, mas_to_pc=oracle.mas_to_pc)
    o = oracle.observe(j, im, p, None, math=oracle.MATH_JSX)
    assert np.array_equal(tau.reshape(-1), o["tau"])
    sizes, _, _ = oracle.image(im)
    pc_in_mas = oracle.mas_to_pc(0.0165)
    omega = (sizes / (pc_in_mas * abi.PC)) ** 2 * abi.MAS_TO_RAD ** 2
    I_jy = o["I"] / (1e-23 * (1.0165 ** 3) / omega)
    assert np.allclose(I.reshape(-1), I_jy, rtol=1e-14, atol=0)
    assert (o["flags"] & abi.PIX_TAU_CROSSED).any()
