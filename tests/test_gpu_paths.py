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
    for name, (j, im, p) in [("paraboloid", models.paraboloid_helical_model(n_side=64, pixel_pc=0.08)),
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
    1e-5 relative of FP64 on every hit pixel, Q/U/V within 5e-5 of peak I.
    The FP64 runs use relative Stokes control (abs 1e-20): with the reference
    absolute tolerances faint pixels are unconverged at the 1e-3 level in FP64
    itself, and any other step sequence lands elsewhere."""
    cases = [("paraboloid_full", models.paraboloid_helical_model(n_side=64, pixel_pc=0.08)),
             ("conical_I", models.conical_toroidal_model(n_side=64, pixel_pc=0.02)),
             ("reference_main", W.reference_main_model())]
    for name, (j, im, p) in cases:
        p.stokes_abs_tol = 1e-20
        g64 = ctx.observe(j, im, p)
        # Stokes I alone needs 1e-7: its error norm lacks the Q/U/V terms
        # that hold full-Stokes steps to ~1e-6 of the result
        pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6 if p.stokes == abi.STOKES_FULL else 1e-7)
        if persistent:
            pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
        gm = ctx.observe(j, im, pm)
        rel, quv, nhit = _mixed_errors(g64, gm)
        print("mixed %s: %d hits, max rel I err %.3g, QUV/peak %.3g" % (name, nhit, rel, quv), flush=True)
        assert rel <= 1e-5, (name, rel)
        assert quv <= 5e-5, (name, quv)


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


def test_split_long_rays_on_gpu(ctx):
    """Mixed full-Stokes batch with parking after 4096 right-hand sides: the
    Faraday-thick apex rays go through the time-parallel pieces; every hit
    pixel's I stays within 1e-5 of FP64 and Q/U/V within 5e-5 of peak I."""
    j, im, p = models.paraboloid_helical_model(n_side=64, pixel_pc=0.08)
    g64 = ctx.observe(j, im, p)
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
    old = ctx.get_option("defer_rhs")
    ctx.set_option("defer_rhs", 4096)
    try:
        gm = ctx.observe(j, im, pm)
        parked, pieces = ctx.get_option("stat_parked"), ctx.get_option("stat_pieces")
    finally:
        ctx.set_option("defer_rhs", old)
    rel, quv, nhit = _mixed_errors(g64, gm)
    print("split: %d parked rays, %d pieces; max rel I %.3g, QUV/peak %.3g" % (parked, pieces, rel, quv), flush=True)
    assert parked > 0 and pieces > parked
    assert rel <= 1e-5
    assert quv <= 5e-5


def test_split_exponential_pieces_on_gpu(ctx):
    """piece_layout 2: the split rays' pieces are integrated by exponential
    (Magnus) propagators; every hit pixel's I stays within 1e-5 of FP64 and
    Q/U/V within 5e-5 of peak I, with fewer coefficient evaluations than the
    Dormand-Prince pieces of the same split take right-hand sides."""
    j, im, p = models.paraboloid_helical_model(n_side=64, pixel_pc=0.08)
    g64 = ctx.observe(j, im, p)
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
    old = ctx.get_option("defer_rhs"), ctx.get_option("piece_layout")
    ctx.set_option("defer_rhs", 4096)
    try:
        rhs = {}
        for layout in (1, 2):
            ctx.set_option("piece_layout", layout)
            gm = ctx.observe(j, im, pm)
            # Dormand-Prince pieces count right-hand sides, exponential
            # pieces coefficient evaluations (two per step)
            rhs[layout] = ctx.get_option("stat_rhs_pieces" if layout == 1 else "stat_rhs_exp")
            parked = ctx.get_option("stat_parked")
    finally:
        ctx.set_option("defer_rhs", old[0])
        ctx.set_option("piece_layout", old[1])
    rel, quv, nhit = _mixed_errors(g64, gm)
    print("exponential pieces: %d parked rays; coefficient evaluations %d (Dormand-Prince rhs %d); max rel I %.3g, QUV/peak %.3g" % (
        parked, rhs[2], rhs[1], rel, quv), flush=True)
    assert parked > 0
    assert 0 < rhs[2] < rhs[1]
    assert rel <= 1e-5
    assert quv <= 5e-5


def test_exponential_single_launch_and_rounds(ctx):
    """The exponential pieces in one launch (mg_part: parts spawned and taken
    while the launch runs) and in the budgeted rounds (mg_part 0): both within
    the bar of FP64 on every hit pixel, with the same flags; the single launch
    takes more pieces (its spawned parts split again) in one loop; a batch of
    two models (pieces of both requests in one launch, wavefronts mixing
    them) gives each model's image within the bar as well."""
    j, im, p = models.paraboloid_helical_model(n_side=64, pixel_pc=0.08)
    g64 = ctx.observe(j, im, p)
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-7)
    pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
    old = ctx.get_option("mg_part")
    out = {}
    try:
        for part in (48, 0):
            ctx.set_option("mg_part", part)
            gm = ctx.observe(j, im, pm)
            out[part] = (gm, ctx.get_option("stat_pieces"), ctx.get_option("stat_mg_trips"))
            rel, quv, nhit = _mixed_errors(g64, gm)
            print("mg_part %d: %d pieces, %d wavefront trips; max rel I %.3g, QUV/peak %.3g" % (
                part, out[part][1], out[part][2], rel, quv), flush=True)
            assert rel <= 1e-5
            assert quv <= 5e-5
        assert np.array_equal(out[48][0]["flags"], out[0][0]["flags"])
        assert out[48][1] > out[0][1] > 0
        # two models in one batched launch, single-launch mode
        ctx.set_option("mg_part", 48)
        j2, im2, p2 = models.paraboloid_helical_model(n_side=48, pixel_pc=0.1)
        g64b = ctx.observe(j2, im2, p2)
        pm2 = models.with_precision(p2, abi.PREC_MIXED, rel_tol=1e-7)
        pm2.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
        ra, rb = ctx.observe_batch([(j, im, pm), (j2, im2, pm2)])
        for ref, got in ((g64, ra), (g64b, rb)):
            rel, quv, nhit = _mixed_errors(ref, got)
            print("batch of two: max rel I %.3g, QUV/peak %.3g over %d hits" % (rel, quv, nhit), flush=True)
            assert rel <= 1e-5
            assert quv <= 5e-5
    finally:
        ctx.set_option("mg_part", old)


def test_exponential_error_estimates_on_gpu(ctx):
    """Both step error estimates of the exponential pieces (mg_err 0: first
    order in the commutator terms; 1: a second, 2nd-order exponential per
    step) hold every hit pixel's I within 1e-5 of FP64 and Q/U/V within 5e-5
    of peak I, at the default tolerance and at a 4x tighter one; the tighter
    tolerance takes more coefficient evaluations."""
    j, im, p = models.paraboloid_helical_model(n_side=64, pixel_pc=0.08)
    g64 = ctx.observe(j, im, p)
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
    names = ("defer_rhs", "piece_layout", "mg_err", "mg_rel")
    old = [ctx.get_option(k) for k in names]
    ctx.set_option("defer_rhs", 4096)
    ctx.set_option("piece_layout", 2)
    try:
        for est in (0, 1):
            ctx.set_option("mg_err", est)
            evals = []
            for tol in (old[3], old[3] / 4):
                ctx.set_option("mg_rel", tol)
                gm = ctx.observe(j, im, pm)
                parked, n_exp = ctx.get_option("stat_parked"), ctx.get_option("stat_rhs_exp")
                rel, quv, nhit = _mixed_errors(g64, gm)
                print("mg_err %d mg_rel %.2g: %d parked, %d coefficient evaluations; max rel I %.3g, QUV/peak %.3g"
                      % (est, tol, parked, n_exp, rel, quv), flush=True)
                assert parked > 0 and n_exp > 0
                assert rel <= 1e-5
                assert quv <= 5e-5
                evals.append(n_exp)
            assert evals[1] > evals[0]
    finally:
        for k, v in zip(names, old):
            ctx.set_option(k, v)


def test_batch_equals_single_images(ctx):
    """jsx_observe_batch over three different models (one launch per
    precision/stokes group) gives each image exactly what the single-image
    persistent launch gives (FP64, and mixed without splitting)."""
    cases = [models.conical_toroidal_model(n_side=64, pixel_pc=0.02),
             models.paraboloid_helical_model(n_side=32, pixel_pc=0.16),
             W.reference_main_model()]
    old = ctx.get_option("defer_rhs")
    ctx.set_option("defer_rhs", 0)
    try:
        for prec in (abi.PREC_FP64, abi.PREC_MIXED):
            items = []
            for j, im, p in cases:
                q = models.with_precision(p, prec, rel_tol=1e-6) if prec == abi.PREC_MIXED else models.copy_struct(p)
                q.stokes = abi.STOKES_FULL
                q.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
                items.append((j, im, q))
            outs = ctx.observe_batch(items)
            for (j, im, q), b in zip(items, outs):
                s = ctx.observe(j, im, q)
                for k in ("tau", "I", "Q", "U", "V", "l", "flags", "n_rhs", "n_steps"):
                    assert np.array_equal(s[k], b[k]), (prec, k)
    finally:
        ctx.set_option("defer_rhs", old)


def test_triage_list_same_planes(ctx):
    """The triage pass (misses written first, live rays queued) changes only
    which lane takes which ray: FP64 and unsplit mixed planes are identical
    with and without it, over a batch of images whose hit fractions differ."""
    cases = [models.conical_toroidal_model(n_side=64, pixel_pc=0.02),
             models.paraboloid_helical_model(n_side=48, pixel_pc=0.1),
             W.reference_main_model()]
    old = ctx.get_option("defer_rhs"), ctx.get_option("triage")
    ctx.set_option("defer_rhs", 0)
    try:
        for prec in (abi.PREC_FP64, abi.PREC_MIXED):
            items = []
            for j, im, p in cases:
                q = models.with_precision(p, prec, rel_tol=1e-6) if prec == abi.PREC_MIXED else models.copy_struct(p)
                q.stokes = abi.STOKES_FULL
                q.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
                items.append((j, im, q))
            res = []
            for tri in (0, 1):
                ctx.set_option("triage", tri)
                res.append(ctx.observe_batch(items))
            for a, b in zip(*res):
                for k in ("tau", "I", "Q", "U", "V", "l", "flags", "n_rhs", "n_steps"):
                    assert np.array_equal(a[k], b[k]), (prec, k)
    finally:
        ctx.set_option("defer_rhs", old[0])
        ctx.set_option("triage", old[1])


def test_server_same_alone_and_in_mix(ctx):
    """Continuous batching: a request gives the same planes alone and when
    submitted among others of different sizes and frequencies -- its rays run
    in a batched launch with the other requests', and a long full-Stokes ray
    may be split into pieces placed by the launch's other long rays, so the
    two agree to the solver's tolerance (I within 1e-5 relative, Q/U/V within
    5e-5 of peak I, identical flags for the hit mask), and both agree with the
    FP64 image to the same bar; every request of the mix completes and reports
    its latency."""
    from jetshow_amd.host import Server
    j0, im0, p0 = models.paraboloid_helical_model(n_side=32, pixel_pc=0.16)
    p0 = models.with_precision(p0, abi.PREC_MIXED, rel_tol=1e-6)
    others = []
    for k, (n, px, nu) in enumerate([(48, 0.1, 8.1), (16, 0.3, 15.4), (64, 0.05, 22.0), (24, 0.2, 5.0)]):
        j, im, p = models.paraboloid_helical_model(n_side=n, pixel_pc=px, nu_ghz=nu)
        others.append((j, im, models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)))
    with Server(ctx, abi.PREC_MIXED, abi.STOKES_FULL, max_active=3) as srv:
        alone, _ = srv.wait(srv.submit(j0, im0, p0))
        ids = [srv.submit(*others[0]), srv.submit(*others[1])]
        mid = srv.submit(j0, im0, p0)
        ids += [srv.submit(*o) for o in others[2:]]
        mixed, t = srv.wait(mid)
        lat = [srv.wait(i)[1]["latency_s"] for i in ids]
        st = srv.stats()
    print("server: %d launches, latencies %s s" % (st["epochs"], ["%.3f" % x for x in lat]), flush=True)
    rel, quv, _ = _mixed_errors(alone, mixed)
    assert rel <= 1e-5 and quv <= 5e-5
    ref = ctx.observe(j0, im0, models.copy_struct(models.with_precision(p0, abi.PREC_FP64, rel_tol=1e-8)))
    for got in (alone, mixed):
        rel, quv, _ = _mixed_errors(ref, got)
        assert rel <= 1e-5 and quv <= 5e-5
    assert st["epochs"] >= 2
    assert all(x > 0 for x in lat)


def test_split_piece_layouts_agree(ctx):
    """The two Stokes-state layouts of a piece (a 5-lane group, or one lane
    with 20 components) evaluate the same coefficients and the same error norm,
    so they take nearly the same steps; where rounding tips one step's
    accept/reject the sequences part, each within the tolerance (1e-6) of the
    solution: the split rays agree to twice that, of peak I."""
    j, im, p = models.paraboloid_helical_model(n_side=64, pixel_pc=0.08)
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
    old = ctx.get_option("defer_rhs"), ctx.get_option("piece_layout")
    ctx.set_option("defer_rhs", 4096)
    try:
        out = []
        for layout in (0, 1):
            ctx.set_option("piece_layout", layout)
            out.append(ctx.observe(j, im, pm))
            assert ctx.get_option("stat_parked") > 0
    finally:
        ctx.set_option("defer_rhs", old[0])
        ctx.set_option("piece_layout", old[1])
    a, b = out
    peak = a["I"].max()
    diff = max(np.abs(a[k] - b[k]).max() for k in ("I", "Q", "U", "V")) / peak
    same = all(np.array_equal(a[k], b[k]) for k in ("I", "Q", "U", "V"))
    print("piece layouts: max diff / peak I %.3g, bitwise identical %s" % (diff, same), flush=True)
    assert np.array_equal(a["flags"], b["flags"])
    assert diff <= 2e-6
