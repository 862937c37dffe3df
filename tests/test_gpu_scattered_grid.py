"""Scattered simulation nodes on the GPU (create_triangulation +
SimulationInterpolater, src/SimulationInterpolater.cpp:24-117): B, N and Gamma
given at random (z, r) nodes, uploaded with jsx_grid_upload_points
(Delaunay-triangulated on the host, looked up through a bucket index in HBM).
Field values at points and FP64 images bit for bit against the oracle on the
same triangulation; mixed precision against FP64; nodes that form a regular
grid take the regular layout."""
import numpy as np
import pytest

from jetshow_amd import abi, grf, models, textio
from oracle import oracle

pytestmark = pytest.mark.gpu

PC = abi.PC


def _nodes(n, seed):
    """random nodes over the cylinder's (z, r) box, its corners and axis ends
    included, with the c0-like profiles of the grid workload (B ~ 0.1 G / d,
    N ~ 500 / d^2, Gamma rising from 2 to 12), d = max(|(z, r)|, 0.1 pc)"""
    rng = np.random.default_rng(seed)
    z = np.concatenate([rng.uniform(0, 20, n), [0, 20, 0, 20]])
    r = np.concatenate([rng.uniform(0, 0.5, n), [0, 0, 0.5, 0.5]])
    d = np.maximum(np.hypot(z, r), 0.1)
    wig = 1.0 + 0.1 * rng.normal(size=z.size)
    B = 0.1 / d * wig
    N = 500.0 / d ** 2 * wig
    G = 2.0 + 10.0 * z / 20.0
    return z, r, (B, N, G)


def _setup(ctx, n=4000, seed=3):
    z, r, vals = _nodes(n, seed)
    # the library takes the nodes in cm and divides by pc: the oracle gets
    # the same coordinates
    z, r = (z * PC) / PC, (r * PC) / PC
    tri = textio.triangulate(z, r)
    handles, tgrids = [], {}
    for v, nv in zip(vals, (0.0, 0.0, 1.0)):
        pts = np.stack([z * PC, r * PC, v], 1)
        h = ctx.upload_points(pts, nan_value=nv)
        handles.append(h)
        tgrids[h] = oracle.TriGrid(z, r, v, tri, nv)
    # the oracle reads handle h from grids[h - 1]
    grids = [tgrids.get(h, oracle.TriGrid(z, r, vals[0], tri, 0.0)) for h in range(1, max(handles) + 1)]
    return handles, grids


def _release(ctx, handles):
    for h in handles:
        ctx.release_grid(h)


def test_point_values_bit_exact(ctx):
    handles, grids = _setup(ctx)
    try:
        j, im, p = models.grid_jet_model(handles, n_side=16, pixel_pc=0.125)
        rng = np.random.default_rng(9)
        m = 3000
        zq = rng.uniform(-0.5, 20.5, m)
        rq = rng.uniform(0.0, 0.6, m)
        ph = rng.uniform(-np.pi, np.pi, m)
        pts = np.stack([rq * np.cos(ph) * PC, rq * np.sin(ph) * PC, zq * PC], 1)
        th = rng.uniform(0.02, 1.2, m)
        nl = np.stack([np.sin(th), np.zeros(m), np.cos(th)], 1)
        g = ctx.fields(j, pts, nl, p.nu)
        o = oracle.point_values(j, p.nu, pts, nl, math=oracle.MATH_JSX, grids=grids)
        bad = ~((g == o) | (np.isnan(g) & np.isnan(o)))
        if bad.any():
            rows, cols = np.nonzero(bad)
            print("mismatch: %d values, columns %s, first point %s: %s vs %s" % (
                bad.sum(), sorted(set(cols.tolist())), pts[rows[0]] / PC, g[rows[0]], o[rows[0]]), flush=True)
        assert not bad.any()
        # some points fall outside the nodes' hull: N there is the nan_value 0
        outside = (zq < 0) | (zq > 20) | (rq > 0.5)
        assert outside.sum() > 100
    finally:
        _release(ctx, handles)


@pytest.mark.parametrize("stokes", [abi.STOKES_I, abi.STOKES_FULL])
def test_fp64_image_bit_exact(ctx, stokes):
    handles, grids = _setup(ctx)
    try:
        j, im, p = models.grid_jet_model(handles, n_side=16, pixel_pc=0.125, stokes=stokes)
        print("scattered grid image, stokes", stokes, flush=True)
        g = ctx.observe(j, im, p)
        o = oracle.observe(j, im, p, None, math=oracle.MATH_JSX, grids=grids, threads=8)
        for k in ("tau", "I", "Q", "U", "V", "l", "flags", "n_rhs", "n_steps"):
            assert np.array_equal(g[k].reshape(-1), o[k]), k
        assert ((o["flags"] & abi.PIX_HIT) != 0).sum() > 50
        assert (o["I"] > 0).sum() > 50
    finally:
        _release(ctx, handles)


def test_mixed_image_vs_fp64(ctx):
    """the piecewise-linear fields have kinks at every triangle edge, which
    the Dormand-Prince error estimate undershoots: FP64 at 1e-10 still lies
    ~5e-6 from the converged image, so the reference point runs at 1e-11 and
    the mixed solver at 2e-10"""
    handles, grids = _setup(ctx)
    try:
        j, im, p = models.grid_jet_model(handles, n_side=16, pixel_pc=0.125)
        p.stokes_rel_tol = 1e-11
        g64 = ctx.observe(j, im, p)
        gm = ctx.observe(j, im, models.with_precision(p, abi.PREC_MIXED, rel_tol=2e-10))
        hit = (g64["flags"] & abi.PIX_HIT) != 0
        assert np.array_equal(hit, (gm["flags"] & abi.PIX_HIT) != 0)
        ok = hit & (g64["I"] > 0)
        rel = np.abs(gm["I"][ok] / g64["I"][ok] - 1.0)
        print("scattered mixed vs fp64: max rel I %.3g over %d pixels" % (rel.max(), ok.sum()), flush=True)
        assert rel.max() <= 1e-5
    finally:
        _release(ctx, handles)


def test_regular_nodes_take_the_regular_layout(ctx):
    """nodes of a complete regular grid, uploaded as points, give the image of
    the grid jsx_grid_from_points lays them out as, uploaded as a grid, bit
    for bit"""
    (B, N, G), geo = grf.jet_grids(96, 24)
    z0, dz, r0, dr = geo
    zz, rr = np.meshgrid(z0 + dz * np.arange(96), r0 + dr * np.arange(24), indexing="ij")
    hp, hg = [], []
    try:
        for v, nv in zip((B, N, G), (0.0, 0.0, 1.0)):
            pts = np.stack([zz.ravel() * PC, rr.ravel() * PC, v.ravel()], 1)
            hp.append(ctx.upload_points(pts, nan_value=nv))
            vals, g2 = textio.grid_from_points(pts, nv)
            hg.append(ctx.upload_grid(vals, *g2, nan_value=nv))
        j1, im, p = models.grid_jet_model(hp, n_side=16, pixel_pc=0.125)
        j2, _, _ = models.grid_jet_model(hg, n_side=16, pixel_pc=0.125)
        a = ctx.observe(j1, im, p)
        b = ctx.observe(j2, im, p)
        for k in ("tau", "I", "flags", "n_rhs"):
            assert np.array_equal(a[k], b[k]), k
        assert ((a["flags"] & abi.PIX_HIT) != 0).sum() > 50
    finally:
        _release(ctx, hp + hg)
