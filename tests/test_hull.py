"""Convex-hull geometry (SimulationGeometry, src/Geometry.cpp:92-132, built as
in src/main.cpp:873-897).

CPU: the oracle's plane-clipping hit against hand-computed entries and exits
of a cube, misses and parallel rays; the hull planes of a cube's corners; the
hull of the reference's 36-point rings around a cylinder against the
analytic cylinder hit (within the polygon's sagitta) and against brute-force
sampling of the line.  GPU: jsx_hit_host and a full FP64 image through the
hull bit for bit against the oracle; the mixed-precision solver through the
hull against FP64."""
import math

import numpy as np
import pytest

from jetshow_amd import abi, hull, models
from oracle import oracle
from tests import workloads as W

PC = 3.0856775814671913e+18


def _table(planes):
    d = abi.GridDesc()
    d.nz, d.nr = planes.shape[0], 4
    d.z0, d.dz, d.r0, d.dr, d.nan_value = 0.0, 1.0, 0.0, 1.0, 0.0
    return (d, np.ascontiguousarray(planes, dtype=np.float64))


def _hull_jet(handle):
    j, im, p = models.conical_toroidal_model(n_side=8, pixel_pc=0.1)
    j.geometry = abi.GEOM_CONVEX_HULL
    j.geo_table = handle
    j.z_clip_min = j.z_clip_max = 0.0
    return j, im, p


CUBE = np.array([[1, 0, 0, PC], [-1, 0, 0, PC], [0, 1, 0, PC], [0, -1, 0, PC], [0, 0, 1, PC], [0, 0, -1, PC]],
                dtype=np.float64)


def test_oracle_cube_known_answers():
    j, _, _ = _hull_jet(1)
    s3 = 1.0 / math.sqrt(3.0)
    rays = np.array([
        [5 * PC, 0.2 * PC, 0.3 * PC, -1, 0, 0],      # along -x: x from 1 to -1 pc
        [0.5 * PC, 0.5 * PC, -7 * PC, 0, 0, 1],      # along +z: z from -1 to 1 pc
        [3 * PC, 3 * PC, 3 * PC, -s3, -s3, -s3],    # diagonal: corner to corner
        [5 * PC, 2 * PC, 0.0, -1, 0, 0],             # y = 2 pc: parallel to y = 1 outside
        [5 * PC, 0.0, 0.0, 0, 1, 0],                 # along y at x = 5 pc: misses
    ])
    ns, seg = oracle.hit(j, rays, grids=[_table(CUBE)])
    assert ns.tolist() == [1, 1, 1, 0, 0]
    assert np.allclose(seg[0, 0], [PC, 0.2 * PC, 0.3 * PC, -PC, 0.2 * PC, 0.3 * PC], rtol=0, atol=1e-3 * PC * 1e-12)
    assert np.allclose(seg[1, 0], [0.5 * PC, 0.5 * PC, -PC, 0.5 * PC, 0.5 * PC, PC], rtol=1e-15, atol=1.0)
    assert np.allclose(seg[2, 0], [PC, PC, PC, -PC, -PC, -PC], rtol=1e-14, atol=1e-14 * PC)


def test_planes_of_cube_corners():
    c = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], dtype=np.float64)
    pl = hull.planes_from_points(c)
    assert pl.shape == (6, 4)
    key = sorted(tuple(np.round(np.r_[r[:3], r[3] / PC], 12)) for r in pl)
    assert key == sorted(tuple(float(v) for v in r[:3]) + (1.0,) for r in CUBE)


def _cylinder_hull(radius_pc=0.5, z1_pc=20.0, nz=41, n_circle=36, z0_pc=0.0):
    z = np.linspace(z0_pc, z1_pc, nz) * PC
    r = np.full(nz, radius_pc * PC)
    return hull.simulation_hull(z, r, n_circle)


def _image_rays(n=24, los=0.1, extent_pc=2.0):
    """Rays of an image plane (direction (-sin los, 0, -cos los)) through a grid
    of across/along offsets."""
    d = np.array([-math.sin(los), 0.0, -math.cos(los)])
    out = []
    for a in np.linspace(-0.6, 0.6, n):
        for b in np.linspace(0.1, extent_pc, n):
            o = np.array([0.0, a * PC, b * PC / math.sin(los)])
            out.append(np.r_[o, d])
    return np.array(out)


def test_ring_hull_against_cylinder_and_sampling():
    planes = _cylinder_hull()
    assert planes.shape[0] == 36 + 2    # 36 side faces and two caps, coplanar triangles merged
    j, _, _ = _hull_jet(1)
    rays = _image_rays()
    ns, seg = oracle.hit(j, rays, grids=[_table(planes)])
    # analytic cylinder r = 0.5 pc clipped to 0 <= z <= 20 pc
    jc, _, _ = _hull_jet(1)
    jc.geometry = abi.GEOM_CYLINDER
    jc.geo_origin[:] = (0.0, 0.0, 0.0)
    jc.geo_axis[:] = (0.0, 0.0, 1.0)
    jc.geo_param = 0.5 * PC
    jc.z_clip_min, jc.z_clip_max = 0.0, 20.0 * PC
    nc, segc = oracle.hit(jc, rays)
    sag = 0.5 * PC * (1 - math.cos(math.pi / 36))
    both = (ns == 1) & (nc == 1)
    # every hull hit lies inside the circumscribed cylinder, every cylinder hit
    # well inside the polygon is a hull hit
    assert both.sum() > 100
    dpin = np.linalg.norm(seg[both, 0, :3] - segc[both, 0, :3], axis=1)
    dpout = np.linalg.norm(seg[both, 0, 3:] - segc[both, 0, 3:], axis=1)
    los = 0.1
    # a radial offset of the wall by up to the sagitta moves the crossing along
    # a ray inclined at los to the axis by up to sag / sin(angle to the wall)
    assert np.median(dpin) < 20 * sag / math.sin(los) and np.median(dpout) < 20 * sag / math.sin(los)
    # brute force: points sampled on the line are inside all planes exactly
    # between the reported entry and exit
    for q in np.nonzero(ns == 1)[0][::37]:
        o, d = rays[q, :3], rays[q, 3:]
        t_in = np.dot(seg[q, 0, :3] - o, d)
        t_out = np.dot(seg[q, 0, 3:] - o, d)
        ts = np.linspace(t_in - 0.05 * (t_out - t_in), t_out + 0.05 * (t_out - t_in), 2001)
        P = o[None, :] + ts[:, None] * d[None, :]
        inside = (P @ planes[:, :3].T <= planes[:, 3][None, :] * (1 + 1e-12)).all(axis=1)
        expect = (ts >= t_in) & (ts <= t_out)
        edge = np.abs(ts - t_in) < 2e-3 * (t_out - t_in)
        edge |= np.abs(ts - t_out) < 2e-3 * (t_out - t_in)
        assert (inside == expect)[~edge].all()


@pytest.mark.gpu
def test_hull_hit_and_image_on_gpu(ctx):
    planes = _cylinder_hull(nz=21)
    h = ctx.upload_hull(planes)
    print("hull handle", h, flush=True)
    try:
        j, _, _ = _hull_jet(h)
        rays = _image_rays(16)
        table = [_table(planes) if i == h - 1 else _table(CUBE) for i in range(h)]
        ns_g, seg_g = ctx.hit(j, rays)
        ns_o, seg_o = oracle.hit(j, rays, grids=table)
        assert np.array_equal(ns_g, ns_o) and np.array_equal(seg_g, seg_o)
        assert (ns_g == 1).sum() > 50
        # a full FP64 image on grid fields inside the hull, bit for bit
        (B, N, G), geo = W.grf_grids(nz=48, nr=12)
        descs = W.grid_descs(B.shape, geo, (0.0, 0.0, 1.0))
        gh = [ctx.upload_grid(v, *geo, nan_value=d.nan_value) for v, d in zip((B, N, G), descs)]
        try:
            jg, im, p = W.grid_model(gh + [gh[0]], n_side=32, pixel_pc=0.06)
            jg.geometry = abi.GEOM_CONVEX_HULL
            jg.geo_table = h
            jg.z_clip_min = jg.z_clip_max = 0.0
            g = {k: v.reshape(-1) for k, v in ctx.observe(jg, im, p).items()}
            full = [None] * max(gh + [h])
            for hh, d, v in zip(gh, descs, (B, N, G)):
                full[hh - 1] = (d, v)
            full[h - 1] = _table(planes)
            full = [t if t is not None else _table(CUBE) for t in full]
            o = oracle.observe(jg, im, p, None, math=oracle.MATH_JSX, grids=full)
            for k in ("tau", "I", "l", "flags", "n_rhs"):
                assert np.array_equal(g[k], o[k]), k
            assert ((g["flags"] & abi.PIX_HIT) != 0).sum() > 50
        finally:
            for hh in gh:
                ctx.release_grid(hh)
    finally:
        ctx.release_grid(h)


@pytest.mark.gpu
def test_hull_mixed_precision_image(ctx):
    """The mixed-precision solver through the hull: identical hit mask, and
    Stokes I within 1e-5 of a converged FP64 image (relative tolerance 1e-11)
    and of the FP64 image at the same tolerance as the mixed run (1e-8).
    These rays cross up to ~20 pc of optically thin cylinder (tau ~ 0.003),
    where the Dormand-Prince local errors, each bounded relative to the
    accumulated intensity, add up along the path (FP64 at 1e-8 is ~5e-6 from
    the converged image); the mixed solver holds the state to the caller's
    tolerance and only the increment's share of the error norm to its FP32
    floor, so at the same tolerance it lands as close as FP64 does.  Both runs
    use relative Stokes control (abs 1e-20)."""
    # from z = 1 pc on: the conical fields (B ~ 1/r, N ~ 1/r^2) are singular at
    # the origin, which a cylinder from z = 0 would contain
    planes = _cylinder_hull(nz=20, z0_pc=1.0)
    h = ctx.upload_hull(planes)
    print("hull mixed", flush=True)
    try:
        j, im, p = models.conical_toroidal_model(n_side=48, pixel_pc=0.05)
        j.geometry = abi.GEOM_CONVEX_HULL
        j.geo_table = h
        j.z_clip_min = j.z_clip_max = 0.0
        p.stokes_abs_tol = 1e-20
        g64 = ctx.observe(j, im, p)
        gconv = ctx.observe(j, im, models.with_precision(p, abi.PREC_FP64, rel_tol=1e-11))
        gm = ctx.observe(j, im, models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-8))
        hit = (g64["flags"] & abi.PIX_HIT) != 0
        assert hit.sum() > 100
        assert np.array_equal(hit, (gm["flags"] & abi.PIX_HIT) != 0)
        pos = hit & (gconv["I"] > 0)
        rel_conv = np.abs(gm["I"][pos] - gconv["I"][pos]) / gconv["I"][pos]
        rel_64 = np.abs(gm["I"][pos] - g64["I"][pos]) / g64["I"][pos]
        fp64_own = np.abs(g64["I"][pos] - gconv["I"][pos]) / gconv["I"][pos]
        print("hull mixed: %d hits, max rel I err %.3g vs converged, %.3g vs FP64 1e-8 (FP64 1e-8 itself %.3g)"
              % (hit.sum(), rel_conv.max(), rel_64.max(), fp64_own.max()), flush=True)
        assert rel_conv.max() <= 1e-5, rel_conv.max()
        assert rel_64.max() <= 1e-5, rel_64.max()
    finally:
        ctx.release_grid(h)
