"""The mixed Stokes pass's optical-depth horizon and its error-norm floor on the
CPU (the kernels' lane code compiled for the host): optically thick grid rays
start their Stokes pass where everything deeper is attenuated by e^-50, take a
fraction of the whole path's derivative evaluations, keep the tau and l of the
whole path, and stay within the bar of a converged FP64 run; a ray whose
polarized mode grows along the path (the reference's coefficients let |eta|
exceed k_I near a field along the line of sight) keeps its whole path."""
import numpy as np

from jetshow_amd import abi, grf, models
from oracle import oracle
from tests import lanehost as LH


def _grid_model():
    (B, N, G), geo = grf.jet_grids(512, 128)
    table = []
    for v, nv in zip((B, N, G), (0.0, 0.0, 1.0)):
        d = abi.GridDesc()
        d.nz, d.nr = v.shape
        d.z0, d.dz, d.r0, d.dr = geo
        d.nan_value = nv
        table.append((d, v))
    j, im, p = models.grid_jet_model((1, 2, 3), n_side=32, pixel_pc=2.0 / 32, nu_ghz=8.1,
                                     precision=abi.PREC_MIXED, stokes=abi.STOKES_I)
    return j, im, p, table


def test_thick_grid_rays_start_at_the_horizon():
    j, im, p, table = _grid_model()
    pix = np.array([513, 449, 482, 544, 256, 300], dtype=np.int64)   # tau 1e4, 1e4, 7e3, 1e4, 444; 0.04
    q64 = models.with_precision(p, abi.PREC_FP64, rel_tol=1e-11)
    o = oracle.observe(j, im, q64, pix, math=oracle.MATH_LIBM, grids=table, threads=8)
    # the Dormand-Prince lane (grid models go cell by cell unless told not to)
    a = LH.observe(j, im, p, pix, grids=table, grid_walk=False)
    full = models.copy_struct(p)
    full.flags |= abi.OPT_FULL_DEPTH
    b = LH.observe(j, im, full, pix, grids=table, grid_walk=False)
    thick = o["tau"] > 50.0
    assert thick.sum() == 5
    assert ((a["flags"][thick] & abi.PIX_HORIZON) != 0).all()
    assert ((a["flags"][~thick] & abi.PIX_HORIZON) == 0).all()
    assert ((b["flags"] & abi.PIX_HORIZON) == 0).all()
    assert np.array_equal(a["flags"] & np.uint32(0xffffffff ^ abi.PIX_HORIZON), b["flags"])
    assert np.array_equal(a["tau"], b["tau"]) and np.array_equal(a["l"], b["l"])
    # the rays tau ~ 1e4 deep: a third of the whole path's evaluations or less
    deep = o["tau"] > 5e3
    assert (a["n_rhs"][deep] < 0.35 * b["n_rhs"][deep]).all()
    # against converged FP64: the horizon's pass within the bar; the whole
    # path's within twice it (its deep stiff stretch adds the error of many
    # steps at the kinks of the interpolated fields: 1.2e-5 on the tau = 444 ray)
    ea = np.abs(a["I"] / o["I"] - 1.0).max()
    eb = np.abs(b["I"] / o["I"] - 1.0).max()
    assert ea <= 1e-5
    assert eb <= 2e-5


def test_growing_polarized_mode_keeps_the_whole_path():
    """sweep model with s = 2.02 at 16.9 GHz, line of sight 0.026 rad inside
    the cone: at this pixel the dichroism |eta| outgrows k_I along the path,
    the decay bound never reaches 50, and the pass runs from tau_max"""
    sw = models.abc_sweep_models(n_models=8, sides=(64,))
    j, im, p, k = [x for x in sw if x[3] == 1 and x[2].nu > 1e10][0]
    j.s = 2.02
    pix = np.array([3778], dtype=np.int64)
    q64 = models.with_precision(p, abi.PREC_FP64, rel_tol=1e-8)
    q64.stokes_abs_tol = 1e-20
    o = oracle.observe(j, im, q64, pix, math=oracle.MATH_JSX)
    a = LH.observe(j, im, p, pix)
    assert o["tau"][0] > 1e3
    assert (a["flags"][0] & abi.PIX_HORIZON) == 0
    assert abs(a["I"][0] / o["I"][0] - 1.0) <= 1e-5


def test_grid_rays_cell_by_cell():
    """the cell-by-cell walk of grid fields: thin and thick rays of both grid
    resolutions within 1e-5 of converged FP64 (rel 1e-11), tau, l and flags
    as the lanes give them, a fraction of the lane's evaluations"""
    j, im, p, table = _grid_model()
    pix = np.array([513, 449, 482, 544, 256, 300, 600, 280, 405], dtype=np.int64)
    q64 = models.with_precision(p, abi.PREC_FP64, rel_tol=1e-11)
    q64.tau_abs_tol, q64.tau_rel_tol = 1e-20, 1e-11      # tau itself converged too
    o = oracle.observe(j, im, q64, pix, math=oracle.MATH_LIBM, grids=table, threads=8)
    w = LH.observe(j, im, p, pix, grids=table)
    lane = LH.observe(j, im, p, pix, grids=table, grid_walk=False)
    assert np.abs(w["I"] / o["I"] - 1.0).max() <= 1e-5
    crossed = (o["flags"] & abi.PIX_TAU_CROSSED) != 0
    assert crossed.sum() == 3 and (w["tau"][crossed] >= p.tau_max).all()
    assert np.abs(w["tau"][~crossed] / o["tau"][~crossed] - 1.0).max() <= 1e-5
    assert np.allclose(w["l"], lane["l"], rtol=1e-12, atol=0)
    assert np.array_equal(w["flags"], lane["flags"])
    assert w["n_rhs"].sum() < 0.5 * lane["n_rhs"].sum()
    # full Stokes: Q, U, V zero (random-direction B), I as for Stokes I
    pf = models.copy_struct(p)
    pf.stokes = abi.STOKES_FULL
    wf = LH.observe(j, im, pf, pix, grids=table)
    assert np.array_equal(wf["I"], w["I"])
    assert not wf["Q"].any() and not wf["U"].any() and not wf["V"].any()
