"""Mixed-precision requests on regular-grid fields, integrated cell by cell
(the walk kernel): every pixel of small images of both grid resolutions the
workloads use, Stokes I and full Stokes, within 1e-5 of converged FP64
(rel 1e-11, tau converged too, tau itself within 1e-5); l and flags as the Dormand-Prince lane
gives them (tau through the crossing step differs where tau_max is reached);
the same images through the batched launch and the request server."""
import numpy as np
import pytest

from jetshow_amd import abi, grf, models

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nz,nr", [(512, 128), (2048, 512)])
@pytest.mark.parametrize("stokes", [abi.STOKES_I, abi.STOKES_FULL])
def test_grid_walk_vs_converged_fp64(ctx, nz, nr, stokes):
    (B, N, G), geo = grf.jet_grids(nz, nr)
    hs = [ctx.upload_grid(v, *geo, nan_value=nv) for v, nv in ((B, 0.0), (N, 0.0), (G, 1.0))]
    try:
        j, im, p = models.grid_jet_model(hs, n_side=48, pixel_pc=2.0 / 48, nu_ghz=15.4, precision=abi.PREC_MIXED,
                                         stokes=stokes)
        print("grid walk %dx%d stokes %d" % (nz, nr, stokes), flush=True)
        q64 = models.with_precision(p, abi.PREC_FP64, rel_tol=1e-11)
        q64.stokes_abs_tol = 1e-20
        q64.tau_abs_tol, q64.tau_rel_tol = 1e-20, 1e-11
        g64 = ctx.observe(j, im, q64)
        ctx.set_option("grid_walk", 1)
        gw = ctx.observe(j, im, p)
        ctx.set_option("grid_walk", 0)
        try:
            gl = ctx.observe(j, im, p)
        finally:
            ctx.set_option("grid_walk", 1)
        hit = (g64["flags"] & abi.PIX_HIT) != 0
        assert np.array_equal(hit, (gw["flags"] & abi.PIX_HIT) != 0)
        ok = hit & (g64["I"] > 0)
        rel = np.abs(gw["I"][ok] / g64["I"][ok] - 1.0)
        print("  walk: max rel I %.3g over %d pixels, evaluations %.0f per hit (lane %.0f)" % (
            rel.max(), ok.sum(), gw["n_rhs"][hit].mean(), gl["n_rhs"][hit].mean()), flush=True)
        assert rel.max() <= 1e-5
        assert np.array_equal(gw["flags"], gl["flags"])
        assert np.allclose(gw["l"], gl["l"], rtol=1e-12, atol=0)
        # tau: the Gauss sum over the pieces against FP64 with tau converged
        # (the lanes' own tau runs at the reference's tau tolerance, ~1e-3)
        crossed = (g64["flags"] & abi.PIX_TAU_CROSSED) != 0
        etau = np.abs(gw["tau"][hit & ~crossed] / g64["tau"][hit & ~crossed] - 1.0).max()
        print("  walk: tau within %.3g" % etau, flush=True)
        assert etau <= 1e-5
        assert gw["n_rhs"][hit].sum() < gl["n_rhs"][hit].sum()
        if stokes == abi.STOKES_FULL:
            assert not gw["Q"].any() and not gw["U"].any() and not gw["V"].any()
        # the batched launch routes it to the same kernel
        pb = models.copy_struct(p)
        pb.flags |= 0x100 | 0x200
        gb = ctx.observe_batch([(j, im, pb)])[0]
        assert np.array_equal(gb["I"], gw["I"]) and np.array_equal(gb["flags"], gw["flags"])
    finally:
        for h in hs:
            ctx.release_grid(h)
