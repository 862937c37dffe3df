"""JetModelZoom (jetmodel.py:11-112) on the GPU solver: pixel grid and
descriptors on the CPU; single and batched runs on the GPU."""
import math

import numpy as np
import pytest

from jetshow_amd import abi, models, zoom
from oracle import oracle


def _zm(theta=0.663225, **kw):
    z = zoom.JetModelZoom(15.4, 0.0165, 50, 10, -3, -1, central_vfield=False, **kw)
    z.set_params_vec(np.array([0.0, 0.0, 0.0, theta, 0.068, math.log(0.1), math.log(1e4), math.log(2.0),
                               math.log(1.0), math.log(2.0)]))
    return z


def test_grid_and_descriptors():
    z = _zm()
    assert z.pixsize_array.shape == (50, 10)
    assert np.allclose(z.pixsize_array[:, 0], np.logspace(-3, -1, 50))
    # pixel centres: r_ob increasing along the jet, d antisymmetric across it
    assert z.r_ob.shape == (50, 10) and z.d.shape == (50, 10)
    assert (np.diff(z.r_ob[:, 0]) > 0).all()
    assert np.allclose(z.d[:, :5], -z.d[:, 5:][:, ::-1])
    j, im, p = z.descriptors()
    j2, im2, p2 = models.get_i_image_model(0.663225, 0.0165, 50, 10, 1e-3, 1e-1, 0.068, z.B1, z.m, z.K1, z.n,
                                           z.gamma, 15.4, 1e5, False)
    assert abs(z.B1 - 0.1) < 1e-15 and abs(z.K1 - 1e4) < 1e-10
    assert bytes(j) == bytes(j2) and bytes(im) == bytes(im2) and bytes(p) == bytes(p2)


@pytest.mark.gpu
def test_run_and_run_many(ctx):
    zs = [_zm(theta=t, ctx=ctx) for t in (0.3, 0.5, 0.663225)]
    single = []
    for z in zs:
        z.run()
        single.append((z.image().copy(), z.image_tau().copy()))
        z.set_params_vec(np.array([0.0, 0.0, 0.0, z.theta, z.phi, math.log(z.B1), math.log(z.K1),
                                   math.log(z.gamma), math.log(z.m), math.log(z.n)]))
    assert zoom.run_many(zs, ctx) == 3
    for z, (i1, t1) in zip(zs, single):
        assert np.array_equal(z.image(), i1)
        assert np.array_equal(z.image_tau(), t1)
    # against the oracle in Jy (the pyjetshow scaling)
    z = zs[2]
    j, im, p = z.descriptors()
    o = oracle.observe(j, im, p, None, math=oracle.MATH_JSX)
    assert np.array_equal(z.image_tau().reshape(-1), o["tau"])
    assert (z.image() > 0).any()
