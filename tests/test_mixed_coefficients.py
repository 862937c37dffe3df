"""The mixed-precision coefficient code (FP32 with FP64 exponents, the
kernels' own m_coefficients compiled for the CPU) against the oracle's FP64
coefficients at points inside the jets of the paraboloid workload and of
sweep models -- including s close to 2, where k_C_r's 2/(s - 2) factor
multiplies a difference of two numbers near 1 (src/utils.cpp:93-98), and a
line of sight inside the cone (LOS < half-angle).

Error scales: absorption-type terms relative to k_I, emission-type relative
to eta_I, k_F relative to the value it would have with the field along the
line of sight (|k_F| / |cos theta|: FP32 rounding of cos theta is absolute),
the conversion pair (k_C, h_Q) relative to its own magnitude."""
import math

import numpy as np
import pytest

from tests import lanehost
from jetshow_amd import abi, models
from oracle import oracle

PC = abi.PC


def _worst(j, im, p, pts):
    cm = lanehost.coefficients_mixed(j, im, p, pts)
    d = np.array([-math.sin(im.los_angle), 0, -math.cos(im.los_angle)])
    co = np.array([oracle.coefficients(j, p.nu, q, -d, math=oracle.MATH_LIBM) for q in pts])
    fl = np.array([oracle.fields(j, p.nu, q, -d, math=oracle.MATH_LIBM) for q in pts])
    b, nl = fl[:, 4:7], fl[:, 9:12]
    cos_t = np.abs((b * nl).sum(1)) / np.linalg.norm(b, axis=1)
    ks, es = np.abs(co[:, 0]), np.abs(co[:, 7])
    err = np.abs(cm - co)
    worst = np.zeros(11)
    for i in range(11):
        sc = es if i >= 7 else ks
        if i == 4:
            sc = np.maximum(ks, np.abs(co[:, 4]) / np.maximum(cos_t, 1e-300))
        if i in (5, 6):
            sc = np.maximum(ks, np.hypot(co[:, 5], co[:, 6]))
        worst[i] = (err[:, i] / sc).max()
    return worst


def test_mixed_coefficients_paraboloid():
    rng = np.random.default_rng(3)
    j, im, p = models.paraboloid_helical_model(precision=abi.PREC_MIXED)
    pts = np.stack([rng.uniform(-.2, .2, 300), rng.uniform(-.2, .2, 300), rng.uniform(.2, 5, 300)], 1) * PC
    w = _worst(j, im, p, pts)
    assert w.max() < 5e-6, w


@pytest.mark.parametrize("s", [2.0005, 2.005, 2.05, None])
def test_mixed_coefficients_sweep_models_s_near_2(s):
    """k_C stays within FP32 rounding as s -> 2 (2e-4 with the difference
    formed in FP32; the kernels form it through expm1 of an FP64 exponent)."""
    rng = np.random.default_rng(4)
    sw = models.abc_sweep_models(n_models=6)
    seen_inside = False
    for (j, im, p, k) in sw[::2]:
        if s is not None:
            j.s = s
        if k == 2:                      # a line of sight inside the cone
            im.los_angle = 0.5 * j.geo_param
            seen_inside = True
        a = j.geo_param
        z = rng.uniform(0.05, 10, 200)
        r = rng.uniform(0, 1, 200) * math.tan(a) * z
        ph = rng.uniform(0, 2 * math.pi, 200)
        pts = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1) * PC
        w = _worst(j, im, p, pts)
        assert w.max() < 5e-6, (k, j.s, w)
    assert seen_inside
