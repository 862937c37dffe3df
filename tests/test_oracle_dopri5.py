"""Pins for the oracle's Dormand-Prince 5(4) restatement (Boost.odeint is not
vendored in the reference).  Each check fails for a plausible transcription
error of the tableau, the error weights, the dense output or the controller."""
import math

import numpy as np

from oracle import oracle


def test_fifth_order_quadrature_is_exact():
    # y' = 5 t^4: the 5th-order weights integrate degree-4 polynomials exactly
    for dt in (0.1, 0.7, 2.0):
        out, nrhs = oracle.dopri5_check(1, dt=dt)
        assert abs(out[0] - dt ** 5) <= 4e-15 * dt ** 5
        # the embedded 4th-order solution is not exact for degree 4: nonzero estimate
        assert abs(out[1]) > 1e-6 * dt ** 5


def test_embedded_error_vanishes_for_cubics():
    # y' = 4 t^3: both solutions exact, error estimate zero up to rounding
    for dt in (0.1, 1.0, 3.0):
        out, _ = oracle.dopri5_check(3, dt=dt)
        assert abs(out[0] - dt ** 4) <= 1e-15 * dt ** 4
        assert abs(out[1]) <= 1e-14 * dt ** 4


def test_controlled_decay_accuracy_and_counts():
    for tol in (1e-6, 1e-9, 1e-12):
        out, nrhs = oracle.dopri5_check(0, t_end=2.0, dt=0.01, abs_tol=tol, rel_tol=tol)
        assert abs(out[0] - math.exp(-2.0)) <= 50 * tol
        tries = int(out[1])
        # FSAL: one evaluation at the start, six per try
        assert nrhs == 1 + 6 * tries
    # tighter tolerance -> more steps, roughly tol^(-1/5)
    _, n6 = oracle.dopri5_check(0, 2.0, 0.01, 1e-6, 1e-6)
    _, n12 = oracle.dopri5_check(0, 2.0, 0.01, 1e-12, 1e-12)
    assert 8 < n12 / n6 < 30


def test_dense_output_endpoints_and_order():
    for dt in (0.2, 0.1):
        o0, _ = oracle.dopri5_check(2, dt=dt, theta=0.0)
        assert o0[0] == 1.0
        o1, _ = oracle.dopri5_check(2, dt=dt, theta=1.0)
        assert abs(o1[0] - o1[3]) <= 1e-16
    errs = []
    for dt in (0.2, 0.1, 0.05):
        om, _ = oracle.dopri5_check(2, dt=dt, theta=0.5)
        errs.append(abs(om[0] - math.exp(0.5 * dt)))
    # continuous extension of order 4: local error ~ dt^5
    assert 20 < errs[0] / errs[1] < 45
    assert 20 < errs[1] / errs[2] < 45


def test_steps_schedule():
    # src/utils.cpp:354-366: n_min below tau_min, n_max above tau_max, quadratic in ln(tau) between
    assert oracle.steps_schedule(0.05, 100, 1000) == 100
    assert oracle.steps_schedule(200.0, 100, 1000) == 1000
    b = -math.log(0.1)
    k = 1000 / (math.log(100.0) + b) ** 2
    for tau in (0.5, 3.0, 42.0):
        assert oracle.steps_schedule(tau, 100, 1000) == math.ceil(k * (math.log(tau) + b) ** 2 + 100)


def test_math_modes_converge_to_the_same_image():
    """The oracle's two math modes (its own polynomial pow/atan2/sincos, which
    the kernels share, and the C library's) differ by last-ulp amounts; run to
    convergence (relative tolerance 1e-12 on Stokes and tau) they must give the
    same images to 1e-9 on every hit pixel with the same flags: a shared-math
    function off by more than rounding would fail this."""
    from jetshow_amd import abi, models
    cases = [models.conical_toroidal_model(), models.paraboloid_helical_model(n_side=16, pixel_pc=0.3)]
    for j, im, p in cases:
        p.stokes_abs_tol, p.stokes_rel_tol = 1e-30, 1e-12
        p.tau_abs_tol, p.tau_rel_tol = 1e-30, 1e-12
        a = oracle.observe(j, im, p, None, math=oracle.MATH_JSX, threads=8)
        b = oracle.observe(j, im, p, None, math=oracle.MATH_LIBM, threads=8)
        assert np.array_equal(a["flags"], b["flags"])
        hit = (b["flags"] & abi.PIX_HIT) != 0
        assert hit.sum() >= 40
        for k in ("tau", "I", "Q", "U", "V"):
            scale = np.abs(b[k]) if k in ("tau", "I") else np.maximum(np.abs(b[k]), np.abs(b["I"]))
            rel = np.abs(a[k] - b[k])[hit] / np.maximum(scale[hit], 1e-300)
            assert rel.max() <= 1e-9, k
