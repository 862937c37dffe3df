"""The kernels' own lane code, built for the CPU (tests/hostbuild), against the
oracle: the FP64 lane must be bit-identical to the oracle's shared-math mode;
the mixed-precision lane and the time-parallel split of long rays must stay
within the mixed-path accuracy bar of the FP64 result (I to 1e-5 relative).

These run without a GPU; the GPU tests check the same kernels on the device."""
import numpy as np
import pytest
from scipy.linalg import expm

from jetshow_amd import abi, models
from oracle import oracle
from tests import lanehost as LH
from tests import workloads as W


def _subset(im, n, seed=0):
    return W.pixel_subset(im, n, seed=seed)


@pytest.mark.parametrize("case", ["conical", "paraboloid_full", "reference_main"])
def test_fp64_lane_bit_exact(case):
    if case == "conical":
        j, im, p = models.conical_toroidal_model()
        idx = np.arange(128 * 128, dtype=np.int64)
    elif case == "paraboloid_full":
        j, im, p = models.paraboloid_helical_model(n_side=32, pixel_pc=0.16)
        idx = np.arange(32 * 32, dtype=np.int64)
    else:
        j, im, p = W.reference_main_model()
        idx = _subset(im, 300)
    o = oracle.observe(j, im, p, idx, math=oracle.MATH_JSX, threads=8)
    h = LH.observe(j, im, p, idx)
    for k in ("tau", "I", "Q", "U", "V", "l", "flags", "n_rhs", "n_steps"):
        assert np.array_equal(h[k], o[k]), k
    assert ((o["flags"] & abi.PIX_HIT) != 0).sum() > 20


def _errors(o, h):
    hit = (o["flags"] & abi.PIX_HIT) != 0
    assert np.array_equal(o["flags"][hit], h["flags"][hit])
    ref = o["I"][hit]
    pos = ref > 0
    rel = np.abs(h["I"][hit] - ref)[pos] / ref[pos]
    peak = ref.max()
    quv = max(np.abs(h[k][hit] - o[k][hit]).max() for k in ("Q", "U", "V")) / peak
    return rel.max(initial=0.0), quv


def test_mixed_lane_full_stokes_accuracy():
    """Paraboloid full Stokes (Faraday rotation and conversion), every pixel
    of a 48 x 48 image of 0.1 pc: I within 1e-5 relative of FP64 on every
    hit pixel; Q/U/V within 5e-5 of peak I."""
    j, im, p = models.paraboloid_helical_model(n_side=48, pixel_pc=0.1)
    idx = np.arange(48 * 48, dtype=np.int64)
    o = oracle.observe(j, im, p, idx, math=oracle.MATH_JSX, threads=8)
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    h = LH.observe(j, im, pm, idx)
    rel, quv = _errors(o, h)
    print("mixed lane: max rel I %.3g, QUV/peak %.3g, rhs mixed/fp64 %.3f" % (
        rel, quv, h["n_rhs"].sum() / o["n_rhs"].sum()))
    assert rel <= 1e-5
    assert quv <= 5e-5


def test_mixed_lane_stokes_i_accuracy():
    """Stokes I on the conical jet.  The reference's absolute tolerance (1e-10)
    dominates the error norm of faint pixels and leaves the FP64 result itself
    unconverged at the 1e-3 level there, so the comparison uses a converged
    FP64 run (relative control, abs 1e-20)."""
    j, im, p = models.conical_toroidal_model(n_side=64, pixel_pc=0.02)
    p.stokes_abs_tol = 1e-20
    idx = np.arange(64 * 64, dtype=np.int64)
    o = oracle.observe(j, im, p, idx, math=oracle.MATH_JSX, threads=8)
    pm = models.with_precision(p, abi.PREC_MIXED)
    h = LH.observe(j, im, pm, idx)
    rel, _ = _errors(o, h)
    assert rel <= 1e-5


def test_split_long_rays():
    """Rays past 4096 right-hand sides are parked and finished as pieces
    (particular solution + propagator, composed in order).  On the heaviest
    rays of the image: I within 1e-6 of FP64, Q/U/V within 5e-5 of peak I (the
    bar of the sequential mixed lane: on these Faraday-thick rays the FP32
    rounding of k_F times ~1e5 rad of rotation sets the Q/U phase error),
    flags identical, and the rays really were split."""
    j, im, p = models.paraboloid_helical_model(n_side=48, pixel_pc=0.1)
    o_all = oracle.observe(j, im, p, None, math=oracle.MATH_JSX, threads=8)
    idx = np.argsort(o_all["n_rhs"])[::-1][:4].astype(np.int64)
    o = {k: v[idx] for k, v in o_all.items()}
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    h = LH.observe(j, im, pm, idx, defer_rhs=4096, w_piece=256.0)
    assert (h["n_pieces"] > 1).all(), h["n_pieces"]
    rel, quv = _errors(o, h)
    print("split: pieces %s, max rel I %.3g, QUV/peak %.3g" % (h["n_pieces"], rel, quv))
    assert rel <= 1e-6
    assert quv <= 5e-5


def _stokes_gen(e, r):
    return np.array([[0, e[0], e[1], e[2]], [e[0], 0, r[2], -r[1]], [e[1], -r[2], 0, r[0]], [e[2], r[1], -r[0], 0]],
                    float)


def test_magnus_exponential_closed_form():
    """The exponential pieces' closed forms (cubic polynomials in the Stokes
    generator, from its eigenvalues) against scipy's general matrix
    exponential of the augmented 5x5 generator [[-mu 1 + G, v], [0, 0]], over
    absorption depths 1e-6..1e3, dichroism 1e-5..300 and Faraday/conversion
    rotation 1e-5..2e3 rad per step (both branches: Taylor series in G and
    hyperbolic/trigonometric closed form): the propagator to 1e-10 absolute,
    the source column to 1e-10 relative.  And the Lorentz-algebra commutator
    [G(e1, r1), G(e2, r2)] = G(r2 x e1 - r1 x e2, e1 x e2 - r1 x r2)."""
    rng = np.random.default_rng(5)
    cases = []
    for _ in range(1500):
        se, sr = 10 ** rng.uniform(-5, 2.5), 10 ** rng.uniform(-5, 3.3)
        e = rng.normal(size=3)
        e *= se / np.linalg.norm(e)
        r = rng.normal(size=3) * sr
        mu = max(10 ** rng.uniform(-6, 3), 1.01 * np.linalg.norm(e))   # k_I >= |eta|
        cases.append((mu, e, r))
    # the closed form's small-argument corners: tiny absorption with rotation
    # far below / above it, all eigenvalues zero (|eta| = |rho|, eta . rho = 0)
    for mu, se, sr in ((1e-12, 1e-14, 1e-13), (1e-9, 1e-12, 1e-6), (1e-6, 1e-8, 1e-12), (0.3, 1e-9, 1e-9),
                       (1e-3, 0.0, 0.0)):
        cases.append((mu, np.array([se, 0.0, 0.0]), np.array([0.0, sr, 0.0])))
    cases.append((1e-4, np.array([0.0, 2e-5, 0.0]), np.array([2e-5, 0.0, 0.0])))
    for mu, e, r in cases:
        v = rng.normal(size=4)
        A = np.zeros((5, 5))
        A[:4, :4] = -mu * np.eye(4) + _stokes_gen(e, r)
        A[:4, 4] = v
        X = expm(A)
        # both forms the kernels use: per-lane choice, and the closed form that a
        # wavefront with any large generator applies to all of its lanes
        for mode in (2, 1):
            E, w, ok = LH.magnus_expphi(mu, e, r, v, mode=mode)
            assert ok
            assert np.abs(E - X[:4, :4]).max() <= 1e-10, (mode, mu, e, r)
            assert np.abs(w - X[:4, 4]).max() <= 1e-10 * np.abs(X[:4, 4]).max(), (mode, mu, e, r)
    for _ in range(20):
        e1, r1, e2, r2 = rng.normal(size=(4, 3))
        eo, ro = LH.magnus_comm(e1, r1, e2, r2)
        C = _stokes_gen(e1, r1) @ _stokes_gen(e2, r2) - _stokes_gen(e2, r2) @ _stokes_gen(e1, r1)
        assert np.abs(C - _stokes_gen(eo, ro)).max() <= 1e-14


def test_split_long_rays_exponential_pieces():
    """The same split as test_split_long_rays with the pieces integrated by
    exponential (4th-order Magnus) propagators: same bars against FP64 (I to
    1e-6, Q/U/V to 5e-5 of peak I, flags identical) with far fewer
    coefficient evaluations than the Dormand-Prince pieces."""
    j, im, p = models.paraboloid_helical_model(n_side=48, pixel_pc=0.1)
    o_all = oracle.observe(j, im, p, None, math=oracle.MATH_JSX, threads=8)
    idx = np.argsort(o_all["n_rhs"])[::-1][:4].astype(np.int64)
    o = {k: v[idx] for k, v in o_all.items()}
    pm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    hd = LH.observe(j, im, pm, idx, defer_rhs=4096, w_piece=256.0, method=0)
    hm = LH.observe(j, im, pm, idx, defer_rhs=4096, w_piece=256.0, method=2)
    assert (hm["n_pieces"] > 1).all(), hm["n_pieces"]
    rel, quv = _errors(o, hm)
    print("exponential pieces: rhs %s (Dormand-Prince pieces %s), max rel I %.3g, QUV/peak %.3g" % (
        hm["n_rhs"], hd["n_rhs"], rel, quv))
    assert rel <= 1e-6
    assert quv <= 5e-5
    assert hm["n_rhs"].sum() * 10 < hd["n_rhs"].sum()
    # budgeted rounds: pieces past 64 right-hand sides hand the rest of their
    # interval to new pieces chained after them; same bars
    hr = LH.observe(j, im, pm, idx, defer_rhs=4096, w_piece=256.0, method=3)
    assert (hr["n_pieces"] > hm["n_pieces"]).all(), (hr["n_pieces"], hm["n_pieces"])
    rel_r, quv_r = _errors(o, hr)
    print("rounds: pieces %s, max rel I %.3g, QUV/peak %.3g" % (hr["n_pieces"], rel_r, quv_r))
    assert rel_r <= 1e-6
    assert quv_r <= 5e-5
    # the single launch's decomposition: every piece, spawned ones too, hands
    # on its rest past 48 right-hand sides; same bars
    hs = LH.observe(j, im, pm, idx, defer_rhs=4096, w_piece=256.0, method=4)
    assert (hs["n_pieces"] > hr["n_pieces"]).all(), (hs["n_pieces"], hr["n_pieces"])
    rel_s, quv_s = _errors(o, hs)
    print("single launch: pieces %s, max rel I %.3g, QUV/peak %.3g" % (hs["n_pieces"], rel_s, quv_s))
    assert rel_s <= 1e-6
    assert quv_s <= 5e-5


def test_mixed_thin_path_holds_callers_tolerance():
    """~20 pc of optically thin cylinder (conical fields, ~40 steps per ray):
    the local errors add up along the path, so the mixed solver must hold the
    state to the caller's tolerance rather than its FP32 floor.  At rel 1e-8
    it is within 1e-5 of the converged FP64 image (rel 1e-11) and of FP64 at
    1e-8 (itself ~5e-6 from converged); at its old uniform 1e-7 floor it was
    2.2e-5 off."""
    j, im, p = models.conical_toroidal_model(n_side=48, pixel_pc=0.05)
    j.geometry = abi.GEOM_CYLINDER
    j.geo_param = 0.5 * abi.PC
    j.z_clip_min, j.z_clip_max = 1.0 * abi.PC, 20.0 * abi.PC
    p.stokes_abs_tol = 1e-20
    pix = np.arange(im.n_across * im.n_along)
    o8 = oracle.observe(j, im, p, pix, math=oracle.MATH_LIBM)
    o11 = oracle.observe(j, im, models.with_precision(p, abi.PREC_FP64, rel_tol=1e-11), pix, math=oracle.MATH_LIBM)
    m = LH.observe(j, im, models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-8), pix)
    hit = (o8["flags"] & abi.PIX_HIT) != 0
    assert hit.sum() > 500
    assert np.array_equal(hit, (m["flags"] & abi.PIX_HIT) != 0)
    pos = hit & (o11["I"] > 0)
    assert (np.abs(m["I"][pos] - o11["I"][pos]) / o11["I"][pos]).max() <= 1e-5
    assert (np.abs(m["I"][pos] - o8["I"][pos]) / o8["I"][pos]).max() <= 1e-5


@pytest.mark.parametrize("w_piece", [256.0, 16384.0])
def test_exponential_pieces_through_cone_apex(w_piece):
    """A sweep model whose line of sight lies inside the cone (its rays pass
    the apex, where the power-law coefficients rise by orders of magnitude
    over a short stretch), split into exponential pieces in budgeted rounds,
    short and long (w_piece: stiffness per piece).  Three past faults of the
    exponential step's control show up here as errors of 1e-3 .. 1e25: a
    first step that could straddle the apex; the growth bound taken from the
    2nd-order generator while the 4th-order one (with the commutator) was
    applied; the particular solution's error bounded by the norm of
    exp(-mu + G) instead of phi1(-mu + G), which is ~1/mu, not e^-mu, on
    optically thick steps.  Every hit pixel within 1e-5 of the FP64 oracle."""
    sw = models.abc_sweep_models(n_models=1, sides=(64,), freqs_ghz=(8.1,))
    j, im, p, _ = sw[0]
    im.los_angle = 0.5 * j.geo_param
    q = models.with_precision(p, abi.PREC_FP64, rel_tol=1e-8)
    q.stokes_abs_tol = 1e-20
    pix = np.sort(np.random.default_rng(0).choice(im.n_across * im.n_along, 160, replace=False))
    o = oracle.observe(j, im, q, pix, math=oracle.MATH_JSX)
    m = LH.observe(j, im, p, pix, defer_rhs=512, method=3, w_piece=w_piece)
    hit = (o["flags"] & abi.PIX_HIT) != 0
    assert hit.sum() > 100 and m["n_pieces"].sum() > 100
    assert np.array_equal(hit, (m["flags"] & abi.PIX_HIT) != 0)
    rel = np.abs(m["I"][hit] - o["I"][hit]) / o["I"][hit]
    assert rel.max() <= 1e-5, rel.max()


def test_server_fails_fast_after_hip_error():
    """A launch whose completion reports a HIP error fails its requests with
    JSX_EHIP, and later submissions are refused at once instead of queueing
    to a scheduler that has stopped (they would wait forever)."""
    j, im, p = models.paraboloid_helical_model(n_side=4, pixel_pc=0.3)
    p = models.with_precision(p, abi.PREC_MIXED)
    ok, failed, after, _ = LH.server_error_path(j, im, p)
    assert ok == abi.JSX_OK
    assert failed == -3 and after == -3      # JSX_EHIP
