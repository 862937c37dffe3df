"""The further reference classes -- B fields (include/BField.h:56-164), N
fields (NField.h:20-60), V fields (VField.h:17-176) and ParabaloidCone -- in
the oracle, pinned by what each class must satisfy: closed forms of the field
laws (src/BField.cpp:53-146, NField.cpp:42-53, VField.cpp:32-268), the
geometric meaning of the flows (radial, along z, tangent to the paraboloid
through the point), continuity where a const-cone field switches parts, the
surfaces a ParabaloidCone segment starts and ends on, and images that agree
between the oracle's C-library and shared-math modes.  CPU only."""
import math

import numpy as np
import pytest
from scipy import special

from jetshow_amd import abi, models
from oracle import oracle

PC = abi.PC
C_LIGHT = 29979245800.0
NU = 15.4e9 * 1.1


def model(name):
    for n, j, im, p in models.reference_class_models():
        if n == name:
            return j, im, p
    raise KeyError(name)


def values(j, pts, opt=0, math_mode=oracle.MATH_LIBM):
    pts = np.asarray(pts, float).reshape(-1, 3)
    nl = np.tile([-math.sin(0.1), 0.0, -math.cos(0.1)], (len(pts), 1))
    return oracle.point_values(j, NU, pts, nl, opt=opt, math=math_mode)


def sample_points(n=400, seed=0, zmax=10.0, rmax=0.4):
    rng = np.random.default_rng(seed)
    z = rng.uniform(0.05, zmax, n) * PC
    r = rng.uniform(0.0, rmax, n) * PC * np.sqrt(z / PC) / 3.0
    phi = rng.uniform(-np.pi, np.pi, n)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def speed(g):
    return C_LIGHT * np.sqrt(1.0 - 1.0 / (g * g))


# ---------------------------------------------------------------- B fields

def test_const_cylinder_and_radial_conical_b():
    pts = sample_points()
    r = np.linalg.norm(pts, axis=1)
    j, _, _ = model("ConstCylinderBField")
    v = values(j, pts)
    assert (v[:, 7] == 1).all()
    np.testing.assert_allclose(v[:, 6], 0.1 * (r / PC) ** -1.0, rtol=1e-14)
    assert (v[:, 4] == 0).all() and (v[:, 5] == 0).all()
    j, _, _ = model("RadialConicalBField")
    v = values(j, pts)
    b = v[:, 4:7]
    np.testing.assert_allclose(np.linalg.norm(b, axis=1), 0.1 * (r / PC) ** -1.0, rtol=1e-14)
    # directed along p / |p|
    np.testing.assert_allclose(np.cross(b, pts / r[:, None]), 0.0, atol=1e-15 * np.abs(b).max())
    assert (np.einsum("ij,ij->i", b, pts) > 0).all()


def test_helical_cylinder_b():
    j, _, _ = model("HelicalCylinderBField")
    pts = sample_points()
    v = values(j, pts)
    rho = np.hypot(pts[:, 0], pts[:, 1])
    np.testing.assert_allclose(v[:, 4], 0.05 * np.tan(0.5 * pts[:, 1] / rho), rtol=1e-14, atol=1e-18)
    np.testing.assert_allclose(v[:, 5], -0.05 * np.tan(0.5 * pts[:, 0] / rho), rtol=1e-14, atol=1e-18)
    assert (v[:, 6] == 0.05).all()
    # on the x = 0 plane the y component vanishes and b_x = b0 tan(pitch sign(y))
    v0 = values(j, [[0.0, 0.1 * PC, PC], [0.0, -0.2 * PC, 2 * PC]])
    assert v0[0, 5] == 0.0 and v0[1, 5] == 0.0
    np.testing.assert_allclose(v0[:, 4], [0.05 * math.tan(0.5), -0.05 * math.tan(0.5)], rtol=1e-15)


def test_spiral_conical_b():
    j, _, _ = model("SpiralConicalBField")
    pts = sample_points()
    v = values(j, pts)
    b = v[:, 4:7]
    z = pts[:, 2]
    bz = 0.5 * (PC / z) ** 2
    np.testing.assert_allclose(b[:, 2], bz, rtol=1e-14)
    # transverse part = radial b_z (x, y)/z plus azimuthal b_z tan(pitch)/pc (y, -x)
    rad = bz[:, None] * pts[:, :2] / z[:, None]
    azi = (bz * math.tan(0.3) / PC)[:, None] * np.stack([pts[:, 1], -pts[:, 0]], axis=1)
    np.testing.assert_allclose(b[:, :2], rad + azi, rtol=1e-12, atol=1e-15 * np.abs(b[:, :2]).max())


def test_force_free_cylindrical_b():
    j, _, _ = model("ForceFreeCylindricalBField")
    pts = sample_points()
    v = values(j, pts)
    b = v[:, 4:7]
    mu = 1.5
    np.testing.assert_allclose(b[:, 2], 0.05 * special.i0(mu), rtol=1e-14)
    s = np.sin(np.arctan(pts[:, 1] / pts[:, 0]))
    np.testing.assert_allclose(b[:, 1], 0.05 * special.i1(mu) * s, rtol=1e-13, atol=1e-18)
    # both transverse components carry sin(fi), as the reference writes them
    np.testing.assert_array_equal(b[:, 0], -b[:, 1])


def test_composite_fields_switch_laws_at_z0():
    pts = sample_points(zmax=8.0)
    z = pts[:, 2]
    j, _, _ = model("CompositeRandomScalarBFieldZ")
    v = values(j, pts)
    assert (v[:, 7] == 0).all()
    expect = np.where(z > 3 * PC, 0.1 * (z / PC) ** -2.0, 0.1 * (z / PC) ** -1.0)
    np.testing.assert_allclose(v[:, 4], expect, rtol=1e-14)
    j, _, _ = model("CompositeBKNField")
    v = values(j, pts)
    r = np.linalg.norm(pts, axis=1)
    expect = np.where(z > 3 * PC, 500.0 * (r / PC) ** -3.0, 500.0 * (r / PC) ** -2.0)
    np.testing.assert_allclose(v[:, 11], expect, rtol=1e-14)
    # plasma-frame density: nf unchanged (n_in_plasma_frame)
    np.testing.assert_array_equal(v[:, 12], v[:, 11])


# ---------------------------------------------------------------- V fields

def test_sheared_and_sheath_flat_v():
    j, _, _ = model("ShearedFlatVField")
    rho = np.array([0.0, 0.25, 0.5, 1.0]) * PC
    pts = np.stack([rho, np.zeros(4), np.full(4, 3 * PC)], axis=1)
    v = values(j, pts)
    g = 10.0 - 7.0 * rho / PC
    np.testing.assert_allclose(v[:, 3], g, rtol=1e-12)
    np.testing.assert_allclose(v[:, 2], speed(g), rtol=1e-14)
    assert (v[:, :2] == 0).all()
    j, _, _ = model("SheathFlatVField")
    rho = np.array([0.0, 0.0999, 0.1001, 0.3]) * PC
    pts = np.stack([np.zeros(4), rho, np.full(4, 3 * PC)], axis=1)
    v = values(j, pts)
    np.testing.assert_allclose(v[:, 3], [10, 10, 3, 3], rtol=1e-12)


def cone_points(n=400, seed=1, half_angle=0.03):
    rng = np.random.default_rng(seed)
    z = rng.uniform(0.05, 10.0, n) * PC
    rho = z * np.tan(rng.uniform(0.0, half_angle, n))
    phi = rng.uniform(-np.pi, np.pi, n)
    return np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)


def test_central_v_fields():
    pts = cone_points()
    r = np.linalg.norm(pts, axis=1)
    theta = np.arccos(pts[:, 2] / r)
    j, _, _ = model("ShearedCentralVField")
    v = values(j, pts)
    g = 10.0 - 7.0 * theta / 0.05
    np.testing.assert_allclose(v[:, 3], g, rtol=1e-11)
    np.testing.assert_allclose(v[:, :3], speed(g)[:, None] * pts / r[:, None], rtol=1e-12, atol=1e-6)
    # the reference's acos(z + o_z / r) is NaN away from the origin
    vb = values(j, pts[:5], opt=abi.OPT_SHEARED_CENTRAL_REF_BUG)
    assert np.isnan(vb[:, 3]).all()
    j, _, _ = model("SheathCentralVField")
    v = values(j, pts)
    np.testing.assert_allclose(v[:, 3], np.where(theta < 0.015, 10.0, 3.0), rtol=1e-12)


def parabolic_tangent(pts, v):
    """flow tangent to the paraboloid rho^2 = k z through the point: v_z / v_rho = 2 z / rho"""
    rho = np.hypot(pts[:, 0], pts[:, 1])
    vr = (v[:, 0] * pts[:, 0] + v[:, 1] * pts[:, 1]) / rho
    np.testing.assert_allclose(v[:, 2] * rho, vr * 2 * pts[:, 2], rtol=1e-12)
    # no azimuthal part
    np.testing.assert_allclose(v[:, 0] * pts[:, 1] - v[:, 1] * pts[:, 0], 0.0, atol=1e-9 * np.abs(v).max() * PC)


def test_parabolic_v_fields():
    pts = sample_points(zmax=8.0)
    z = pts[:, 2]
    j, _, _ = model("ConstParabolicVField")
    v = values(j, pts)
    parabolic_tangent(pts, v)
    np.testing.assert_allclose(np.linalg.norm(v[:, :3], axis=1), speed(5.0), rtol=1e-13)
    on_axis = values(j, [[0.0, 0.0, 2 * PC]])
    np.testing.assert_allclose(on_axis[0, :3], [0, 0, speed(5.0)], rtol=1e-15)
    j, _, _ = model("AccParabolicVField")
    v = values(j, pts)
    parabolic_tangent(pts, v)
    np.testing.assert_allclose(v[:, 3], 1 + 9.0 * np.sqrt(z / (5 * PC)), rtol=1e-12)
    j, _, _ = model("ShearedAccParabolicVField")
    v = values(j, pts)
    parabolic_tangent(pts, v)
    rho = np.hypot(pts[:, 0], pts[:, 1])
    frac = rho / (0.1 * PC * np.sqrt(z / PC))            # r / R of the enclosing paraboloid
    g0 = np.maximum(10.0 - 7.0 * frac ** 0.25, 1.0)
    np.testing.assert_allclose(v[:, 3], 1 + (g0 - 1) * np.sqrt(z / (5 * PC)), rtol=1e-11)


@pytest.mark.parametrize("name", ["AccParabolicConstConeVField", "ShearedAccParabolicConstConeVField"])
def test_const_cone_v_fields(name):
    j, _, _ = model(name)
    z0 = 2 * PC
    rho = np.array([0.02, 0.05, 0.1]) * PC
    below = np.stack([rho, np.zeros(3), np.full(3, z0 * (1 - 1e-9))], axis=1)
    above = np.stack([rho, np.zeros(3), np.full(3, z0 * (1 + 1e-9))], axis=1)
    vb, va = values(j, below), values(j, above)
    # the cone part is radial from (0, 0, -z0), tangent to the paraboloid flow
    # at z0; the accelerating field also matches in speed there (Gamma = gamma0
    # at z = R0 = z0), the sheared one does not (its paraboloid part shears as
    # (r/R)^(1/4), its cone part linearly in the polar angle)
    ub = vb[:, :3] / np.linalg.norm(vb[:, :3], axis=1)[:, None]
    ua = va[:, :3] / np.linalg.norm(va[:, :3], axis=1)[:, None]
    np.testing.assert_allclose(ua, ub, atol=1e-8)
    if name.startswith("Acc"):
        np.testing.assert_allclose(va[:, :3], vb[:, :3], rtol=1e-8, atol=1e-8 * speed(10.0))
    far = np.array([[0.3 * PC, 0.1 * PC, 6 * PC]])
    vf = values(j, far)[0]
    d = far[0] + np.array([0, 0, z0])
    np.testing.assert_allclose(np.cross(vf[:3], d), 0.0, atol=1e-9 * speed(10.0) * PC)
    if name.startswith("Acc"):
        np.testing.assert_allclose(vf[3], 10.0, rtol=1e-13)


# ---------------------------------------------------------------- ParabaloidCone

def test_paraboloid_cone_segments_on_surfaces():
    j, im, p = model("ParabaloidCone")
    r0, z0 = 0.1 * PC, 2 * PC
    tan_a = 0.5 * r0 / math.sqrt(z0 * PC)
    rng = np.random.default_rng(7)
    n = 2000
    o = np.stack([rng.uniform(-0.5, 0.5, n) * PC, rng.uniform(-0.5, 0.5, n) * PC, rng.uniform(0.0, 8.0, n) * PC], 1)
    th = rng.uniform(0.05, 0.6, n)
    d = np.stack([-np.sin(th), np.zeros(n), -np.cos(th)], 1)
    rays = np.concatenate([o - 30 * PC * d, d], axis=1)
    ns, seg = oracle.hit(j, rays)
    assert (ns > 0).sum() > 200
    # every end lies on the cone (apex -z0, tangent at z0) or, below z0, on
    # the paraboloid; an exit can be on the cone below z0: the reference takes
    # the cone's exit whenever the paraboloid's lies above z0
    for q in np.nonzero(ns == 1)[0]:
        for pt in (seg[q, 0, :3], seg[q, 0, 3:]):
            rho = math.hypot(pt[0], pt[1])
            on_cone = abs(rho - tan_a * (pt[2] + z0)) <= 1e-9 * PC
            on_para = pt[2] >= 0 and abs(rho - r0 * math.sqrt(pt[2] / PC)) <= 1e-9 * PC
            assert on_cone or (on_para and pt[2] <= z0)
    # the cone is tangent to the paraboloid at z0
    assert abs(r0 * math.sqrt(z0 / PC) - tan_a * 2 * z0) <= 1e-12 * PC
    # points between entry and exit are inside; a little outside are not
    pv_in, pv_out = [], []
    for q in np.nonzero(ns == 1)[0][:400]:
        a, b = seg[q, 0, :3], seg[q, 0, 3:]
        L = np.linalg.norm(b - a)
        if L < 1e-3 * PC:
            continue
        u = (b - a) / L
        pv_in.append(a + 0.5 * (b - a))
        out = b + 1e-3 * L * u
        if out[2] > z0:          # below z0 is_within is the paraboloid's (below)
            pv_out.append(out)
    assert len(pv_out) > 20
    assert values(j, pv_in)[:, 29].all()
    assert not values(j, pv_out)[:, 29].any()
    # Parabaloid::is_within compares hypot(x, y) with r0 sqrt(z), z in cm
    # (Parabaloid.cpp:36): below z0 a point 10 pc off the axis counts as inside
    assert values(j, [[10 * PC, 0.0, 1 * PC]])[0, 29] == 1.0
    assert values(j, [[10 * PC, 0.0, 3 * PC]])[0, 29] == 0.0


def test_paraboloid_cone_above_z0_is_the_cone():
    """rays that meet only the cone part keep the cone's segment"""
    j, _, _ = model("ParabaloidCone")
    jc = models.copy_struct(j)
    jc.geometry = abi.GEOM_CONE
    r0, z0 = 0.1 * PC, 2 * PC
    jc.geo_param = math.atan(0.5 * r0 * (z0 * PC) ** -0.5)
    jc.geo_origin[2] = -z0
    d = np.array([-math.sin(1.2), 0.0, -math.cos(1.2)])
    o = np.array([0.05 * PC, 0.01 * PC, 6 * PC]) - 10 * PC * d
    rays = np.concatenate([o, d])[None, :]
    n1, s1 = oracle.hit(j, rays)
    n2, s2 = oracle.hit(jc, rays)
    assert n1[0] == 1 and n2[0] == 1
    np.testing.assert_allclose(s1, s2, rtol=1e-15)


# ---------------------------------------------------------------- images

DISCONTINUOUS = {"ForceFreeCylindricalBField", "CompositeRandomScalarBFieldZ", "CompositeBKNField",
                 "SheathFlatVField", "SheathCentralVField", "AccParabolicConstConeVField",
                 "ShearedAccParabolicConstConeVField"}


@pytest.mark.parametrize("name", [m[0] for m in models.reference_class_models()])
def test_images_shared_math_vs_libm(name):
    j, im, p = model(name)
    a = oracle.observe(j, im, p, None, math=oracle.MATH_JSX, threads=4)
    b = oracle.observe(j, im, p, None, math=oracle.MATH_LIBM, threads=4)
    np.testing.assert_array_equal(a["flags"], b["flags"])
    hit = (a["flags"] & abi.PIX_HIT) != 0
    assert hit.sum() > 100
    peak = np.abs(b["I"]).max()
    assert peak > 0 and np.isfinite(a["I"]).all()
    # the shared math stays within a few ulp of the C library: on smooth
    # fields the step sequences coincide and the images agree to rounding; a
    # field with a jump (a sheath's Gamma step, a composite law switching at
    # z0, the const-cone fields' switch, sin(atan(y/x)) flipping sign at x = 0)
    # lets those ulp decide on which side a stage lands, and the step
    # sequences part there (measured: up to 1e-6 of the peak for the sheath
    # in polar angle)
    jumps = name in DISCONTINUOUS
    for k in ("I", "Q", "U", "V"):
        assert np.abs(a[k] - b[k]).max() <= (2e-6 if jumps else 1e-12) * peak
