"""The C++ class API on the GPU (tests/cxx/class_api.cpp): field values of
every reference B / N / V class, is_within, ParabaloidCone::hit, Jet's
getters and an Observation::run, each against the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

from jetshow_amd import abi, models
from oracle import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PC = abi.PC
C_LIGHT = 29979245800.0


@pytest.fixture(scope="module")
def printed(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cxx") / "class_api")
    r = subprocess.run(["bash", os.path.join(HERE, "cxx", "build.sh"), exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vals = {}
    for ln in r.stdout.splitlines():
        tag, *xs = ln.split()
        vals[tag] = np.array([float(x) for x in xs])
    return vals


def desc(**kw):
    """the class API's stand-alone descriptor (a class lowered over plain
    defaults), with the class's own fields set"""
    j = abi.JetDesc()
    j.geometry, j.geo_param, j.geo_big_scale = abi.GEOM_CONE, 0.1, 100.0
    j.geo_axis[2] = 1.0
    j.bfield, j.b0 = abi.B_RANDOM_SCALAR, 1.0
    j.nfield, j.n0 = abi.N_CONST, 1.0
    j.vfield, j.gamma0 = abi.V_CONST_FLAT, 2.0
    j.s, j.gamma_min = 2.5, 100.0
    for k, v in kw.items():
        if k in ("geo_origin", "v_origin"):
            for i in range(3):
                getattr(j, k)[i] = v[i]
        else:
            setattr(j, k, v)
    return j


P = np.array([0.02 * PC, 0.01 * PC, 2.0 * PC])
Q = np.array([-0.05 * PC, 0.03 * PC, 4.0 * PC])


def at(j, pt, nl=(0.0, 0.0, 1.0), nu=1e9, opt=0):
    return oracle.point_values(j, nu, np.array(pt)[None], np.array(nl)[None], opt=opt, math=oracle.MATH_JSX)[0]


B_CASES = {
    "RandomScalarBField": (dict(bfield=abi.B_RANDOM_SCALAR, b0=0.1, b_exp=1.0, b_in_plasma_frame=1), P, 1),
    "RandomScalarBFieldZ": (dict(bfield=abi.B_RANDOM_SCALAR_Z, b0=0.1, b_exp=1.0, b_in_plasma_frame=1), P, 1),
    "CompositeRandomScalarBFieldZ.p": (dict(bfield=abi.B_COMPOSITE_RANDOM_SCALAR_Z, b0=0.1, b_exp=1.0, b_exp2=2.0,
                                            b_z0=3 * PC, b_in_plasma_frame=1), P, 1),
    "CompositeRandomScalarBFieldZ.q": (dict(bfield=abi.B_COMPOSITE_RANDOM_SCALAR_Z, b0=0.1, b_exp=1.0, b_exp2=2.0,
                                            b_z0=3 * PC, b_in_plasma_frame=1), Q, 1),
    "ConstCylinderBField": (dict(bfield=abi.B_CONST_CYLINDER, b0=0.1, b_exp=1.0, b_in_plasma_frame=1), P, 3),
    "ConstCylinderBFieldZ": (dict(bfield=abi.B_CONST_CYLINDER_Z, b0=0.1, b_exp=1.0, b_in_plasma_frame=1), P, 3),
    "RadialConicalBField": (dict(bfield=abi.B_RADIAL_CONICAL, b0=0.1, b_exp=1.0), P, 3),
    "ToroidalBField": (dict(bfield=abi.B_TOROIDAL, b0=0.1, b_exp=1.0, b_in_plasma_frame=1), P, 3),
    "HelicalCylinderBField": (dict(bfield=abi.B_HELICAL_CYLINDER, b0=0.05, b_pitch=0.5, b_in_plasma_frame=1), P, 3),
    "SpiralConicalBField": (dict(bfield=abi.B_SPIRAL_CONICAL, b0=0.5, b_pitch=0.3, b_in_plasma_frame=1), P, 3),
    "ForceFreeCylindricalBField": (dict(bfield=abi.B_FORCE_FREE_CYLINDRICAL, b0=0.05, b_mu=1.5,
                                        b_in_plasma_frame=1), P, 3),
}


def test_b_fields(printed):
    for tag, (kw, pt, n) in B_CASES.items():
        got = printed[tag]
        want = at(desc(**kw), pt)[4:4 + n]
        assert np.array_equal(got, want), (tag, got, want)


def test_b_plasma_frame(printed):
    v = np.array([0.0, 0.0, 0.9 * C_LIGHT])
    g = 1.0 / math.sqrt(1.0 - 0.81)
    for tag, kw in (("RadialConicalBField.plasma", B_CASES["RadialConicalBField"][0]),
                    ("ToroidalBField.plasma", B_CASES["ToroidalBField"][0])):
        b = at(desc(**kw), P)[4:7]
        if kw.get("b_in_plasma_frame", 0):
            want = b
        else:   # getBprime (src/utils.cpp:164-168)
            want = b / g + g / ((1 + g) * C_LIGHT ** 2) * v * np.dot(v, b)
        np.testing.assert_allclose(printed[tag], want, rtol=1e-14)
    assert printed["RandomScalarBField.plasma"][0] == printed["RandomScalarBField"][0]


def test_n_fields(printed):
    cases = {
        "ConstNField": (dict(nfield=abi.N_CONST, n0=300.0, n_in_plasma_frame=1), P),
        "BKNField": (dict(nfield=abi.N_BK, n0=500.0, n_exp=2.0, n_in_plasma_frame=1), P),
        "BKNFieldZ": (dict(nfield=abi.N_BK_Z, n0=500.0, n_exp=2.0), P),
        "CompositeBKNField.p": (dict(nfield=abi.N_COMPOSITE_BK, n0=500.0, n_exp=2.0, n_exp2=3.0, n_z0=3 * PC), P),
        "CompositeBKNField.q": (dict(nfield=abi.N_COMPOSITE_BK, n0=500.0, n_exp=2.0, n_exp2=3.0, n_z0=3 * PC), Q),
    }
    for tag, (kw, pt) in cases.items():
        assert printed[tag][0] == at(desc(**kw), pt)[11], tag
    assert printed["BKNFieldZ.plasma"][0] == printed["BKNFieldZ"][0] / 5.0


def test_v_fields(printed):
    cases = {
        "ConstFlatVField": (dict(vfield=abi.V_CONST_FLAT, gamma0=5.0), P),
        "ShearedFlatVField": (dict(vfield=abi.V_SHEARED_FLAT, gamma0=10.0, gamma1=3.0, v_r=PC), P),
        "SheathFlatVField": (dict(vfield=abi.V_SHEATH_FLAT, gamma0=10.0, gamma1=3.0, v_r=0.015 * PC), P),
        "ConstCentralVField": (dict(vfield=abi.V_CONST_CENTRAL, gamma0=5.0), P),
        "ShearedCentralVField": (dict(vfield=abi.V_SHEARED_CENTRAL, gamma0=10.0, gamma1=3.0, v_theta=0.05), P),
        "SheathCentralVField": (dict(vfield=abi.V_SHEATH_CENTRAL, gamma0=10.0, gamma1=3.0, v_theta=0.015), P),
        "ConstParabolicVField": (dict(vfield=abi.V_CONST_PARABOLIC, gamma0=5.0, v_Rz0=0.1 * PC), P),
        "AccParabolicVField": (dict(vfield=abi.V_ACC_PARABOLIC, gamma0=10.0, v_R0=5 * PC, v_Rz0=0.1 * PC), P),
        "ShearedAccParabolicVField": (dict(vfield=abi.V_SHEARED_ACC_PARABOLIC, gamma0=10.0, gamma1=3.0,
                                           v_R0=5 * PC, v_Rz0=0.1 * PC), P),
        "AccParabolicConstConeVField.p": (dict(vfield=abi.V_ACC_PARABOLIC_CONST_CONE, gamma0=10.0, v_Rz0=0.1 * PC,
                                               v_z0=2 * PC), P),
        "AccParabolicConstConeVField.q": (dict(vfield=abi.V_ACC_PARABOLIC_CONST_CONE, gamma0=10.0, v_Rz0=0.1 * PC,
                                               v_z0=2 * PC), Q),
        "ShearedAccParabolicConstConeVField.q": (dict(vfield=abi.V_SHEARED_ACC_PARABOLIC_CONST_CONE, gamma0=10.0,
                                                      gamma1=3.0, v_Rz0=0.1 * PC, v_z0=2 * PC), Q),
    }
    for tag, (kw, pt) in cases.items():
        assert np.array_equal(printed[tag], at(desc(**kw), pt)[0:3]), tag


def test_geometries(printed):
    o = (0.0, 0.0, 0.0)
    cone = desc(geometry=abi.GEOM_CONE, geo_param=0.03, geo_big_scale=100 * PC)
    cyl = desc(geometry=abi.GEOM_CYLINDER, geo_param=0.3 * PC)
    para = desc(geometry=abi.GEOM_PARABOLOID, geo_param=0.1 * PC, geo_big_scale=100 * PC)
    pcone = desc(geometry=abi.GEOM_PARABOLOID_CONE, geo_param=0.1 * PC, geo_z0=2 * PC, geo_big_scale=100 * PC)
    assert printed["Cone.is_within"][0] == at(cone, P)[29]
    assert printed["Cylinder.is_within"][0] == at(cyl, Q)[29]
    assert printed["Parabaloid.is_within"][0] == at(para, P)[29]
    assert printed["ParabaloidCone.is_within.p"][0] == at(pcone, P)[29]
    assert printed["ParabaloidCone.is_within.q"][0] == at(pcone, Q)[29]
    d = np.array([-math.sin(0.3), 0.0, -math.cos(0.3)])
    ray = np.concatenate([[0.01 * PC, 0.0, 20 * PC], d])[None]
    ns, seg = oracle.hit(pcone, ray)
    assert printed["ParabaloidCone.hit.n"][0] == ns[0] == 1
    assert np.array_equal(printed["ParabaloidCone.hit.in"], seg[0, 0, :3])
    assert np.array_equal(printed["ParabaloidCone.hit.out"], seg[0, 0, 3:])


def test_jet_getters(printed):
    j = desc(geometry=abi.GEOM_PARABOLOID_CONE, geo_param=0.1 * PC, geo_z0=2 * PC, geo_big_scale=100 * PC,
             bfield=abi.B_HELICAL_R, b0=0.1, b_exp=1.0, b_pitch=0.7853981633974483, b_in_plasma_frame=0,
             nfield=abi.N_BK, n0=500.0, n_exp=2.0, n_in_plasma_frame=1,
             vfield=abi.V_ACC_PARABOLIC_CONST_CONE, gamma0=10.0, v_Rz0=0.1 * PC, v_z0=2 * PC)
    j.s, j.gamma_min = 2.5, 100.0
    nl = (math.sin(0.1), 0.0, math.cos(0.1))
    row = at(j, P, nl, 15.4e9)
    for tag, k in (("Jet.getKI", 18), ("Jet.getEtaI", 25), ("Jet.getKQ", 19), ("Jet.getKF", 22),
                   ("Jet.getEtaV", 28)):
        assert printed[tag][0] == row[k], tag
    assert np.array_equal(printed["Jet.getV"], at(j, P)[0:3])
    # utils.h frame transforms on the host against the kernels' values
    np.testing.assert_allclose(printed["getG"][0], row[3], rtol=1e-15)
    np.testing.assert_allclose(printed["getD"][0], row[13], rtol=1e-14)
    np.testing.assert_allclose(printed["get_n_los_prime"], row[14:17], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(printed["getBprime"], row[8:11], rtol=1e-13)


def test_observation_like_main(printed):
    j, im, p = models.get_i_image_model(0.3, 0.05, 24, 16, 0.005, 0.05, 0.05, 0.2, 1.0, 2e3, 2.0, 8.0, 8.1, 1e4,
                                        True, mas_to_pc=oracle.mas_to_pc)
    sizes, _, _ = oracle.image(im)
    np.testing.assert_allclose(printed["pixel_size.0.0"][0], sizes[0], rtol=1e-15)
    np.testing.assert_allclose(printed["pixel_size.last"][0], sizes[-1], rtol=1e-15)
    o = oracle.observe(j, im, p, None, math=oracle.MATH_JSX)
    assert ((o["flags"] & abi.PIX_HIT) != 0).sum() > 20
    tsum = isum = 0.0
    for x in o["tau"]:
        tsum += x
    for x in o["I"]:
        isum += x
    assert printed["Observation.tau.sum"][0] == tsum
    assert printed["Observation.I.sum"][0] == isum
    assert printed["Observation.I.max"][0] == o["I"].max()
