"""Jet-model builders: POD descriptors for the solver's C ABI.

Each builder returns (JetDesc, ImageDesc, SolverParams).  `get_i_image_model`
reproduces exactly what the reference's pybind11 entry point builds
(src/pyjetshow.cpp:69-156); the others are the synthetic workloads the
benchmarks and parity tests run, with every size and constant stated.
"""
import math

import numpy as np

from . import abi

PC = abi.PC


def _jet():
    j = abi.JetDesc()
    j.geo_axis[2] = 1.0
    j.gamma_min = 100.0
    j.s = 3.42
    return j


def get_i_image_model(los_angle, redshift, number_of_pixels_along, number_of_pixels_across,
                      pixel_size_mas_start, pixel_size_mas_stop, cone_half_angle, B_1, m, K_1, n,
                      Gamma, nu_observed_ghz, tau_max=1e4, central_vfield=False, mas_to_pc=None):
    """Descriptors for pyjetshow.get_i_image (src/pyjetshow.cpp:69-156):
    Cone(origin 0, +z, alpha, big_scale = 100 [cm, as passed at :79]),
    RandomScalarBFieldZ(B_1, m), BKNFieldZ(K_1, n, plasma frame),
    ConstFlat or ConstCentral V(Gamma), s = 3.42 (src/Jet.cpp:32)."""
    j = _jet()
    j.geometry = abi.GEOM_CONE
    j.geo_param = cone_half_angle
    j.geo_big_scale = 100.0
    j.bfield = abi.B_RANDOM_SCALAR_Z
    j.b_in_plasma_frame = 1
    j.b0, j.b_exp = B_1, m
    j.nfield = abi.N_BK_Z
    j.n_in_plasma_frame = 1
    j.n0, j.n_exp = K_1, n
    j.vfield = abi.V_CONST_CENTRAL if central_vfield else abi.V_CONST_FLAT
    j.gamma0 = Gamma
    if mas_to_pc is None:
        from .host import mas_to_pc as _m2p
        mas_to_pc = _m2p
    pc_in_mas = mas_to_pc(redshift)
    im = abi.ImageDesc()
    im.n_across, im.n_along = number_of_pixels_across, number_of_pixels_along
    im.lg_pixel_size_start_cm = math.log10(pixel_size_mas_start * pc_in_mas * PC)
    im.lg_pixel_size_stop_cm = math.log10(pixel_size_mas_stop * pc_in_mas * PC)
    im.los_angle = los_angle
    nu = nu_observed_ghz * 1E+09
    nu = nu * (1. + redshift)
    p = abi.default_params(nu, tau_max=tau_max)
    return j, im, p


def conical_toroidal_model(n_side=128, pixel_pc=0.05, los=0.1, redshift=0.1, nu_ghz=15.4):
    """Plain Stokes-I conical jet: Cone apex 0, axis +z, half-angle 0.01 rad,
    clipped to 0 <= z <= 10 pc; toroidal B = 0.1 G (r/pc)^-1; N = 500 cm^-3
    (r/pc)^-2 with s = 2.5 (r = |p|, both in the plasma frame); Gamma = 10
    along +z; uniform pixels; tau_max = 1e4; tau tolerance 1e-7/1e-14;
    I tolerance abs 1e-10 / rel 1e-8."""
    j = _jet()
    j.geometry = abi.GEOM_CONE
    j.geo_param = 0.01
    j.geo_big_scale = 100.0 * PC
    j.z_clip_min, j.z_clip_max = 0.0, 10.0 * PC
    j.bfield = abi.B_TOROIDAL_R
    j.b_in_plasma_frame = 1
    j.b0, j.b_exp = 0.1, 1.0
    j.nfield = abi.N_BK
    j.n_in_plasma_frame = 1
    j.n0, j.n_exp = 500.0, 2.0
    j.vfield = abi.V_CONST_FLAT
    j.gamma0 = 10.0
    j.s = 2.5
    im = abi.ImageDesc()
    im.n_across = im.n_along = n_side
    im.lg_pixel_size_start_cm = im.lg_pixel_size_stop_cm = math.log10(pixel_pc * PC)
    im.los_angle = los
    p = abi.default_params(nu_ghz * 1e9 * (1.0 + redshift))
    p.stokes_abs_tol, p.stokes_rel_tol = 1e-10, 1e-8
    return j, im, p


def paraboloid_helical_model(n_side=512, pixel_pc=0.01, los=0.05, redshift=0.1, nu_ghz=15.4,
                             stokes=abi.STOKES_FULL, precision=abi.PREC_FP64):
    """Full-Stokes paraboloid jet: x^2 + y^2 = (r0^2/pc) z with r0 = 0.1 pc;
    helical B, pitch 45 deg, |B| = 0.2 G (r/pc)^-1; relativistic N = 500
    (r/pc)^-2, s = 2.5, plus cold thermal electrons at 0.1 x that density
    (Faraday rotation and conversion); Gamma linear from 2 at z = 0 to 15 at
    z = 5 pc, clamped; LOS 0.05 rad; uniform 0.01 pc pixels (512 x 512 covers
    5.12 pc across and z up to ~100 pc along).  Finer pixels put more rays
    through the Faraday-thick apex, where single rays need up to 1e7 right-hand
    side evaluations."""
    j = _jet()
    j.geometry = abi.GEOM_PARABOLOID
    j.geo_param = 0.1 * PC
    j.geo_big_scale = 100.0 * PC
    j.bfield = abi.B_HELICAL_R
    j.b_in_plasma_frame = 1
    j.b0, j.b_exp, j.b_pitch = 0.2, 1.0, math.pi / 4.0
    j.nfield = abi.N_BK
    j.n_in_plasma_frame = 1
    j.n0, j.n_exp = 500.0, 2.0
    j.thermal_fraction = 0.1
    j.vfield = abi.V_LINEAR_GAMMA_Z
    j.gamma0, j.gamma1 = 2.0, 15.0
    j.gamma_z0, j.gamma_z1 = 0.0, 5.0 * PC
    j.s = 2.5
    im = abi.ImageDesc()
    im.n_across = im.n_along = n_side
    im.lg_pixel_size_start_cm = im.lg_pixel_size_stop_cm = math.log10(pixel_pc * PC)
    im.los_angle = los
    p = abi.default_params(nu_ghz * 1e9 * (1.0 + redshift), stokes=stokes, precision=precision)
    # Intensities in this image span ~1e-10 .. 1e-4 (cgs), so an absolute
    # tolerance of 1e-10 would dominate the error norm and leave faint pixels
    # with ~3e-4 relative error; relative control (abs 1e-20) converges the
    # FP64 image to ~1e-8.
    p.stokes_abs_tol, p.stokes_rel_tol = 1e-20, 1e-8
    return j, im, p


def with_precision(params, precision, rel_tol=None, abs_tol=None):
    """Copy of `params` with another precision mode (and tolerances)."""
    q = abi.SolverParams.from_buffer_copy(params)
    q.precision = precision
    if rel_tol is not None:
        q.stokes_rel_tol = rel_tol
    if abs_tol is not None:
        q.stokes_abs_tol = abs_tol
    return q


def copy_struct(s):
    return type(s).from_buffer_copy(s)


def abc_sweep_models(n_models=4096, seed=20261203, freqs_ghz=(8.1, 15.4), redshift=0.1, sides=(64, 128, 256, 512),
                     precision=abi.PREC_MIXED, rel_tol=1e-6, mas_to_pc=None):
    """The batched ABC-style sweep: cone jets with parameters drawn from the
    priors, each imaged at every frequency in `freqs_ghz` on a JetModelZoom-
    style log-spaced pixel grid (pixel sizes 10^-3 .. 10^-1 mas along the jet,
    jetmodel.py:212, converted at `redshift`), full Stokes.

    Priors (numpy default_rng(seed), drawn in this order per model): LOS
    U[0.02, 0.3] rad; half-angle U[0.005, 0.05] rad; Gamma U[2, 30]; B0
    log-U[0.01, 1] G; N0 log-U[10, 1e4] cm^-3; s U[2, 3]; image side uniform
    over `sides` (square images).  Fields: helical B (pitch 45 deg) of
    magnitude B0 (|p|/pc)^-1, relativistic N0 (|p|/pc)^-2, both in the plasma
    frame; ConstFlat Gamma along +z; cone clipped to 0 <= z <= 10 pc.

    Returns a list of (JetDesc, ImageDesc, SolverParams, model_index)."""
    if mas_to_pc is None:
        from .host import mas_to_pc as _m2p
        mas_to_pc = _m2p
    rng = np.random.default_rng(seed)
    pc_in_mas = mas_to_pc(redshift)
    out = []
    for k in range(n_models):
        los = rng.uniform(0.02, 0.3)
        alpha = rng.uniform(0.005, 0.05)
        gamma = rng.uniform(2.0, 30.0)
        b0 = 10 ** rng.uniform(-2.0, 0.0)
        n0 = 10 ** rng.uniform(1.0, 4.0)
        s = rng.uniform(2.0, 3.0)
        side = int(sides[rng.integers(0, len(sides))])
        j = _jet()
        j.geometry = abi.GEOM_CONE
        j.geo_param = alpha
        j.geo_big_scale = 100.0 * PC
        j.z_clip_min, j.z_clip_max = 0.0, 10.0 * PC
        j.bfield = abi.B_HELICAL_R
        j.b_in_plasma_frame = 1
        j.b0, j.b_exp, j.b_pitch = b0, 1.0, math.pi / 4.0
        j.nfield = abi.N_BK
        j.n_in_plasma_frame = 1
        j.n0, j.n_exp = n0, 2.0
        j.vfield = abi.V_CONST_FLAT
        j.gamma0 = gamma
        j.s = s
        im = abi.ImageDesc()
        im.n_across = im.n_along = side
        im.lg_pixel_size_start_cm = math.log10(1e-3 * pc_in_mas * PC)
        im.lg_pixel_size_stop_cm = math.log10(1e-1 * pc_in_mas * PC)
        im.los_angle = los
        for f in freqs_ghz:
            p = abi.default_params(f * 1e9 * (1.0 + redshift), stokes=abi.STOKES_FULL, precision=precision)
            p.stokes_abs_tol = 1e-20
            p.stokes_rel_tol = rel_tol if precision == abi.PREC_MIXED else 1e-8
            out.append((j, im, p, k))
    return out


def grid_jet_model(handles, n_side=1024, pixel_pc=0.002, los=0.1, redshift=0.1, nu_ghz=15.4,
                   precision=abi.PREC_FP64, stokes=abi.STOKES_I):
    """Grid-defined fields (SimulationInterpolater path): a cylinder r = 0.5 pc
    clipped to 0 <= z <= 20 pc, with tangled |B| (grid), N (grid) and Gamma
    (grid, along +z) interpolated from the node grids referenced by
    `handles` = (B, N, Gamma) -- see grf.jet_grids; s = 2.5; LOS 0.1 rad;
    uniform pixels; tau_max 1e4; Stokes I with relative control (abs 1e-20,
    rel 1e-8 in FP64, 1e-9 mixed).  The interpolated fields are piecewise
    linear, with kinks at every cell edge and diagonal that the
    Dormand-Prince error estimate undershoots: FP64 at 1e-8 lands ~1e-4 from
    the converged image (48^2 image, 512 x 128 grids), FP64 at 1e-10 ~2e-6,
    the mixed solver at 1e-9 ~8e-6."""
    j = _jet()
    j.geometry = abi.GEOM_CYLINDER
    j.geo_param = 0.5 * PC
    j.geo_big_scale = 100.0 * PC
    j.z_clip_min, j.z_clip_max = 0.0, 20.0 * PC
    hb, hn, hv = handles[:3]
    j.bfield = abi.B_GRID_SCALAR
    j.grid_b = hb
    j.b_in_plasma_frame = 1
    j.nfield = abi.N_GRID
    j.grid_n = hn
    j.n_in_plasma_frame = 1
    j.vfield = abi.V_GRID
    j.grid_v = hv
    j.s = 2.5
    im = abi.ImageDesc()
    im.n_across = im.n_along = n_side
    im.lg_pixel_size_start_cm = im.lg_pixel_size_stop_cm = math.log10(pixel_pc * PC)
    im.los_angle = los
    p = abi.default_params(nu_ghz * 1e9 * (1.0 + redshift), stokes=stokes, precision=precision)
    p.stokes_abs_tol, p.stokes_rel_tol = 1e-20, 1e-8 if precision == abi.PREC_FP64 else 1e-9
    return j, im, p


def reference_class_models(n_side=24, pixel_pc=0.05, los=0.1, redshift=0.1, nu_ghz=15.4,
                           stokes=abi.STOKES_FULL, precision=abi.PREC_FP64):
    """One small model per further reference class (the B, N, V fields and the
    ParabaloidCone geometry of include/BField.h:56-164, NField.h:20-60,
    VField.h:17-176, ParabaloidCone.h), each in a jet where it shapes the
    image: a cone of half-angle 0.03 rad (apex 0, +z, clipped to 0..12 pc), a
    paraboloid r = 0.1 pc (z/pc)^1/2, or a ParabaloidCone of r0 = 0.1 pc
    switching to its tangent cone at z0 = 2 pc.  Unless the class under test
    replaces them: helical B 0.1 G (|p|/pc)^-1 with pitch 45 deg given in the
    lab frame, N = 500 (|p|/pc)^-2 in the plasma frame, ConstFlat Gamma = 5,
    s = 2.5.  LOS 0.1 rad, uniform pixels, relative Stokes control (abs 1e-20,
    rel 1e-8).  Returns a list of (name, JetDesc, ImageDesc, SolverParams)."""
    def base(geometry="cone"):
        j = _jet()
        j.geo_big_scale = 100.0 * PC
        if geometry == "cone":
            j.geometry = abi.GEOM_CONE
            j.geo_param = 0.03
            j.z_clip_min, j.z_clip_max = 0.0, 12.0 * PC
        elif geometry == "paraboloid":
            j.geometry = abi.GEOM_PARABOLOID
            j.geo_param = 0.1 * PC
        else:
            j.geometry = abi.GEOM_PARABOLOID_CONE
            j.geo_param = 0.1 * PC
            j.geo_z0 = 2.0 * PC
        j.bfield = abi.B_HELICAL_R
        j.b_in_plasma_frame = 0
        j.b0, j.b_exp, j.b_pitch = 0.1, 1.0, math.pi / 4.0
        j.nfield = abi.N_BK
        j.n_in_plasma_frame = 1
        j.n0, j.n_exp = 500.0, 2.0
        j.vfield = abi.V_CONST_FLAT
        j.gamma0 = 5.0
        j.s = 2.5
        return j

    out = []

    def add(name, j):
        im = abi.ImageDesc()
        im.n_across = im.n_along = n_side
        im.lg_pixel_size_start_cm = im.lg_pixel_size_stop_cm = math.log10(pixel_pc * PC)
        im.los_angle = los
        p = abi.default_params(nu_ghz * 1e9 * (1.0 + redshift), stokes=stokes, precision=precision)
        p.stokes_abs_tol, p.stokes_rel_tol = 1e-20, 1e-8
        out.append((name, j, im, p))

    j = base("paraboloid_cone")
    add("ParabaloidCone", j)
    j = base(); j.bfield = abi.B_CONST_CYLINDER; j.b0, j.b_exp = 0.1, 1.0
    add("ConstCylinderBField", j)
    j = base(); j.bfield = abi.B_RADIAL_CONICAL; j.b0, j.b_exp = 0.1, 1.0
    add("RadialConicalBField", j)
    j = base(); j.bfield = abi.B_HELICAL_CYLINDER; j.b0, j.b_pitch = 0.05, 0.5
    add("HelicalCylinderBField", j)
    j = base(); j.bfield = abi.B_SPIRAL_CONICAL; j.b0, j.b_pitch = 0.5, 0.3
    add("SpiralConicalBField", j)
    j = base(); j.bfield = abi.B_FORCE_FREE_CYLINDRICAL; j.b0, j.b_mu = 0.05, 1.5
    add("ForceFreeCylindricalBField", j)
    j = base(); j.bfield = abi.B_COMPOSITE_RANDOM_SCALAR_Z; j.b_in_plasma_frame = 1
    j.b0, j.b_exp, j.b_exp2, j.b_z0 = 0.1, 1.0, 2.0, 3.0 * PC
    add("CompositeRandomScalarBFieldZ", j)
    j = base(); j.nfield = abi.N_COMPOSITE_BK; j.n_exp2, j.n_z0 = 3.0, 3.0 * PC
    add("CompositeBKNField", j)
    j = base(); j.vfield = abi.V_SHEARED_FLAT; j.gamma0, j.gamma1, j.v_r = 10.0, 3.0, 1.0 * PC
    add("ShearedFlatVField", j)
    j = base(); j.vfield = abi.V_SHEATH_FLAT; j.gamma0, j.gamma1, j.v_r = 10.0, 3.0, 0.1 * PC
    add("SheathFlatVField", j)
    j = base(); j.vfield = abi.V_SHEARED_CENTRAL; j.gamma0, j.gamma1, j.v_theta = 10.0, 3.0, 0.05
    add("ShearedCentralVField", j)
    j = base(); j.vfield = abi.V_SHEATH_CENTRAL; j.gamma0, j.gamma1, j.v_theta = 10.0, 3.0, 0.015
    add("SheathCentralVField", j)
    j = base("paraboloid"); j.vfield = abi.V_CONST_PARABOLIC; j.gamma0, j.v_Rz0 = 5.0, 0.1 * PC
    add("ConstParabolicVField", j)
    j = base("paraboloid"); j.vfield = abi.V_ACC_PARABOLIC
    j.gamma0, j.v_R0, j.v_Rz0 = 10.0, 5.0 * PC, 0.1 * PC
    add("AccParabolicVField", j)
    j = base("paraboloid"); j.vfield = abi.V_SHEARED_ACC_PARABOLIC
    j.gamma0, j.gamma1, j.v_R0, j.v_Rz0 = 10.0, 3.0, 5.0 * PC, 0.1 * PC
    add("ShearedAccParabolicVField", j)
    j = base("paraboloid_cone"); j.vfield = abi.V_ACC_PARABOLIC_CONST_CONE
    j.gamma0, j.v_Rz0, j.v_z0 = 10.0, 0.1 * PC, 2.0 * PC
    add("AccParabolicConstConeVField", j)
    j = base("paraboloid_cone"); j.vfield = abi.V_SHEARED_ACC_PARABOLIC_CONST_CONE
    j.gamma0, j.gamma1, j.v_Rz0, j.v_z0 = 10.0, 3.0, 0.1 * PC, 2.0 * PC
    add("ShearedAccParabolicConstConeVField", j)
    return out
