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
