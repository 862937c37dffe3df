"""Parity workloads: small or sub-sampled versions of the benchmark models,
plus variants that reach every geometry / field / branch of the path."""
import math

import numpy as np

from jetshow_amd import abi, models

PC = abi.PC


def pixel_subset(image, n, seed=0, extra_cols=4):
    """A deterministic pixel subset: n random pixels plus every pixel of the
    first `extra_cols` columns of the middle rows (the apex, where rays are
    longest and tau crosses tau_max)."""
    na, nl = image.n_across, image.n_along
    rng = np.random.default_rng(seed)
    idx = set(rng.choice(na * nl, size=min(n, na * nl), replace=False).tolist())
    mid = na // 2
    for r in range(max(0, mid - 2), min(na, mid + 2)):
        for c in range(min(extra_cols, nl)):
            idx.add(r * nl + c)
    return np.array(sorted(idx), dtype=np.int64)


def reference_main_model():
    """The reference executable's hard-coded run (src/main.cpp:1440-1441):
    LOS 0.663225 rad, z = 0.0165, 500 x 60 log pixels 0.001-0.1 mas, cone
    half-angle 0.06832789, B_1 = 0.1, m = 1, K_1 = 1e4, n = 2, Gamma = 5,
    15.35 GHz, tau_max = 1e5, Stokes I (through get_i_image's model)."""
    return models.get_i_image_model(0.663225, 0.0165, 500, 60, 0.001, 0.1, 0.06832789, 0.1, 1.0, 1e4, 2.0, 5.0,
                                    15.35, 1e5, False)


def central_vfield_model():
    j, im, p = models.get_i_image_model(0.3, 0.05, 64, 32, 0.005, 0.05, 0.05, 0.2, 1.0, 2e3, 2.0, 8.0, 8.1, 1e4,
                                        True)
    j.v_origin[2] = 0.05 * PC
    return j, im, p


def toroidal_lab_frame_model():
    """Reference ToroidalBField (atan(y/x) azimuth, (z/pc)^-n) given in the lab
    frame, so getBprime runs; full Stokes."""
    j, im, p = models.conical_toroidal_model(n_side=64, pixel_pc=0.02)
    j.bfield = abi.B_TOROIDAL
    j.b_in_plasma_frame = 0
    j.n_in_plasma_frame = 0
    j.nfield = abi.N_BK_Z
    j.thermal_fraction = 0.05
    p.stokes = abi.STOKES_FULL
    return j, im, p


def cylinder_model(ref_bug=False):
    """Cylinder r = 0.3 pc clipped to 0.5 <= z <= 5 pc (the clip keeps the
    rays away from |p| = 0, where the |p|^-1 / |p|^-2 profiles diverge)."""
    j, im, p = models.conical_toroidal_model(n_side=32, pixel_pc=0.05)
    j.geometry = abi.GEOM_CYLINDER
    j.geo_param = 0.3 * PC
    j.z_clip_min, j.z_clip_max = 0.5 * PC, 5.0 * PC
    if ref_bug:
        p.flags |= abi.OPT_CYLINDER_REF_BUG
    return j, im, p


def grf_grids(nz=96, nr=24, z1=20.0, r1=0.5, seed=(1, 2, 3), sigma=0.3):
    """Small B, N, Gamma grids: the conical profiles times exp(sigma G) with G a
    unit Gaussian random field of power spectrum k^(-11/3)."""
    from jetshow_amd import grf
    z = np.linspace(0.0, z1, nz)
    r = np.linspace(0.0, r1, nr)
    Z, R = np.meshgrid(z, r, indexing="ij")
    rs = np.maximum(np.hypot(Z, R), 0.05)
    gb = grf.gaussian_random_field((nz, nr), seed[0])
    gn = grf.gaussian_random_field((nz, nr), seed[1])
    gg = grf.gaussian_random_field((nz, nr), seed[2])
    B = 0.1 / rs * np.exp(sigma * gb)
    N = 500.0 / rs ** 2 * np.exp(sigma * gn)
    G = np.maximum(1.0, 10.0 * np.exp(sigma * gg))
    dz, dr = z[1] - z[0], r[1] - r[0]
    return (B, N, G), (0.0, dz, 0.0, dr)


def grid_descs(shape, geo, nan_values=(0.0, 0.0, 1.0)):
    out = []
    for nv in nan_values:
        g = abi.GridDesc()
        g.nz, g.nr = shape
        g.z0, g.dz, g.r0, g.dr = geo
        g.nan_value = nv
        out.append(g)
    return out


def grid_model(handles, n_side=64, pixel_pc=0.04, vector_b=False):
    """Grid-defined fields on a cylinder r = 0.5 pc clipped to 0 <= z <= 20 pc
    (grid path of SimulationInterpolater)."""
    j, im, p = models.conical_toroidal_model(n_side=n_side, pixel_pc=pixel_pc)
    j.geometry = abi.GEOM_CYLINDER
    j.geo_param = 0.5 * PC
    j.z_clip_min, j.z_clip_max = 0.0, 20.0 * PC
    hb, hn, hv = handles[:3]
    if vector_b:
        j.bfield = abi.B_GRID_VECTOR
        j.grid_b, j.grid_b2 = handles[3], hb
    else:
        j.bfield = abi.B_GRID_SCALAR
        j.grid_b = hb
    j.nfield = abi.N_GRID
    j.grid_n = hn
    j.vfield = abi.V_GRID
    j.grid_v = hv
    p.stokes_abs_tol, p.stokes_rel_tol = 1e-10, 1e-8
    return j, im, p
