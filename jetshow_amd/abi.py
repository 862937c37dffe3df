"""ctypes mirror of include/jetshow_amd.h.

The structures here must match the C header field for field; a test checks
their sizes against the shared library's own sizeof values.
"""
import ctypes as C

JSX_OK = 0
STATUS_NAMES = {
    0: "JSX_OK", -1: "JSX_EINVAL", -2: "JSX_ENOMEM", -3: "JSX_EHIP",
    -4: "JSX_EUNSUPPORTED", -5: "JSX_EBUDGET", -6: "JSX_EBUSY",
    -7: "JSX_ENOTFOUND", -8: "JSX_ENODEV",
}

ABI_VERSION = 2
GEOM_CONE, GEOM_PARABOLOID, GEOM_CYLINDER, GEOM_CONVEX_HULL, GEOM_PARABOLOID_CONE = 1, 2, 3, 4, 5
B_RANDOM_SCALAR, B_RANDOM_SCALAR_Z, B_TOROIDAL, B_TOROIDAL_R, B_HELICAL_R, \
    B_GRID_SCALAR, B_GRID_VECTOR, B_CONST_CYLINDER_Z = 1, 2, 3, 4, 5, 6, 7, 8
B_CONST_CYLINDER, B_RADIAL_CONICAL, B_HELICAL_CYLINDER, B_SPIRAL_CONICAL, \
    B_FORCE_FREE_CYLINDRICAL, B_COMPOSITE_RANDOM_SCALAR_Z = 9, 10, 11, 12, 13, 14
N_CONST, N_BK, N_BK_Z, N_GRID, N_COMPOSITE_BK = 1, 2, 3, 4, 5
V_CONST_FLAT, V_CONST_CENTRAL, V_LINEAR_GAMMA_Z, V_GRID = 1, 2, 3, 4
V_SHEARED_FLAT, V_SHEATH_FLAT, V_SHEARED_CENTRAL, V_SHEATH_CENTRAL, V_CONST_PARABOLIC, \
    V_ACC_PARABOLIC, V_SHEARED_ACC_PARABOLIC, V_ACC_PARABOLIC_CONST_CONE, \
    V_SHEARED_ACC_PARABOLIC_CONST_CONE = 5, 6, 7, 8, 9, 10, 11, 12, 13
STOKES_I, STOKES_FULL = 1, 4
PREC_FP64, PREC_MIXED = 0, 1
OPT_CYLINDER_REF_BUG = 0x1
OPT_SHEARED_CENTRAL_REF_BUG = 0x2
OPT_FULL_DEPTH = 0x4

PIX_HIT = 0x01
PIX_NSEG2 = 0x02
PIX_TAU_CROSSED = 0x04
PIX_I_SKIPPED = 0x08
PIX_FAIL_MAXTRIALS = 0x10
PIX_NAN = 0x20
PIX_CLIPPED = 0x40
PIX_FAIL_BISECT = 0x80
PIX_HORIZON = 0x100

# constants of include/utils.h:17-36 in the reference
PC = 3.0856775814671913e+18
MAS_TO_RAD = 4.84813681109536e-09
C_LIGHT = 29979245800.0


class JetDesc(C.Structure):
    _fields_ = [
        ("geometry", C.c_int32), ("bfield", C.c_int32), ("nfield", C.c_int32), ("vfield", C.c_int32),
        ("geo_origin", C.c_double * 3), ("geo_axis", C.c_double * 3),
        ("geo_param", C.c_double), ("geo_big_scale", C.c_double),
        ("z_clip_min", C.c_double), ("z_clip_max", C.c_double),
        ("b_in_plasma_frame", C.c_int32), ("grid_b", C.c_int32), ("grid_b2", C.c_int32), ("geo_table", C.c_int32),
        ("b0", C.c_double), ("b_exp", C.c_double), ("b_pitch", C.c_double),
        ("n_in_plasma_frame", C.c_int32), ("grid_n", C.c_int32),
        ("n0", C.c_double), ("n_exp", C.c_double), ("thermal_fraction", C.c_double),
        ("grid_v", C.c_int32), ("pad1", C.c_int32),
        ("gamma0", C.c_double), ("gamma1", C.c_double), ("gamma_z0", C.c_double), ("gamma_z1", C.c_double),
        ("v_origin", C.c_double * 3),
        ("s", C.c_double), ("gamma_min", C.c_double),
        ("geo_z0", C.c_double), ("b_exp2", C.c_double), ("b_z0", C.c_double), ("b_mu", C.c_double),
        ("n_exp2", C.c_double), ("n_z0", C.c_double), ("v_r", C.c_double), ("v_theta", C.c_double),
        ("v_R0", C.c_double), ("v_Rz0", C.c_double), ("v_z0", C.c_double),
    ]


class ImageDesc(C.Structure):
    _fields_ = [
        ("n_across", C.c_int64), ("n_along", C.c_int64),
        ("lg_pixel_size_start_cm", C.c_double), ("lg_pixel_size_stop_cm", C.c_double),
        ("los_angle", C.c_double),
    ]


class SolverParams(C.Structure):
    _fields_ = [
        ("stokes", C.c_int32), ("precision", C.c_int32), ("n", C.c_int32), ("max_failed_tries", C.c_int32),
        ("nu", C.c_double), ("tau_max", C.c_double), ("tau_min", C.c_double), ("dt_max", C.c_double),
        ("tau_abs_tol", C.c_double), ("tau_rel_tol", C.c_double),
        ("stokes_abs_tol", C.c_double), ("stokes_rel_tol", C.c_double),
        ("tau_precision", C.c_double), ("tau_n_min", C.c_double), ("tau_n_max", C.c_double),
        ("n_max_factor", C.c_int32), ("flags", C.c_uint32),
    ]


class Outputs(C.Structure):
    _fields_ = [
        ("tau", C.POINTER(C.c_double)), ("I", C.POINTER(C.c_double)), ("Q", C.POINTER(C.c_double)),
        ("U", C.POINTER(C.c_double)), ("V", C.POINTER(C.c_double)), ("l", C.POINTER(C.c_double)),
        ("flags", C.POINTER(C.c_uint32)), ("n_rhs", C.POINTER(C.c_uint32)), ("n_steps", C.POINTER(C.c_uint32)),
    ]


class GridDesc(C.Structure):
    _fields_ = [
        ("nz", C.c_int32), ("nr", C.c_int32),
        ("z0", C.c_double), ("dz", C.c_double), ("r0", C.c_double), ("dr", C.c_double),
        ("nan_value", C.c_double),
    ]


class Request(C.Structure):
    _fields_ = [
        ("jet", JetDesc), ("image", ImageDesc), ("params", SolverParams), ("out_host", Outputs),
    ]


FIELD_B, FIELD_B2, FIELD_N, FIELD_V = 1, 2, 3, 4


class GridPayload(C.Structure):
    _fields_ = [("field", C.c_int32), ("pad", C.c_int32), ("desc", GridDesc), ("values", C.POINTER(C.c_double))]


class RequestTiming(C.Structure):
    _fields_ = [("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("n_pixels", C.c_int64), ("status", C.c_int32), ("pad", C.c_int32)]


def default_params(nu_hz, stokes=STOKES_I, precision=PREC_FP64, tau_max=1e4):
    """Solver constants Observation::run uses when called from pyjetshow
    (src/pyjetshow.cpp:144-156, src/Observation.cpp:23,132,156,269)."""
    p = SolverParams()
    p.stokes = stokes
    p.precision = precision
    p.n = 100
    p.max_failed_tries = 500
    p.nu = nu_hz
    p.tau_max = tau_max
    p.tau_min = 10.0 ** -10.0
    p.dt_max = (PC * 0.001) * PC        # Observation.cpp:23 scales pc*0.001 by pc again
    p.tau_abs_tol, p.tau_rel_tol = 1e-7, 1e-14
    p.stokes_abs_tol, p.stokes_rel_tol = 1e-9, 1e-9
    p.tau_precision = 1.0
    p.tau_n_min, p.tau_n_max = 0.1, 100.0
    p.n_max_factor = 10
    p.flags = 0
    return p
