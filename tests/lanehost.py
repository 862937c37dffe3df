"""ctypes wrapper of tests/hostbuild/liblanehost.so -- the kernels' lane code
compiled for the CPU (test infrastructure; see tests/hostbuild/lane_host.cpp).
Built on first use with g++ (a few seconds)."""
import ctypes as C
import os
import subprocess

import numpy as np

from jetshow_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "hostbuild", "lane_host.cpp")
LIB = os.path.join(HERE, "hostbuild", "liblanehost.so")
DEPS = [SRC, os.path.join(HERE, "hostbuild", "hip", "hip_runtime.h"), os.path.join(ROOT, "include", "jetshow_amd.h")]
_lib = None


def _sources():
    csrc = os.path.join(ROOT, "jetshow_amd", "csrc")
    return DEPS + [os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith((".h", ".hip"))]


def build(force=False):
    if not force and os.path.exists(LIB):
        t = os.path.getmtime(LIB)
        if all(os.path.getmtime(s) <= t for s in _sources()):
            return LIB
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-w",
           "-I", os.path.join(HERE, "hostbuild"), "-I", os.path.join(ROOT, "include"), SRC, os.path.join(ROOT, "jetshow_amd", "csrc", "jsx_textio.cpp"), "-o", LIB]
    subprocess.run(cmd, check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(build())
        _lib.lh_observe.restype = C.c_int
        _lib.lh_observe.argtypes = [C.POINTER(abi.JetDesc), C.POINTER(abi.ImageDesc), C.POINTER(abi.SolverParams),
                                    C.POINTER(C.c_int64), C.c_int64, C.c_int32, C.POINTER(abi.GridDesc),
                                    C.POINTER(C.POINTER(C.c_double)), C.c_uint, C.c_float, C.c_int32,
                                    C.POINTER(abi.Outputs), C.POINTER(C.c_int32), C.c_int32]
        _lib.lh_magnus_expphi.restype = C.c_int
        _lib.lh_magnus_expphi.argtypes = [C.c_double] + [C.POINTER(C.c_double)] * 5
        _lib.lh_split_costs.restype = C.c_int64
        _lib.lh_split_costs.argtypes = [C.POINTER(C.c_double), C.c_int64]
        _lib.lh_magnus_comm.restype = None
        _lib.lh_magnus_comm.argtypes = [C.POINTER(C.c_double)] * 6
        _lib.lh_coefficients_mixed.restype = C.c_int
        _lib.lh_coefficients_mixed.argtypes = [C.POINTER(abi.JetDesc), C.POINTER(abi.ImageDesc),
                                               C.POINTER(abi.SolverParams), C.POINTER(C.c_double), C.c_int64,
                                               C.POINTER(C.c_double), C.c_int32, C.POINTER(abi.GridDesc),
                                               C.POINTER(C.POINTER(C.c_double))]
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def observe(jet, image, params, pixels, grids=None, defer_rhs=0, w_piece=256.0, method=0, grid_walk=True):
    """Lane-by-lane run of the listed pixels; grids = [(GridDesc, values)] in
    handle order 1..n.  defer_rhs > 0 splits rays past that many right-hand
    sides into pieces of w_piece stiffness, integrated by Dormand-Prince
    (method 0) or exponential propagators (method 2).  Returns planes like
    oracle.observe plus 'n_pieces'."""
    pix = np.ascontiguousarray(pixels, dtype=np.int64)
    n = len(pix)
    res = {k: np.zeros(n) for k in ("tau", "I", "Q", "U", "V", "l")}
    for k in ("flags", "n_rhs", "n_steps"):
        res[k] = np.zeros(n, dtype=np.uint32)
    o = abi.Outputs()
    for k in ("tau", "I", "Q", "U", "V", "l"):
        setattr(o, k, _dp(res[k]))
    for k in ("flags", "n_rhs", "n_steps"):
        setattr(o, k, res[k].ctypes.data_as(C.POINTER(C.c_uint32)))
    npcs = np.zeros(n, dtype=np.int32)
    grids = grids or []
    gd = (abi.GridDesc * max(1, len(grids)))()
    vals = [np.ascontiguousarray(v, dtype=np.float64) for _, v in grids]
    gv = (C.POINTER(C.c_double) * max(1, len(grids)))()
    for i, (d, _) in enumerate(grids):
        gd[i] = d
        gv[i] = _dp(vals[i])
    rc = lib().lh_observe(C.byref(jet), C.byref(image), C.byref(params), pix.ctypes.data_as(C.POINTER(C.c_int64)),
                          n, len(grids), gd, gv, defer_rhs, w_piece, method, C.byref(o),
                          npcs.ctypes.data_as(C.POINTER(C.c_int32)), 1 if grid_walk else 0)
    if rc:
        raise RuntimeError("lh_observe failed: %d" % rc)
    res["n_pieces"] = npcs
    return res


def coefficients_mixed(jet, image, params, pts_cm, grids=None):
    """grids = [(GridDesc, values)] in handle order 1..n"""
    pts = np.ascontiguousarray(pts_cm, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(pts), 11))
    grids = grids or []
    gd = (abi.GridDesc * max(1, len(grids)))()
    vals = [np.ascontiguousarray(v, dtype=np.float64) for _, v in grids]
    gv = (C.POINTER(C.c_double) * max(1, len(grids)))()
    for i, (d, _) in enumerate(grids):
        gd[i] = d
        gv[i] = _dp(vals[i])
    rc = lib().lh_coefficients_mixed(C.byref(jet), C.byref(image), C.byref(params), _dp(pts), len(pts), _dp(out),
                                     len(grids), gd, gv)
    if rc:
        raise RuntimeError("lh_coefficients_mixed failed: %d" % rc)
    return out


def magnus_expphi(mu, eta, rho, v, mode=2):
    """(E, w, ok): exp(-mu 1 + G(eta, rho)) and phi1(-mu 1 + G) v; mode 1:
    the closed form for every generator, 2: the Taylor series where
    a^2 + b^2 <= 1 (what one lane of the kernels computes)."""
    E = np.zeros(16)
    w = np.zeros(4)
    ok = lib().lh_magnus_expphi(float(mu), _dp(np.ascontiguousarray(eta, float)), _dp(np.ascontiguousarray(rho, float)),
                                _dp(np.ascontiguousarray(v, float)), _dp(E), _dp(w), C.c_int32(mode))
    return E.reshape(4, 4).T, w, ok


def magnus_comm(e1, r1, e2, r2):
    eo, ro = np.zeros(3), np.zeros(3)
    a = [np.ascontiguousarray(x, float) for x in (e1, r1, e2, r2)]
    lib().lh_magnus_comm(*[_dp(x) for x in a], _dp(eo), _dp(ro))
    return eo, ro


def split_costs(cap=1 << 16):
    """(Dormand-Prince stiffness, exponential steps, rotation, optical depth)
    estimates of the rays the last observe() call split, in pixel order."""
    out = np.zeros(4 * cap)
    n = lib().lh_split_costs(_dp(out), cap)
    return out[:4 * n].reshape(n, 4)


def vis_sincos(ph):
    """The mixed visibility kernel's (sin, cos) of 2 pi ph, in FP32."""
    ph = np.ascontiguousarray(ph, dtype=np.float64)
    s = np.zeros(ph.size, dtype=np.float32)
    c = np.zeros(ph.size, dtype=np.float32)
    lib().lh_vis_sincos(_dp(ph), C.c_int64(ph.size), s.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p))
    return s, c


def vis_mixed(x, y, c, u, v, isign=-1):
    """The mixed visibility kernel's arithmetic for a real image, one lane per
    visibility, on the CPU."""
    a = [np.ascontiguousarray(t, dtype=np.float64).reshape(-1) for t in (x, y, c, u, v)]
    out = np.zeros(2 * a[3].size)
    lib().lh_vis_mixed(_dp(a[0]), _dp(a[1]), _dp(a[2]), C.c_int64(a[0].size), _dp(a[3]), _dp(a[4]),
                       C.c_int64(a[3].size), C.c_int32(isign), _dp(out))
    return out[0::2] + 1j * out[1::2]


def server_error_path(jet, image, params):
    """The server scheduler on the stand-in runtime with a forced HIP error:
    statuses of (a normal request, a request whose launch fails, a request
    submitted after the failure)."""
    rc = np.zeros(4, dtype=np.int32)
    lib().lh_server_error_path(C.byref(jet), C.byref(image), C.byref(params), rc.ctypes.data_as(C.POINTER(C.c_int32)))
    return rc.tolist()
