"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MATH_LIBM, MATH_JSX = 0, 1
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_mas_to_pc.restype = C.c_double
        L.orc_mas_to_pc.argtypes = [C.c_double]
        L.orc_grid_interp.restype = C.c_double
        L.orc_dopri5_check.restype = C.c_int64
        L.orc_dopri5_check.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p]
        L.orc_steps_schedule.restype = C.c_int32
        L.orc_steps_schedule.argtypes = [C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _grid_type():
    from jetshow_amd import abi

    class OrcGrid(C.Structure):
        _fields_ = [("desc", abi.GridDesc), ("values", C.c_void_p), ("pz", C.c_void_p), ("pr", C.c_void_p),
                    ("tri", C.c_void_p), ("boff", C.c_void_p), ("blist", C.c_void_p), ("nbz", C.c_int32),
                    ("nbr", C.c_int32), ("bz0", C.c_double), ("bdz", C.c_double), ("br0", C.c_double),
                    ("bdr", C.c_double), ("bzl", C.c_double), ("brl", C.c_double)]
    return OrcGrid


class TriGrid:
    """A triangulated point set for the oracle: node coordinates z, r
    [pc], node values, counter-clockwise triangles [m, 3] and nan_value, with
    a bucket index built here (its own layout: ~3 triangles per bucket; the
    lookup's rule -- lowest-numbered triangle holding the point -- does not
    depend on the layout)."""

    def __init__(self, z, r, values, tri, nan_value=0.0):
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        self.r = np.ascontiguousarray(r, dtype=np.float64)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
        self.nan_value = float(nan_value)
        T = self.tri
        tz, tr = self.z[T], self.r[T]
        self.bz0, self.bzl = float(tz.min()), float(tz.max())
        self.br0, self.brl = float(tr.min()), float(tr.max())
        wz, wr = max(self.bzl - self.bz0, 1e-300), max(self.brl - self.br0, 1e-300)
        nb = max(1.0, len(T) / 3.0)
        nbz = int(min(max(round(np.sqrt(nb * wz / wr)), 1), 1 << 15))
        nbr = int(min(max(round(nb / nbz), 1), 1 << 15))
        self.nbz, self.nbr = nbz, nbr
        self.bdz, self.bdr = wz / nbz, wr / nbr
        lo_z, hi_z, lo_r, hi_r = tz.min(1), tz.max(1), tr.min(1), tr.max(1)
        mz = 1e-9 * (hi_z - lo_z) + 1e-12 * wz
        mr = 1e-9 * (hi_r - lo_r) + 1e-12 * wr
        i0 = np.clip(np.floor((lo_z - mz - self.bz0) / self.bdz), 0, nbz - 1).astype(np.int64)
        i1 = np.clip(np.floor((hi_z + mz - self.bz0) / self.bdz), 0, nbz - 1).astype(np.int64)
        j0 = np.clip(np.floor((lo_r - mr - self.br0) / self.bdr), 0, nbr - 1).astype(np.int64)
        j1 = np.clip(np.floor((hi_r + mr - self.br0) / self.bdr), 0, nbr - 1).astype(np.int64)
        ni, nj = i1 - i0 + 1, j1 - j0 + 1
        cnt = ni * nj
        tid = np.repeat(np.arange(len(T), dtype=np.int64), cnt)
        k = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        bi = np.repeat(i0, cnt) + k // np.repeat(nj, cnt)
        bj = np.repeat(j0, cnt) + k % np.repeat(nj, cnt)
        bucket = bi * nbr + bj
        order = np.lexsort((tid, bucket))
        self.blist = np.ascontiguousarray(tid[order], dtype=np.int32)
        self.boff = np.zeros(nbz * nbr + 1, dtype=np.int32)
        np.add.at(self.boff, bucket + 1, 1)
        self.boff = np.ascontiguousarray(np.cumsum(self.boff), dtype=np.int32)


_ORC_GRID = None


def make_grids(grids):
    """grids: list of (abi.GridDesc, values ndarray) or TriGrid -> (ctypes array, keepalive)."""
    global _ORC_GRID
    if _ORC_GRID is None:
        _ORC_GRID = _grid_type()
    if not grids:
        return None, 0, []
    from jetshow_amd import abi
    arr = (_ORC_GRID * len(grids))()
    keep = []
    for i, g in enumerate(grids):
        if isinstance(g, TriGrid):
            d = abi.GridDesc()
            d.nz = d.nr = 0
            d.nan_value = g.nan_value
            keep.append(g)
            arr[i].desc = d
            arr[i].values = g.values.ctypes.data
            arr[i].pz, arr[i].pr = g.z.ctypes.data, g.r.ctypes.data
            arr[i].tri, arr[i].boff, arr[i].blist = g.tri.ctypes.data, g.boff.ctypes.data, g.blist.ctypes.data
            arr[i].nbz, arr[i].nbr = g.nbz, g.nbr
            arr[i].bz0, arr[i].bdz, arr[i].br0, arr[i].bdr, arr[i].bzl, arr[i].brl = (
                g.bz0, g.bdz, g.br0, g.bdr, g.bzl, g.brl)
            continue
        d, v = g
        v = np.ascontiguousarray(v, dtype=np.float64)
        keep.append(v)
        arr[i].desc = d
        arr[i].values = v.ctypes.data
    return arr, len(grids), keep


def observe(jet, image, params, pixel_idx=None, math=MATH_JSX, threads=0, grids=None):
    """Observation::run on the CPU for the listed pixel indices (all if None).
    Returns dict of 1-D arrays indexed like pixel_idx."""
    n_all = image.n_across * image.n_along
    if pixel_idx is None:
        idx = None
        n = n_all
    else:
        idx = np.ascontiguousarray(pixel_idx, dtype=np.int64)
        n = len(idx)
    res = {k: np.zeros(n) for k in ("tau", "I", "Q", "U", "V", "l")}
    for k in ("flags", "n_rhs", "n_steps"):
        res[k] = np.zeros(n, dtype=np.uint32)
    garr, ng, keep = make_grids(grids or [])
    rc = lib().orc_observe(C.byref(jet), C.byref(image), C.byref(params), garr, ng, _p(idx), C.c_int64(n),
                           _p(res["tau"]), _p(res["I"]), _p(res["Q"]), _p(res["U"]), _p(res["V"]), _p(res["l"]),
                           _p(res["flags"]), _p(res["n_rhs"]), _p(res["n_steps"]), int(math), int(threads))
    del keep
    if rc:
        raise RuntimeError("orc_observe failed: %d" % rc)
    return res


def hit(jet, rays, params=None, grids=None):
    """Geometry::hit of rays [n, 6]; grids (list of (GridDesc, values), handle
    h = grids[h - 1]) hold the hull planes of GEOM_CONVEX_HULL."""
    r = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
    n = r.shape[0]
    ns = np.zeros(n, dtype=np.int32)
    seg = np.zeros((n, 2, 6))
    arr, ng, keep = make_grids(grids)
    lib().orc_hit(C.byref(jet), C.byref(params) if params is not None else None, arr, ng, _p(r), C.c_int64(n),
                  _p(ns), _p(seg))
    return ns, seg


def tau_stripe(jet, image, params, math=MATH_JSX, grids=None):
    """Observation::run_stripe + getStripe("tau"): n_along - n_along/2 values."""
    out = np.zeros(image.n_along - image.n_along // 2)
    garr, ng, keep = make_grids(grids or [])
    rc = lib().orc_tau_stripe(C.byref(jet), C.byref(image), C.byref(params), garr, ng, int(math), _p(out))
    if rc:
        raise ValueError("orc_tau_stripe failed: %d" % rc)
    return out


def image(img):
    n = img.n_across * img.n_along
    sizes = np.zeros(n)
    coords = np.zeros((n, 3))
    rays = np.zeros((n, 6))
    rc = lib().orc_image(C.byref(img), _p(sizes), _p(coords), _p(rays))
    if rc:
        raise RuntimeError("orc_image failed: %d" % rc)
    return sizes, coords, rays


def coefficients(jet, nu, point, n_los, math=MATH_JSX, grids=None):
    out = np.zeros(11)
    garr, ng, keep = make_grids(grids or [])
    lib().orc_coefficients(C.byref(jet), garr, ng, C.c_double(nu), _p(np.asarray(point, float)),
                           _p(np.asarray(n_los, float)), int(math), _p(out))
    return out


def plasma_coefficients(jet, b, n_los_prime, nu_prime, n, math=MATH_LIBM):
    """src/utils.cpp:8-144 in the plasma frame.  b: 3-vector, or a scalar for
    the random-B overloads.  Returns k_i, eta_i, k_q, eta_q, k_v, eta_v, k_F_c,
    k_C_c, k_F_r, k_C_r, k_F, k_C (zeros past eta_i for a scalar b)."""
    out = np.zeros(12)
    bb = np.atleast_1d(np.asarray(b, float))
    vec = 1 if bb.size == 3 else 0
    bb = np.ascontiguousarray(np.resize(bb, 3))
    lib().orc_plasma_coefficients(C.byref(jet), _p(bb), vec, _p(np.asarray(n_los_prime, float)),
                                  C.c_double(nu_prime), C.c_double(n), int(math), _p(out))
    return out


def stokes_slab(coef, L, S0=(0.0, 0.0, 0.0, 0.0), abs_tol=1e-14, rel_tol=1e-12):
    """The full-Stokes system with constant coefficients (kI, kQ, kU, kV, kF,
    kC, hQ, etaI, etaQ, etaU, etaV) over [0, L] from S0; returns (IQUV, n_rhs)."""
    out = np.zeros(4)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    s0 = np.ascontiguousarray(S0, dtype=np.float64)
    n = lib().orc_stokes_slab(_p(c), C.c_double(L), _p(s0), C.c_double(abs_tol), C.c_double(rel_tol), _p(out))
    if n < 0:
        raise RuntimeError("orc_stokes_slab: step failure")
    return out, n


def sky_angle(jet, b, n_los_prime):
    out = np.zeros(2)
    lib().orc_sky_angle(C.byref(jet), _p(np.asarray(b, float)), _p(np.asarray(n_los_prime, float)), _p(out))
    return out


def fields(jet, nu, point, n_los, math=MATH_JSX, grids=None):
    out = np.zeros(14)
    garr, ng, keep = make_grids(grids or [])
    lib().orc_fields(C.byref(jet), garr, ng, C.c_double(nu), _p(np.asarray(point, float)),
                     _p(np.asarray(n_los, float)), int(math), _p(out))
    return out


FIELDS_STRIDE = 32


def point_values(jet, nu, points, n_los, opt=0, math=MATH_JSX, grids=None):
    """Field, frame, coefficient and is_within values at points [n, 3] (cm)
    with unit LOS vectors n_los [n, 3], in the layout of jsx_fields_host."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    nl = np.ascontiguousarray(n_los, dtype=np.float64).reshape(-1, 3)
    n = pts.shape[0]
    out = np.zeros((n, FIELDS_STRIDE))
    garr, ng, keep = make_grids(grids or [])
    lib().orc_point_values(C.byref(jet), garr, ng, C.c_uint32(opt), C.c_double(nu), _p(pts), _p(nl), C.c_int64(n),
                           int(math), _p(out))
    return out


def grid_interp(desc, values, point_cm, math=MATH_JSX):
    garr, ng, keep = make_grids([(desc, values)])
    return lib().orc_grid_interp(C.byref(garr[0]), _p(np.asarray(point_cm, float)), int(math))


def tri_grid_interp(grid, points_cm, math=MATH_JSX):
    """SimulationInterpolater::interpolated_value of a TriGrid at points [n, 3] (cm)."""
    garr, ng, keep = make_grids([grid])
    pts = np.ascontiguousarray(points_cm, dtype=np.float64).reshape(-1, 3)
    L = lib()
    return np.array([L.orc_grid_interp(C.byref(garr[0]), _p(q), int(math)) for q in pts])


def mas_to_pc(z):
    return lib().orc_mas_to_pc(float(z))


def steps_schedule(tau, n_min, n_max, tau_min=0.1, tau_max=100.0):
    return lib().orc_steps_schedule(float(tau), n_min, n_max, tau_min, tau_max)


def dopri5_check(problem, t_end=1.0, dt=0.1, abs_tol=1e-10, rel_tol=1e-10, theta=0.5):
    out = np.zeros(4)
    out[1] = theta
    nrhs = lib().orc_dopri5_check(problem, t_end, dt, abs_tol, rel_tol, _p(out))
    return out, nrhs


def jsx_math(fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), dtype=np.float64)
    pair = fn in (3, 4, 7)
    out = np.zeros(len(x) * (2 if pair else 1))
    lib().orc_jsx_math(int(fn), _p(x), _p(y), _p(out), C.c_int64(len(x)))
    return out.reshape(-1, 2) if pair else out
