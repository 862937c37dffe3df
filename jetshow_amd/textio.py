"""The reference's text formats from Python (host only; the C ABI's
jsx_write_map / jsx_write_stripe / jsx_read_field_text / jsx_grid_from_points).

write_map(path, plane)      write_2dvector into an appended map file
                            (src/utils.cpp:211-224, src/main.cpp:1447-1462)
write_stripe(path, values)  write_vector (src/utils.cpp:227-236)
save_maps(dirname, planes)  the map_{tau,i,q,u,v,l}.txt set the Python drivers
                            load with numpy.loadtxt (abc_simulations.py:188-263)
read_map(path)              numpy.loadtxt of such a file, as those drivers do
read_field_text(path)       read_from_txt (src/utils.cpp:381-388): [n, 3] rows
                            (z [cm], r [cm], value)
grid_from_points(points)    create_triangulation (src/SimulationInterpolater.cpp:
                            105-117) for a complete regular grid: node values
                            [nz, nr] and (z0, dz, r0, dr) in parsec
load_field_grid(ctx, path)  the two above and Context.upload_grid: a grid
                            handle for the Simulation*Field descriptors
"""
import ctypes as C
import os

import numpy as np

from . import abi, host

MAP_NAMES = ("tau", "i", "q", "u", "v", "l")
_PLANE_OF = {"tau": "tau", "i": "I", "q": "Q", "u": "U", "v": "V", "l": "l"}


def _check(rc, what):
    if rc:
        raise host.JsxError(rc, what)


def write_map(path, plane, scale=1.0, append=True):
    """Rows of `plane` (2-D) as write_2dvector writes them: "%g " per value,
    a newline per row; appended by default, as the reference opens the file."""
    a = np.ascontiguousarray(plane, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("plane must be 2-D")
    _check(host.lib().jsx_write_map(os.fsencode(path), a.ctypes.data_as(C.c_void_p), C.c_int64(a.shape[0]),
                                    C.c_int64(a.shape[1]), C.c_double(scale), C.c_int32(1 if append else 0)),
           "cannot write %s" % path)


def write_stripe(path, values, scale=1.0, append=True):
    """write_vector: "%g " and a newline per value."""
    a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    _check(host.lib().jsx_write_stripe(os.fsencode(path), a.ctypes.data_as(C.c_void_p), C.c_int64(a.size),
                                       C.c_double(scale), C.c_int32(1 if append else 0)),
           "cannot write %s" % path)


def save_maps(dirname, planes, names=MAP_NAMES, suffix="", append=True):
    """map_<name><suffix>.txt for each plane the result holds (planes as
    returned by Context.observe: tau, I, Q, U, V, l [n_across, n_along])."""
    out = []
    for n in names:
        key = _PLANE_OF[n]
        if key in planes and planes[key] is not None:
            p = os.path.join(dirname, "map_%s%s.txt" % (n, suffix))
            write_map(p, planes[key], append=append)
            out.append(p)
    return out


def read_map(path):
    """A map file as the reference's Python drivers read it (numpy.loadtxt)."""
    return np.loadtxt(path, ndmin=2)


def read_field_text(path):
    """read_from_txt: [n, 3] rows (z [cm], r [cm], value) until the first row
    that does not parse."""
    L = host.lib()
    n = C.c_int64()
    _check(L.jsx_read_field_text(os.fsencode(path), None, C.c_int64(0), C.byref(n)), "cannot read %s" % path)
    out = np.zeros((n.value, 3))
    if n.value:
        _check(L.jsx_read_field_text(os.fsencode(path), out.ctypes.data_as(C.c_void_p), C.c_int64(n.value),
                                     C.byref(n)), "cannot read %s" % path)
    return out


def grid_from_points(points, nan_value=0.0):
    """Node values [nz, nr] and (z0, dz, r0, dr) in parsec of points (z [cm],
    r [cm], value) that form a complete regular grid; JsxError
    (JSX_EUNSUPPORTED) otherwise."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    g = abi.GridDesc()
    L = host.lib()
    _check(L.jsx_grid_from_points(p.ctypes.data_as(C.c_void_p), C.c_int64(p.shape[0]), C.c_double(nan_value),
                                  C.byref(g), None, C.c_int64(0)), "points are not a complete regular (z, r) grid")
    vals = np.zeros((g.nz, g.nr))
    _check(L.jsx_grid_from_points(p.ctypes.data_as(C.c_void_p), C.c_int64(p.shape[0]), C.c_double(nan_value),
                                  C.byref(g), vals.ctypes.data_as(C.c_void_p), C.c_int64(vals.size)),
           "points are not a complete regular (z, r) grid")
    return vals, (g.z0, g.dz, g.r0, g.dr)


def triangulate(z, r):
    """The Delaunay triangles [m, 3] (counter-clockwise point indices) that
    jsx_grid_upload_points builds for the points (z, r)."""
    zr = np.ascontiguousarray(np.stack([np.asarray(z, dtype=np.float64), np.asarray(r, dtype=np.float64)], 1))
    n = C.c_int64()
    L = host.lib()
    _check(L.jsx_triangulate(zr.ctypes.data_as(C.c_void_p), C.c_int64(zr.shape[0]), None, C.c_int64(0), C.byref(n)),
           "the points span no triangle")
    tri = np.zeros((n.value, 3), dtype=np.int32)
    _check(L.jsx_triangulate(zr.ctypes.data_as(C.c_void_p), C.c_int64(zr.shape[0]), tri.ctypes.data_as(C.c_void_p),
                             C.c_int64(n.value), C.byref(n)), "the points span no triangle")
    return tri


def load_field_grid(ctx, path, nan_value=0.0):
    """Upload a simulation field text file (nodes at any positions) to HBM;
    returns the grid handle."""
    return ctx.upload_points(read_field_text(path), nan_value)
