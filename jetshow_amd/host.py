"""Python host side of the solver: loads libjetshow_amd.so (the HIP build) and
wraps its C ABI with numpy buffers.

There is no CPU fallback: if the shared library is missing or no GPU is
visible, every compute call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjetshow_amd.so")
_lib = None

# exported symbols of include/jetshow_amd.h
ABI_SYMBOLS = [
    "jsx_ctx_create", "jsx_ctx_destroy", "jsx_last_error", "jsx_abi_version",
    "jsx_grid_upload", "jsx_grid_release", "jsx_alloc_outputs", "jsx_free_outputs",
    "jsx_observe", "jsx_observe_host", "jsx_hit_host", "jsx_image_pixel_sizes",
    "jsx_mas_to_pc", "jsx_observe_batch", "jsx_observe_batch_device", "jsx_ctx_set_option", "jsx_ctx_get_option",
    "jsx_server_create", "jsx_server_submit", "jsx_server_poll", "jsx_server_wait", "jsx_server_stats",
    "jsx_server_destroy", "jsx_tau_stripe", "jsx_visibilities", "jsx_visibilities_device",
    "jsx_write_map", "jsx_write_stripe", "jsx_read_field_text", "jsx_grid_from_points", "jsx_hull_upload",
    "jsx_grid_upload_points", "jsx_triangulate",
]


class JsxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (abi.STATUS_NAMES.get(code, str(code)), msg))
        self.code = code


def lib():
    """The loaded HIP library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is missing: build it with __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.jsx_last_error.restype = C.c_char_p
        L.jsx_mas_to_pc.restype = C.c_double
        L.jsx_mas_to_pc.argtypes = [C.c_double]
        L.jsx_sizeof.restype = C.c_size_t
        L.jsx_ctx_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]
        for name in ("jsx_ctx_destroy",):
            getattr(L, name).argtypes = [C.c_void_p]
        L.jsx_last_error.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def mas_to_pc(redshift):
    """Parsecs per milliarcsecond at `redshift` (src/utils.cpp:205-208)."""
    return lib().jsx_mas_to_pc(float(redshift))


def pixel_sizes(image):
    """Pixel sizes [n_across, n_along] in cm (Image::getPixelSizes)."""
    out = np.zeros((image.n_across, image.n_along))
    rc = lib().jsx_image_pixel_sizes(C.byref(image), out.ctypes.data_as(C.c_void_p))
    if rc:
        raise JsxError(rc, "bad image description")
    return out


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


PLANES = ("tau", "I", "Q", "U", "V", "l", "flags", "n_rhs", "n_steps")


def _host_planes(image, keep=PLANES):
    """Host output planes [n_across, n_along] (those named in `keep`; I always)
    and the jsx_outputs view of them (NULL for the others)."""
    shape = (image.n_across, image.n_along)
    res = {}
    o = abi.Outputs()
    for k in PLANES:
        if k != "I" and k not in keep:
            continue
        if k in ("flags", "n_rhs", "n_steps"):
            res[k] = np.zeros(shape, dtype=np.uint32)
            setattr(o, k, res[k].ctypes.data_as(C.POINTER(C.c_uint32)))
        else:
            res[k] = np.zeros(shape)
            setattr(o, k, res[k].ctypes.data_as(C.POINTER(C.c_double)))
    return res, o


class Context:
    """A solver context on one GPU (jsx_ctx)."""

    def __init__(self, device=0, hbm_budget=0):
        self._h = C.c_void_p()
        rc = lib().jsx_ctx_create(int(device), C.c_size_t(hbm_budget), C.byref(self._h))
        if rc:
            raise JsxError(rc, "jsx_ctx_create(device=%d) failed (no GPU visible?)" % device)

    def _check(self, rc):
        if rc:
            raise JsxError(rc, lib().jsx_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().jsx_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def upload_grid(self, values, z0, dz, r0, dr, nan_value=0.0):
        """Copy a [nz, nr] node grid (parsec coordinates) to HBM; returns its handle."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        g = abi.GridDesc()
        g.nz, g.nr = v.shape
        g.z0, g.dz, g.r0, g.dr, g.nan_value = z0, dz, r0, dr, nan_value
        h = C.c_int32()
        self._check(lib().jsx_grid_upload(self._h, C.byref(g), _ptr(v), C.byref(h)))
        return h.value

    def upload_points(self, points, nan_value=0.0):
        """Field nodes [n, 3] = (z [cm], r [cm], value) at any positions
        (create_triangulation + SimulationInterpolater): a complete regular
        grid keeps the regular layout, any other set is Delaunay-triangulated
        on the host and uploaded with its bucket index; returns the handle."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        h = C.c_int32()
        self._check(lib().jsx_grid_upload_points(self._h, _ptr(pts), C.c_int64(pts.shape[0]), C.c_double(nan_value),
                                                 C.byref(h)))
        return h.value

    def upload_hull(self, planes):
        """Face planes [n, 4] = (n_x, n_y, n_z, h) in cm of a convex polyhedron
        (n . p <= h inside) to HBM for GEOM_CONVEX_HULL; returns its handle
        (jet.geo_table).  Released with release_grid."""
        pl = np.ascontiguousarray(planes, dtype=np.float64).reshape(-1, 4)
        h = C.c_int32()
        self._check(lib().jsx_hull_upload(self._h, _ptr(pl), C.c_int32(pl.shape[0]), C.byref(h)))
        return h.value

    def release_grid(self, handle):
        self._check(lib().jsx_grid_release(self._h, C.c_int32(handle)))

    def observe(self, jet, image, params):
        """Run the hot path for one image; returns numpy planes [n_across, n_along]."""
        res, o = _host_planes(image)
        self._check(lib().jsx_observe_host(self._h, C.byref(jet), C.byref(image), C.byref(params), C.byref(o)))
        return res

    def set_option(self, name, value):
        """Launch-shape / scheduling knob of the batched solver (jsx_ctx_set_option)."""
        self._check(lib().jsx_ctx_set_option(self._h, name.encode(), C.c_double(value)))

    def get_option(self, name):
        v = C.c_double()
        self._check(lib().jsx_ctx_get_option(self._h, name.encode(), C.byref(v)))
        return v.value

    def observe_batch(self, items):
        """Many images through one persistent launch per (precision, stokes)
        group: items = [(jet, image, params)]; returns a list of plane dicts."""
        reqs = (abi.Request * len(items))()
        out = []
        for i, (jet, image, params) in enumerate(items):
            planes, o = _host_planes(image)
            reqs[i].jet, reqs[i].image, reqs[i].params, reqs[i].out_host = jet, image, params, o
            out.append(planes)
        self._check(lib().jsx_observe_batch(self._h, reqs, len(items)))
        return out

    def visibilities(self, x, y, c, u, v, isign=-1, precision=abi.PREC_FP64):
        """Exact direct-sum visibilities (jsx_visibilities):
        vis[k] = sum_j c[j] exp(sign(isign) 2 pi i (u[k] x[j] + v[k] y[j])),
        u, v in cycles per unit of x, y.  c real or complex; returns complex128
        [len(u)]."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        c = np.asarray(c).reshape(-1)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        if not (x.size == y.size == c.size) or u.size != v.size:
            raise ValueError("x, y, c and u, v must have matching lengths")
        cre = np.ascontiguousarray(c.real, dtype=np.float64)
        cim = np.ascontiguousarray(c.imag, dtype=np.float64) if np.iscomplexobj(c) else None
        out = np.zeros(u.size, dtype=np.complex128)
        self._check(lib().jsx_visibilities(self._h, _ptr(x), _ptr(y), _ptr(cre), _ptr(cim), C.c_int64(x.size),
                                           _ptr(u), _ptr(v), C.c_int64(u.size), C.c_int32(isign),
                                           C.c_int32(precision), _ptr(out)))
        return out

    def visibilities_device(self, x, y, c_re, c_im, n_src, u, v, n_vis, vis, isign=-1, precision=abi.PREC_FP64,
                            stream=None):
        """jsx_visibilities_device on device addresses (ints, e.g. a torch
        tensor's data_ptr()); asynchronous on `stream` (a raw hipStream_t or
        None for the context's stream)."""
        vp = lambda a: C.c_void_p(a) if a else None
        self._check(lib().jsx_visibilities_device(self._h, vp(x), vp(y), vp(c_re), vp(c_im), C.c_int64(n_src),
                                                  vp(u), vp(v), C.c_int64(n_vis), C.c_int32(isign),
                                                  C.c_int32(precision), vp(vis), vp(stream)))

    def visibilities_device_noalloc(self, x, y, c_re, c_im, n_src, u, v, n_vis, vis, isign=-1,
                                    precision=abi.PREC_FP64, stream=None):
        """jsx_visibilities_device_noalloc: the same as visibilities_device,
        with the partial sums in the library's fixed per-device array (no
        device allocation; calls overlapping on one GPU race)."""
        vp = lambda a: C.c_void_p(a) if a else None
        self._check(lib().jsx_visibilities_device_noalloc(
            self._h, vp(x), vp(y), vp(c_re), vp(c_im), C.c_int64(n_src), vp(u), vp(v), C.c_int64(n_vis),
            C.c_int32(isign), C.c_int32(precision), vp(vis), vp(stream)))

    def tau_stripe(self, jet, image, params):
        """Observation::run_stripe + getStripe("tau"): optical depth of the
        middle row's pixels n_along/2 .. n_along-1 (fixed-step RK4, params.n
        steps per segment)."""
        out = np.zeros(image.n_along - image.n_along // 2)
        self._check(lib().jsx_tau_stripe(self._h, C.byref(jet), C.byref(image), C.byref(params), _ptr(out)))
        return out

    def hit(self, jet, rays, params=None):
        """Geometry::hit for rays [n, 6]; returns (n_seg [n], segments [n, 2, 6])."""
        r = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        n = r.shape[0]
        ns = np.zeros(n, dtype=np.int32)
        seg = np.zeros((n, 2, 6))
        self._check(lib().jsx_hit_host(self._h, C.byref(jet), C.byref(params) if params is not None else None,
                                       _ptr(r), C.c_int64(n), _ptr(ns), _ptr(seg)))
        return ns, seg


    FIELDS_STRIDE = 32

    def fields(self, jet, points, n_los, nu, params=None):
        """Field, frame and coefficient values at points [n, 3] (cm) with unit
        LOS vectors n_los [n, 3]: [n, 32] in the layout of jsx_fields_host
        (v, Gamma, bf, vec, bf_plasma_frame, nf, nf_plasma_frame, D, n_los',
        nu', the 11 lab-frame coefficients, is_within)."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        nl = np.ascontiguousarray(n_los, dtype=np.float64).reshape(-1, 3)
        if nl.shape != pts.shape:
            raise ValueError("points and n_los must have the same shape")
        out = np.zeros((pts.shape[0], self.FIELDS_STRIDE))
        self._check(lib().jsx_fields_host(self._h, C.byref(jet), C.byref(params) if params is not None else None,
                                          _ptr(pts), _ptr(nl), C.c_int64(pts.shape[0]), C.c_double(nu), _ptr(out)))
        return out


class Server:
    """Continuous-batching request server on a context (jsx_server_*).

    submit() queues an image and returns its id; its planes are filled in
    host memory when wait(id) returns.  Grids may travel with the request
    (uploaded on admission, released when done)."""

    def __init__(self, ctx, precision=abi.PREC_MIXED, stokes=abi.STOKES_FULL, max_active=64, hbm_budget=0,
                 epoch_budget=0):
        self.ctx = ctx
        self._h = C.c_void_p()
        self._keep = {}
        ctx._check(lib().jsx_server_create(ctx.handle, precision, stokes, max_active, C.c_size_t(hbm_budget),
                                           epoch_budget, C.byref(self._h)))

    def submit(self, jet, image, params, grids=(), keep=PLANES):
        """grids: [(field, GridDesc, values)] with field in abi.FIELD_*;
        keep: the output planes to return (I always)."""
        planes, o = _host_planes(image, keep)
        r = abi.Request()
        r.jet, r.image, r.params, r.out_host = jet, image, params, o
        pl = (abi.GridPayload * max(1, len(grids)))()
        vals = []
        for i, (field, desc, values) in enumerate(grids):
            v = np.ascontiguousarray(values, dtype=np.float64)
            vals.append(v)
            pl[i].field, pl[i].desc = field, desc
            pl[i].values = v.ctypes.data_as(C.POINTER(C.c_double))
        rid = C.c_uint64()
        self.ctx._check(lib().jsx_server_submit(self._h, C.byref(r), len(grids), pl, C.byref(rid)))
        self._keep[rid.value] = (planes, vals, pl)
        return rid.value

    def poll(self, rid):
        done = C.c_int32()
        self.ctx._check(lib().jsx_server_poll(self._h, C.c_uint64(rid), C.byref(done), None))
        return bool(done.value)

    def wait(self, rid):
        """Blocks until the request is done; returns (planes, timing dict)."""
        t = abi.RequestTiming()
        rc = lib().jsx_server_wait(self._h, C.c_uint64(rid), C.byref(t))
        planes, _, _ = self._keep.pop(rid)
        if rc:
            raise JsxError(rc, "request %d failed: %s" % (rid, lib().jsx_last_error(self.ctx.handle).decode()))
        return planes, {"t_submit": t.t_submit, "t_admit": t.t_admit, "t_done": t.t_done,
                        "latency_s": t.t_done - t.t_submit, "n_pixels": t.n_pixels}

    def stats(self):
        ep = C.c_int64()
        used = C.c_size_t()
        lib().jsx_server_stats(self._h, C.byref(ep), C.byref(used))
        return {"epochs": ep.value, "hbm_used": used.value}

    def close(self):
        if self._h:
            lib().jsx_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
