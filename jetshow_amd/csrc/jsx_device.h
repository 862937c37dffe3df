/*
 * jsx_device.h -- per-point physics on gfx950: geometry hits, field models,
 * frame transforms and transfer coefficients in FP64.
 *
 * Reference: src/Cone.cpp:56-118, src/Parabaloid.cpp:39-60,
 * src/Cylinder.cpp:35-95, src/Geometry.cpp:28-89, src/BField.cpp,
 * src/NField.cpp, src/VField.cpp, src/Jet.cpp:15-46,218-237 (and the
 * commented Q/U/V/F/C getters at :114-214, :239-288), src/utils.cpp:8-173.
 *
 * Every expression keeps the reference's evaluation order (left to right;
 * 3-vector reductions as (x0 + x1) + x2) and uses jsx_math for the
 * point-dependent transcendental terms, so for the same inputs the FP64 path
 * reproduces the CPU restatement bit for bit.  Compile with
 * -ffp-contract=off.  Constants that depend only on model parameters
 * (gamma-function factors, cos of the cone angle, ...) arrive precomputed in
 * DModel; the host evaluates them with the C library exactly as the reference
 * does on every call.
 */
#ifndef JSX_DEVICE_H
#define JSX_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "jsx_math.h"
#include "../../include/jetshow_amd.h"

#define JSX_PC 3.0856775814671913e+18
#define JSX_ME 9.10938356e-28
#define JSX_QE 4.80320467299766e-10
#define JSX_CL 29979245800.0
#define JSX_PI 3.141592653589793238462643383279502884
#define JSX_DBL_EPS 2.220446049250313080847e-16

struct DGrid {
    const double *v;      /* device values [nz][nr]; triangulated: per node */
    int nz, nr;           /* nz == 0: a triangulated point set */
    double z0, dz, r0, dr, nan_value, zlast, rlast;   /* triangulated: the bucket grid */
    /* triangulated point set: node coordinates [pc], counter-clockwise
       triangles (3 node indices each), and per bucket (nbz x nbr over the
       nodes' bounding box, bucket (i, j) at i * nbr + j) the triangles that
       reach into it, in increasing order: blist[boff[b] .. boff[b + 1]) */
    const double *pz, *pr;
    const int32_t *tri, *boff, *blist;
    int nbz, nbr;
};

/* POD model: descriptor + host-evaluated parameter constants */
struct DModel {
    int geometry, bfield, nfield, vfield;
    int b_plasma, n_plasma, clip, opt;
    double origin[3], axis[3];
    double M[3][3];
    double cos_angle, geo_param, big_scale;
    double zmin, zmax;
    double b0, b_negexp, sin_pitch, cos_pitch;
    double n0, n_negexp, thermal_fraction;
    double vz_flat, gamma0, gamma1, gz0, gz1, v_origin[3];
    double f_ki_v, f_ki_s, f_ei_v, f_ei_s, r_kq, r_eq, f_kv, f_ev, f_kFr, kCr_pre, kCr_gm;
    double e_ki, e_ei, e_kv, e_kCr;
    DGrid gb, gb2, gn, gv;
    int has_gb, has_gb2, has_gn, has_gv;
    DGrid hull;          /* JSX_GEOM_CONVEX_HULL: face planes [nz][4] = (n, h) in cm */
    int has_hull;
    /* further reference classes: ParabaloidCone switch height (its cone's
     * origin, angle and matrix sit in origin / cos_angle / M); composite and
     * force-free B; composite N; sheared, sheath and parabolic V (for the
     * const-cone V fields v_origin / v_theta hold the cone part's values) */
    double geo_z0;
    double b_negexp2, b_z0, b_i0, b_i1, b_pitch, tan_pitch;
    double n_negexp2, n_z0;
    double v_r, v_theta, v_R0, v_Rz0, v_z0;
    int ext;             /* a kind outside the workloads' set: the mixed path takes
                            its fields from the FP64 field code */
};

/* ---------------- 3-vector helpers ---------------- */
__device__ inline double d_dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
__device__ inline double d_sqn3(const double *a) { return (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]; }
__device__ inline double d_norm3(const double *a) { return sqrt(d_sqn3(a)); }
__device__ inline void d_cross3(const double *a, const double *b, double *o) {
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}
__device__ inline void d_normalize3(double *a) {
    double z = d_sqn3(a);
    if (z > 0.0) { double r = sqrt(z); a[0] /= r; a[1] /= r; a[2] /= r; }
}
__device__ inline void d_ray_point(const double *o, const double *d, double t, double *p) {
    p[0] = o[0] + d[0] * t; p[1] = o[1] + d[1] * t; p[2] = o[2] + d[2] * t;
}

/* ---------------- geometry ---------------- */
struct DSeg { double pin[3], pout[3]; };

__device__ inline void d_full_infinite(const DModel &m, const double *o, const double *d, DSeg &s) {
    double diff[3] = {o[0] - m.origin[0], o[1] - m.origin[1], o[2] - m.origin[2]};
    double dd = d_dot3(diff, d);
    double ref[3];
    for (int i = 0; i < 3; ++i) ref[i] = diff[i] - dd * d[i];
    for (int i = 0; i < 3; ++i) { s.pin[i] = ref[i] - m.big_scale * d[i]; s.pout[i] = ref[i] + m.big_scale * d[i]; }
}

__device__ inline int d_cone_is_within(const DModel &m, const double *p) {
    double diff[3] = {p[0] - m.origin[0], p[1] - m.origin[1], p[2] - m.origin[2]};
    d_normalize3(diff);
    double cosp = d_dot3(m.axis, diff);
    return cosp > m.cos_angle || cosp < -m.cos_angle;
}

__device__ inline void d_half_infinite(const DModel &m, const double *d, const double *pt, DSeg &s) {
    double chk[3] = {pt[0] - d[0], pt[1] - d[1], pt[2] - d[2]};
    if (d_cone_is_within(m, chk)) {
        for (int i = 0; i < 3; ++i) { s.pin[i] = pt[i] - m.big_scale * d[i]; s.pout[i] = pt[i]; }
    } else {
        for (int i = 0; i < 3; ++i) { s.pin[i] = pt[i]; s.pout[i] = pt[i] + m.big_scale * d[i]; }
    }
}

__device__ inline double d_quad_form(const double (&M)[3][3], const double *d, const double *e) {
    double r[3];
    for (int j = 0; j < 3; ++j) r[j] = (d[0] * M[0][j] + d[1] * M[1][j]) + d[2] * M[2][j];
    return (r[0] * e[0] + r[1] * e[1]) + r[2] * e[2];
}

/* Cone::hit (src/Cone.cpp:56-118) */
__device__ inline int d_cone_hit(const DModel &m, const double *o, const double *d, DSeg *segs) {
    double delta[3] = {o[0] - m.origin[0], o[1] - m.origin[1], o[2] - m.origin[2]};
    double c2 = d_quad_form(m.M, d, d);
    double c1 = d_quad_form(m.M, d, delta);
    double c0 = d_quad_form(m.M, delta, delta);
    double disc = c1 * c1 - c0 * c2;
    if (disc < 0) return 0;
    if (disc == 0) {
        double t = -c1 / c2;
        double p[3];
        d_ray_point(o, d, t, p);
        double dp[3] = {p[0] - m.origin[0], p[1] - m.origin[1], p[2] - m.origin[2]};
        if (d_sqn3(dp) <= (1E-06 * 1E-06) * (1E-06 * 1E-06) && c1 == 0.0) {
            d_full_infinite(m, o, d, segs[0]);
            return 1;
        }
        for (int i = 0; i < 3; ++i) { segs[0].pin[i] = p[i]; segs[0].pout[i] = p[i]; }
        return 1;
    }
    double eps = (c1 > 0 ? 1 : -1);
    double t1 = (-c1 - eps * sqrt(c1 * c1 - c2 * c0)) / (c2);
    double t2 = c0 / (-c1 - eps * sqrt(c1 * c1 - c2 * c0));
    double pin[3], pout[3];
    d_ray_point(o, d, t1 < t2 ? t1 : t2, pin);
    d_ray_point(o, d, t1 < t2 ? t2 : t1, pout);
    double cos_ray_cone = d_dot3(d, m.axis);
    if (fabs(cos_ray_cone) != 1.) {
        for (int i = 0; i < 3; ++i) { segs[0].pin[i] = pin[i]; segs[0].pout[i] = pout[i]; }
        return 1;
    }
    d_half_infinite(m, d, pout, segs[0]);
    d_half_infinite(m, d, pin, segs[1]);
    return 2;
}

/* intersection() (src/Geometry.cpp:28-69) */
__device__ inline int d_quadric_roots(const double *R0, const double *Rd, double A, double B, double C, double D,
                                      double E, double F, double G, double H, double I, double J, double *t) {
    double x0 = R0[0], y0 = R0[1], z0 = R0[2];
    double xd = Rd[0], yd = Rd[1], zd = Rd[2];
    double Aq = A * xd * xd + B * yd * yd + C * zd * zd + D * xd * yd + E * xd * zd + F * yd * zd;
    double Bq = 2 * A * x0 * xd + 2 * B * y0 * yd + 2 * C * z0 * zd + D * (x0 * yd + y0 * xd) + E * x0 * zd +
                F * (y0 * zd + yd * z0) + G * xd + H * yd + I * zd;
    double Cq = A * x0 * x0 + B * y0 * y0 + C * z0 * z0 + D * x0 * y0 + E * x0 * z0 + F * y0 * z0 + G * x0 +
                H * y0 + I * z0 + J;
    double Dscr = Bq * Bq - 4 * Aq * Cq;
    if (Aq == 0) { t[0] = -Cq / Bq; return 1; }
    if (Dscr < 0) return 0;
    double t0 = (-Bq - sqrt(Dscr)) / (2 * Aq);
    if (t0 > 0) { t[0] = t0; return 1; }
    t[0] = t0;
    t[1] = (-Bq + sqrt(Dscr)) / (2 * Aq);
    return 2;
}

/* Parabaloid::hit (src/Parabaloid.cpp:39-60) */
__device__ inline int d_paraboloid_hit(const DModel &m, const double *o, const double *d, DSeg *segs) {
    double r0 = m.geo_param;
    double t[2];
    int nr = d_quadric_roots(o, d, 1., 1., 0., 0., 0., 0., 0., 0., -r0 * r0 / JSX_PC, 0., t);
    if (nr != 2) return 0;
    double t1 = t[0], t2 = t[1];
    d_ray_point(o, d, t1 < t2 ? t1 : t2, segs[0].pin);
    d_ray_point(o, d, t1 < t2 ? t2 : t1, segs[0].pout);
    return 1;
}

/* Cylinder::hit (src/Cylinder.cpp:35-95); r^2 unless JSX_OPT_CYLINDER_REF_BUG */
__device__ inline int d_cylinder_hit(const DModel &m, const double *o, const double *d, DSeg *segs) {
    double ro[3] = {o[0] / JSX_PC, o[1] / JSX_PC, o[2] / JSX_PC};
    double co[3] = {m.origin[0] / JSX_PC, m.origin[1] / JSX_PC, m.origin[2] / JSX_PC};
    const double *a = m.axis;
    double delta[3] = {ro[0] - co[0], ro[1] - co[1], ro[2] - co[2]};
    double da = d_dot3(d, a);
    double v[3] = {d[0] - da * a[0], d[1] - da * a[1], d[2] - da * a[2]};
    double c2 = d_sqn3(v);
    double dla = d_dot3(delta, a);
    double w[3] = {delta[0] - dla * a[0], delta[1] - dla * a[1], delta[2] - dla * a[2]};
    double c1 = d_dot3(v, w);
    double r = m.geo_param;
    double c0 = (m.opt & JSX_OPT_CYLINDER_REF_BUG) ? d_sqn3(w) - r / JSX_PC : d_sqn3(w) - (r / JSX_PC) * (r / JSX_PC);
    double disc = c1 * c1 - c0 * c2;
    if (disc < 0) return 0;
    if (disc > 0) {
        double eps = (c1 > 0 ? 1 : -1);
        double sqrt_d = sqrt(disc);
        double t1 = (-c1 - eps * sqrt_d) / (c2);
        double t2 = c0 / (-c1 - eps * sqrt_d);
        d_ray_point(o, d, JSX_PC * (t1 < t2 ? t1 : t2), segs[0].pin);
        d_ray_point(o, d, JSX_PC * (t1 < t2 ? t2 : t1), segs[0].pout);
        return 1;
    }
    if (c2 == 1 && c1 == 0 && c0 == 0) {
        double t = -c1 / c2;
        d_ray_point(o, d, t, segs[0].pin);
        d_ray_point(o, d, t, segs[0].pout);
        return 1;
    }
    if (c2 == 0 && c1 == 0 && c0 <= 0) { d_full_infinite(m, o, d, segs[0]); return 1; }
    return 0;
}

/* z clip (extension); returns 0 drop, 1 unchanged, 2 clipped */
__device__ inline int d_clip_seg(const DModel &m, DSeg &s) {
    double zA = s.pin[2], zB = s.pout[2];
    if (zA >= m.zmin && zA <= m.zmax && zB >= m.zmin && zB <= m.zmax) return 1;
    if (zB == zA) return 0;
    double u0 = (m.zmin - zA) / (zB - zA);
    double u1 = (m.zmax - zA) / (zB - zA);
    if (u0 > u1) { double tt = u0; u0 = u1; u1 = tt; }
    double lo = u0 > 0.0 ? u0 : 0.0;
    double hi = u1 < 1.0 ? u1 : 1.0;
    if (!(lo < hi)) return 0;
    double A[3], B[3];
    for (int i = 0; i < 3; ++i) {
        double df = s.pout[i] - s.pin[i];
        A[i] = lo == 0.0 ? s.pin[i] : s.pin[i] + lo * df;
        B[i] = hi == 1.0 ? s.pout[i] : s.pin[i] + hi * df;
    }
    for (int i = 0; i < 3; ++i) { s.pin[i] = A[i]; s.pout[i] = B[i]; }
    return 2;
}

/* SimulationGeometry::hit (src/Geometry.cpp:97-132) on the convex hull given
 * by its face planes n . p <= h: the line o + t d enters where the last plane
 * it crosses inwards is crossed (largest t over planes with n . d < 0) and
 * leaves at the first outward crossing (smallest t over n . d > 0); a line
 * parallel to a plane outside it misses.  One segment, entry to exit along d,
 * when entry < exit. */
__device__ inline int d_hull_hit(const DModel &m, const double *o, const double *d, DSeg *segs) {
    double t_in = -INFINITY, t_out = INFINITY;
    const double *pl = m.hull.v;
    for (int i = 0; i < m.hull.nz; ++i) {
        const double *q = pl + 4 * i;
        const double nd = d_dot3(q, d);
        const double gap = q[3] - d_dot3(q, o);
        if (nd == 0.0) {
            if (gap < 0.0) return 0;
        } else {
            const double t = gap / nd;
            if (nd < 0.0) t_in = t > t_in ? t : t_in;
            else t_out = t < t_out ? t : t_out;
        }
    }
    if (!(t_in < t_out) || !(t_in > -INFINITY) || !(t_out < INFINITY)) return 0;
    d_ray_point(o, d, t_in, segs[0].pin);
    d_ray_point(o, d, t_out, segs[0].pout);
    return 1;
}

/* ParabaloidCone::hit (src/ParabaloidCone.cpp:45-94): the paraboloid r0
 * (geo_param) below z0, the cone tangent to it there above (its origin and
 * angle in the model's cone fields).  A ray that misses the paraboloid keeps
 * the cone's result when the cone's first segment enters above z0; one that
 * hits it enters on the paraboloid below z0, else on the cone, and leaves on
 * the paraboloid below z0, else on the cone.  The paraboloid lies inside the
 * cone, so a paraboloid hit without a cone hit (the reference reads an empty
 * list there) can only come from rounding at a tangency: the paraboloid's own
 * points are kept. */
__device__ inline int d_paraboloid_cone_hit(const DModel &m, const double *o, const double *d, DSeg *segs) {
    DSeg cs[2], ps[1];
    int nc = d_cone_hit(m, o, d, cs);
    int np = d_paraboloid_hit(m, o, d, ps);
    if (nc == 0 && np == 0) return 0;
    if (np == 0) {
        if (cs[0].pin[2] > m.geo_z0) {
            segs[0] = cs[0];
            if (nc > 1) segs[1] = cs[1];
            return nc;
        }
        return 0;
    }
    const DSeg &c0 = nc > 0 ? cs[0] : ps[0];
    for (int i = 0; i < 3; ++i) {
        segs[0].pin[i] = ps[0].pin[2] > m.geo_z0 ? c0.pin[i] : ps[0].pin[i];
        segs[0].pout[i] = ps[0].pout[2] < m.geo_z0 ? ps[0].pout[i] : c0.pout[i];
    }
    return 1;
}

/* Geometry::is_within (Cone.cpp:48-53, Parabaloid.cpp:35-37 with its mixed
 * units, Cylinder.cpp:29-32, ParabaloidCone.cpp:36-43); the convex hull: on
 * the inner side of every face plane */
__device__ inline int d_is_within(const DModel &m, const double *p) {
    switch (m.geometry) {
    case JSX_GEOM_CONE: return d_cone_is_within(m, p);
    case JSX_GEOM_PARABOLOID: return sqrt(p[0] * p[0] + p[1] * p[1]) < m.geo_param * sqrt(p[2]);
    case JSX_GEOM_CYLINDER: {
        double w[3] = {p[0] - m.origin[0], p[1] - m.origin[1], p[2] - m.origin[2]}, cr[3];
        d_cross3(w, m.axis, cr);
        return d_norm3(cr) < m.geo_param;
    }
    case JSX_GEOM_PARABOLOID_CONE:
        if (p[2] < m.geo_z0) return sqrt(p[0] * p[0] + p[1] * p[1]) < m.geo_param * sqrt(p[2]);
        return d_cone_is_within(m, p);
    case JSX_GEOM_CONVEX_HULL:
        for (int i = 0; i < m.hull.nz; ++i) {
            const double *q = m.hull.v + 4 * i;
            if (d_dot3(q, p) > q[3]) return 0;
        }
        return 1;
    default: return 0;
    }
}

/* Jet::hit + clip; *clipped set if the clip changed anything.  The two
 * possible segments are handled with constant indices so they stay in
 * registers. */
template <bool EXT = false>
__device__ inline int d_model_hit(const DModel &m, const double *o, const double *d, DSeg *segs, int *clipped) {
    int n = 0;
    *clipped = 0;
    if (m.geometry == JSX_GEOM_CONE) n = d_cone_hit(m, o, d, segs);
    else if (m.geometry == JSX_GEOM_PARABOLOID) n = d_paraboloid_hit(m, o, d, segs);
    else if (m.geometry == JSX_GEOM_CYLINDER) n = d_cylinder_hit(m, o, d, segs);
    else if (m.geometry == JSX_GEOM_CONVEX_HULL) n = d_hull_hit(m, o, d, segs);
    else if constexpr (EXT) {
        if (m.geometry == JSX_GEOM_PARABOLOID_CONE) n = d_paraboloid_cone_hit(m, o, d, segs);
    }
    if (m.clip && n > 0) {
        DSeg a = segs[0], b = segs[1];
        int ra = d_clip_seg(m, a);
        int rb = n > 1 ? d_clip_seg(m, b) : 1;
        if (ra != 1 || rb != 1) *clipped = 1;
        if (n == 1) {
            segs[0] = a;
            n = ra ? 1 : 0;
        } else if (ra && rb) {
            segs[0] = a; segs[1] = b;
        } else if (ra) {
            segs[0] = a; n = 1;
        } else if (rb) {
            segs[0] = b; n = 1;
        } else {
            n = 0;
        }
    }
    return n;
}

/* ---------------- grid interpolation ---------------- */
__device__ inline double d_tri_area(double px, double py, double qx, double qy, double rx, double ry) {
    return ((qx - px) * (ry - py) - (rx - px) * (qy - py)) / 2;
}

/* SimulationInterpolater::interpolated_value on a regular (z, r_p) grid
 * (src/SimulationInterpolater.cpp:43-102): fixed (i,j)-(i+1,j+1) diagonal. */
/* SimulationInterpolater::interpolated_value (src/SimulationInterpolater.cpp:
 * 33-102) on a triangulated point set: the lowest-numbered triangle of the
 * query's bucket whose barycentric weights are all >= -1e-12 (a point on a
 * shared edge, where either triangle gives the value up to rounding, takes
 * the lower number), weighted as Triangle_coordinates_2 forms the weights;
 * outside every triangle (the reference's infinite face): nan_value, as for
 * a NaN result. */
__device__ inline double d_tri_interp(const DGrid &G, double z, double rp) {
    if (!(z >= G.z0 && z <= G.zlast && rp >= G.r0 && rp <= G.rlast)) return G.nan_value;
    int i = (int)floor((z - G.z0) / G.dz);
    int j = (int)floor((rp - G.r0) / G.dr);
    if (i > G.nbz - 1) i = G.nbz - 1;
    if (j > G.nbr - 1) j = G.nbr - 1;
    if (i < 0) i = 0;
    if (j < 0) j = 0;
    const int b = i * G.nbr + j;
    const int k1 = G.boff[b + 1];
    for (int k = G.boff[b]; k < k1; ++k) {
        const int32_t *t = G.tri + 3 * (int64_t)G.blist[k];
        const int a0 = t[0], a1 = t[1], a2 = t[2];
        const double v0x = G.pz[a0], v0y = G.pr[a0], v1x = G.pz[a1], v1y = G.pr[a1];
        const double v2x = G.pz[a2], v2y = G.pr[a2];
        double total = d_tri_area(v0x, v0y, v1x, v1y, v2x, v2y);
        double inv = 1.0 / total;
        double b0 = d_tri_area(v1x, v1y, v2x, v2y, z, rp) * inv;
        double b1 = d_tri_area(v2x, v2y, v0x, v0y, z, rp) * inv;
        double b2 = 1.0 - b0 - b1;
        if (b0 >= -1e-12 && b1 >= -1e-12 && b2 >= -1e-12) {
            double val = 0;
            val += b0 * G.v[a0];
            val += b1 * G.v[a1];
            val += b2 * G.v[a2];
            if (val != val) val = G.nan_value;
            return val;
        }
    }
    return G.nan_value;
}

/* EXT: the model may hold triangulated point sets (only the kernels built
 * for the further kinds carry their lookup) */
template <bool EXT = false>
__device__ inline double d_grid_interp(const DGrid &G, const double *p) {
    double x = p[0] / JSX_PC, y = p[1] / JSX_PC, z = p[2] / JSX_PC;
    double rp = sqrt(x * x + y * y);
    if constexpr (EXT) {
        if (G.nz == 0) return d_tri_interp(G, z, rp);
    }
    if (!(z >= G.z0 && z <= G.zlast && rp >= G.r0 && rp <= G.rlast)) return G.nan_value;
    int i = (int)floor((z - G.z0) / G.dz);
    int j = (int)floor((rp - G.r0) / G.dr);
    if (i > G.nz - 2) i = G.nz - 2;
    if (j > G.nr - 2) j = G.nr - 2;
    if (i < 0) i = 0;
    if (j < 0) j = 0;
    double za = G.z0 + (double)i * G.dz, zb = G.z0 + (double)(i + 1) * G.dz;
    double ra = G.r0 + (double)j * G.dr, rb = G.r0 + (double)(j + 1) * G.dr;
    const double *row0 = G.v + (int64_t)i * G.nr;
    const double *row1 = row0 + G.nr;
    double f00 = row0[j], f11 = row1[j + 1];
    double v1x, v1y, v2x, v2y, f1, f2;
    if ((rp - ra) * (zb - za) <= (z - za) * (rb - ra)) {
        v1x = zb; v1y = ra; f1 = row1[j];
        v2x = zb; v2y = rb; f2 = f11;
    } else {
        v1x = zb; v1y = rb; f1 = f11;
        v2x = za; v2y = rb; f2 = row0[j + 1];
    }
    double total = d_tri_area(za, ra, v1x, v1y, v2x, v2y);
    double inv = 1.0 / total;
    double b0 = d_tri_area(v1x, v1y, v2x, v2y, z, rp) * inv;
    double b1 = d_tri_area(v2x, v2y, za, ra, z, rp) * inv;
    double b2 = 1.0 - b0 - b1;
    double val = 0;
    val += b0 * f00;
    val += b1 * f1;
    val += b2 * f2;
    if (val != val) val = G.nan_value;
    return val;
}

/* pow(x, y) sharing a logarithm; same bits as jsx_pow(x, y) */
__device__ inline double d_pow_shared(double x, double lh, double ll, double y) {
    return jsx_pow_is_generic(x, y) ? jsx_pow_from_log(lh, ll, y) : jsx_pow(x, y);
}

/* ---------------- fields ---------------- */
/* flow along +z with Lorentz factor g (VField.cpp:28,40,57) */
__device__ inline void d_v_axial(double g, double *v) {
    v[0] = 0; v[1] = 0; v[2] = JSX_CL * sqrt(1. - 1. / (g * g));
}

/* radial flow away from -o, r = |p + o| (VField.cpp:68-71,85-86,105-106) */
__device__ inline void d_v_radial(const double *p, const double *o, double r, double g, double *v) {
    double v_r = JSX_CL * sqrt(1. - 1. / (g * g));
    v[0] = v_r * (p[0] + o[0]) / r;
    v[1] = v_r * (p[1] + o[1]) / r;
    v[2] = v_r * (p[2] + o[2]) / r;
}

/* ShearedCentralVField::vf (VField.cpp:78-87): the reference writes
 * acos(z + o_z / r); JSX_OPT_SHEARED_CENTRAL_REF_BUG keeps that, the default
 * takes the polar angle acos((z + o_z) / r) */
__device__ inline void d_v_sheared_central(const DModel &m, const double *p, double *v) {
    double q[3] = {p[0] + m.v_origin[0], p[1] + m.v_origin[1], p[2] + m.v_origin[2]};
    double r = d_norm3(q);
    double theta = (m.opt & JSX_OPT_SHEARED_CENTRAL_REF_BUG) ? jsx_acos(p[2] + m.v_origin[2] / r)
                                                            : jsx_acos((p[2] + m.v_origin[2]) / r);
    double g = m.gamma0 - (m.gamma0 - m.gamma1) * theta / m.v_theta;
    d_v_radial(p, m.v_origin, r, g, v);
}

/* the flow of the parabolic fields (VField.cpp:118-132): along the
 * paraboloid through p, r = hypot(x, y), R the radius of the enclosing
 * paraboloid at z */
__device__ inline void d_v_parabolic(const DModel &m, const double *p, double r, double R, double g, double *v) {
    double v_r = JSX_CL * sqrt(1. - 1. / (g * g));
    double r0 = m.v_Rz0 * r / R;
    double sa, ca;
    jsx_sincos_atan(2 * JSX_PC * r / r0 / r0, &sa, &ca);
    if (r == 0) {
        v[0] = 0; v[1] = 0; v[2] = v_r;
    } else {
        v[0] = v_r * ca * p[0] / r;
        v[1] = v_r * ca * p[1] / r;
        v[2] = v_r * sa;
    }
}

/* AccParabolicVField (VField.cpp:138-161) and ShearedAccParabolicVField
 * (VField.cpp:171-206); R0 = v_R0 */
__device__ inline void d_v_acc_parabolic(const DModel &m, const double *p, int sheared, double *v) {
    double r = sqrt(p[0] * p[0] + p[1] * p[1]);
    double R = sqrt(m.v_Rz0 * m.v_Rz0 * p[2] / JSX_PC);
    double g0 = m.gamma0;
    if (sheared) {
        double R0_border = sqrt(m.v_Rz0 * m.v_Rz0 * m.v_R0 / JSX_PC);
        double rR0 = R0_border * r / R;
        g0 = m.gamma0 - (m.gamma0 - m.gamma1) * jsx_pow(rR0 / R0_border, 0.25);
        if (g0 < 1) g0 = 1;
    }
    double g = 1 + (g0 - 1) / sqrt(m.v_R0) * sqrt(p[2]);
    d_v_parabolic(m, p, r, R, g, v);
}

/* the V fields beyond the workloads' set (VField.cpp:32-59,75-268) */
__device__ inline void d_vfield_ext(const DModel &m, const double *p, double *v) {
    switch (m.vfield) {
    case JSX_V_SHEARED_FLAT: {
        double r = sqrt(p[0] * p[0] + p[1] * p[1]);
        d_v_axial(m.gamma0 - (m.gamma0 - m.gamma1) * r / m.v_r, v);
        return;
    }
    case JSX_V_SHEATH_FLAT: {
        double r = sqrt(p[0] * p[0] + p[1] * p[1]);
        d_v_axial(r < m.v_r ? m.gamma0 : m.gamma1, v);
        return;
    }
    case JSX_V_SHEARED_CENTRAL: d_v_sheared_central(m, p, v); return;
    case JSX_V_SHEATH_CENTRAL: {
        double q[3] = {p[0] + m.v_origin[0], p[1] + m.v_origin[1], p[2] + m.v_origin[2]};
        double r = d_norm3(q);
        double theta = jsx_acos(p[2] / r);
        d_v_radial(p, m.v_origin, r, theta < m.v_theta ? m.gamma0 : m.gamma1, v);
        return;
    }
    case JSX_V_CONST_PARABOLIC: {
        double r = sqrt(p[0] * p[0] + p[1] * p[1]);
        d_v_parabolic(m, p, r, sqrt(m.v_Rz0 * m.v_Rz0 * p[2] / JSX_PC), m.gamma0, v);
        return;
    }
    case JSX_V_ACC_PARABOLIC: d_v_acc_parabolic(m, p, 0, v); return;
    case JSX_V_SHEARED_ACC_PARABOLIC: d_v_acc_parabolic(m, p, 1, v); return;
    case JSX_V_ACC_PARABOLIC_CONST_CONE:
        if (p[2] > m.v_z0) {
            double q[3] = {p[0] + m.v_origin[0], p[1] + m.v_origin[1], p[2] + m.v_origin[2]};
            d_v_radial(p, m.v_origin, d_norm3(q), m.gamma0, v);
        } else {
            d_v_acc_parabolic(m, p, 0, v);
        }
        return;
    case JSX_V_SHEARED_ACC_PARABOLIC_CONST_CONE:
        if (p[2] > m.v_z0) d_v_sheared_central(m, p, v);
        else d_v_acc_parabolic(m, p, 1, v);
        return;
    default:
        v[0] = v[1] = v[2] = 0;
    }
}

template <bool EXT = false>
__device__ inline void d_vfield(const DModel &m, const double *p, double *v) {
    switch (m.vfield) {
    case JSX_V_CONST_FLAT:
        v[0] = 0; v[1] = 0; v[2] = m.vz_flat;
        return;
    case JSX_V_CONST_CENTRAL: {
        double q[3] = {p[0] + m.v_origin[0], p[1] + m.v_origin[1], p[2] + m.v_origin[2]};
        double r = d_norm3(q);
        double v_r = JSX_CL * sqrt(1. - 1. / (m.gamma0 * m.gamma0));
        v[0] = v_r * (p[0] + m.v_origin[0]) / r;
        v[1] = v_r * (p[1] + m.v_origin[1]) / r;
        v[2] = v_r * (p[2] + m.v_origin[2]) / r;
        return;
    }
    case JSX_V_LINEAR_GAMMA_Z: {
        double z = p[2], g;
        if (z <= m.gz0) g = m.gamma0;
        else if (z >= m.gz1) g = m.gamma1;
        else g = m.gamma0 + (m.gamma1 - m.gamma0) * ((z - m.gz0) / (m.gz1 - m.gz0));
        v[0] = 0; v[1] = 0; v[2] = JSX_CL * sqrt(1. - 1. / (g * g));
        return;
    }
    case JSX_V_GRID: {
        double gam = m.has_gv ? d_grid_interp<EXT>(m.gv, p) : 1.0;
        v[0] = 0; v[1] = 0; v[2] = JSX_CL * sqrt(1. - 1. / (gam * gam));
        return;
    }
    default:
        if constexpr (EXT) d_vfield_ext(m, p, v);
        else v[0] = v[1] = v[2] = 0;
    }
}

__device__ inline double d_getG(const double *v) {
    double beta[3] = {v[0] / JSX_CL, v[1] / JSX_CL, v[2] / JSX_CL};
    return sqrt(1. / (1. - d_sqn3(beta)));
}

/* fields with a power law in |p| / pc share log(|p| / pc) */
__device__ __forceinline__ int d_need_r(const DModel &m) {
    return m.bfield == JSX_B_RANDOM_SCALAR || m.bfield == JSX_B_TOROIDAL_R || m.bfield == JSX_B_HELICAL_R ||
           m.bfield == JSX_B_CONST_CYLINDER || m.bfield == JSX_B_RADIAL_CONICAL || m.nfield == JSX_N_BK ||
           m.nfield == JSX_N_COMPOSITE_BK;
}
__device__ __forceinline__ void d_radius_log(const double *p, double *rr, double *lh, double *ll) {
    double r = d_norm3(p) / JSX_PC, h = 0.0, l = 0.0;
    if (r > 0.0 && r < 1.79769313486231570815e+308) h = jsx_log_dd(r, &l);
    *rr = r; *lh = h; *ll = l;
}

/* the further reference B fields (BField.cpp:53-146) */
__device__ inline int d_bfield_ext(const DModel &m, const double *p, double rr, double lr_h, double lr_l, double *bl) {
    switch (m.bfield) {
    case JSX_B_COMPOSITE_RANDOM_SCALAR_Z: {  /* BField.cpp:56-64 */
        double r = fabs(p[2]);
        bl[0] = m.b0 * jsx_pow(r / JSX_PC, r > m.b_z0 ? m.b_negexp2 : m.b_negexp);
        return 0;
    }
    case JSX_B_CONST_CYLINDER:                /* BField.cpp:70-73 */
        bl[0] = 0.0; bl[1] = 0.0; bl[2] = m.b0 * d_pow_shared(rr, lr_h, lr_l, m.b_negexp);
        return 1;
    case JSX_B_RADIAL_CONICAL: {              /* BField.cpp:88-93 */
        double r = d_norm3(p);
        double bb = m.b0 * d_pow_shared(rr, lr_h, lr_l, m.b_negexp);
        bl[0] = bb * p[0] / r; bl[1] = bb * p[1] / r; bl[2] = bb * p[2] / r;
        return 1;
    }
    case JSX_B_HELICAL_CYLINDER: {            /* BField.cpp:112-117 */
        double r = sqrt(p[0] * p[0] + p[1] * p[1]);
        bl[0] = m.b0 * jsx_tan(m.b_pitch * p[1] / r);
        bl[1] = -m.b0 * jsx_tan(m.b_pitch * p[0] / r);
        bl[2] = m.b0;
        return 1;
    }
    case JSX_B_SPIRAL_CONICAL: {              /* BField.cpp:123-131 */
        double z = p[2], x = p[0], y = p[1];
        double b_z = m.b0 / (z * z / (JSX_PC * JSX_PC));
        bl[0] = b_z * (x / z + y * m.tan_pitch / JSX_PC);
        bl[1] = b_z * (y / z - x * m.tan_pitch / JSX_PC);
        bl[2] = b_z;
        return 1;
    }
    case JSX_B_FORCE_FREE_CYLINDRICAL: {      /* BField.cpp:137-146 */
        double sf, cf;
        jsx_sincos_atan(p[1] / p[0], &sf, &cf);
        bl[0] = -m.b0 * m.b_i1 * sf;
        bl[1] = m.b0 * m.b_i1 * sf;
        bl[2] = m.b0 * m.b_i0;
        return 1;
    }
    default:
        bl[0] = 0;
        return 0;
    }
}

/* the B field's bf(point) (BField.cpp:35-146,196-207): returns 1 with the
 * vector in bl[0..2], or 0 with a tangled field's magnitude in bl[0] */
template <bool EXT = false>
__device__ __forceinline__ int d_bfield(const DModel &m, const double *p, double rr, double lr_h, double lr_l, double *bl) {
    switch (m.bfield) {
    case JSX_B_RANDOM_SCALAR:
        bl[0] = m.b0 * d_pow_shared(rr, lr_h, lr_l, m.b_negexp);
        return 0;
    case JSX_B_RANDOM_SCALAR_Z: {
        double r = fabs(p[2]);
        bl[0] = m.b0 * jsx_pow(r / JSX_PC, m.b_negexp);
        return 0;
    }
    case JSX_B_GRID_SCALAR:
        bl[0] = m.has_gb ? d_grid_interp<EXT>(m.gb, p) : 0.0;
        return 0;
    case JSX_B_TOROIDAL: {
        double sf, cf;
        jsx_sincos_atan(p[1] / p[0], &sf, &cf);
        double bb = m.b0 * jsx_pow(p[2] / JSX_PC, m.b_negexp);
        bl[0] = -sf * bb; bl[1] = cf * bb; bl[2] = 0;
        return 1;
    }
    case JSX_B_TOROIDAL_R:
    case JSX_B_HELICAL_R: {
        double bb = m.b0 * d_pow_shared(rr, lr_h, lr_l, m.b_negexp);
        double rho = sqrt(p[0] * p[0] + p[1] * p[1]);
        double ex = -p[1] / rho, ey = p[0] / rho;
        if (m.bfield == JSX_B_TOROIDAL_R) {
            bl[0] = bb * ex; bl[1] = bb * ey; bl[2] = 0;
        } else {
            double bphi = bb * m.sin_pitch;
            bl[0] = bphi * ex; bl[1] = bphi * ey; bl[2] = bb * m.cos_pitch;
        }
        return 1;
    }
    case JSX_B_CONST_CYLINDER_Z:
        bl[0] = 0.0; bl[1] = 0.0; bl[2] = m.b0 * jsx_pow(fabs(p[2]) / JSX_PC, m.b_negexp);
        return 1;
    case JSX_B_GRID_VECTOR: {
        double sf, cf;
        jsx_sincos_atan(p[1] / p[0], &sf, &cf);
        double vp = m.has_gb ? d_grid_interp<EXT>(m.gb, p) : 0.0;
        double vf = m.has_gb2 ? d_grid_interp<EXT>(m.gb2, p) : 0.0;
        bl[0] = -sf * vf; bl[1] = cf * vf; bl[2] = vp;
        return 1;
    }
    default:
        if constexpr (EXT) return d_bfield_ext(m, p, rr, lr_h, lr_l, bl);
        bl[0] = 0;
        return 0;
    }
}

/* the N field's nf(point) (NField.cpp:17-59) */
template <bool EXT = false>
__device__ __forceinline__ double d_nfield(const DModel &m, const double *p, double rr, double lr_h, double lr_l) {
    switch (m.nfield) {
    case JSX_N_CONST: return m.n0;
    case JSX_N_BK: return m.n0 * d_pow_shared(rr, lr_h, lr_l, m.n_negexp);
    case JSX_N_BK_Z: return m.n0 * jsx_pow(fabs(p[2]) / JSX_PC, m.n_negexp);
    case JSX_N_GRID: return m.has_gn ? d_grid_interp<EXT>(m.gn, p) : 0.0;
    default:
        if constexpr (EXT) {
            if (m.nfield == JSX_N_COMPOSITE_BK)   /* NField.cpp:45-53 */
                return m.n0 * d_pow_shared(rr, lr_h, lr_l, fabs(p[2]) > m.n_z0 ? m.n_negexp2 : m.n_negexp);
        }
        return 0;
    }
}

/* frame quantities of Jet::getKI (src/Jet.cpp:25-31) */
struct DFrame {
    double gamma, D, n, nu;
    double nl[3];
    double b[3];
    int vec;
};

template <bool EXT = false>
__device__ inline void d_frame(const DModel &m, const double *p, const double *n_los, double nu, DFrame &f) {
    double v[3];
    d_vfield<EXT>(m, p, v);
    double gamma = d_getG(v);
    f.gamma = gamma;
    double rr = 0.0, lr_h = 0.0, lr_l = 0.0;
    if (d_need_r(m)) d_radius_log(p, &rr, &lr_h, &lr_l);
    /* B field, plasma frame (BField.cpp:35-146,196-224) */
    double bl[3] = {0.0, 0.0, 0.0};
    f.vec = d_bfield<EXT>(m, p, rr, lr_h, lr_l, bl);
    if (f.vec && !m.b_plasma) { /* getBprime (utils.cpp:164-168) */
        double s1 = gamma / ((1. + gamma) * (JSX_CL * JSX_CL));
        double vb = d_dot3(v, bl);
        for (int i = 0; i < 3; ++i) f.b[i] = bl[i] / gamma + (s1 * v[i]) * vb;
    } else {
        f.b[0] = bl[0]; f.b[1] = bl[1]; f.b[2] = bl[2];
    }
    /* getD (utils.cpp:152-155) */
    double beta[3] = {v[0] / JSX_CL, v[1] / JSX_CL, v[2] / JSX_CL};
    double D = 1. / (gamma * (1. - d_dot3(beta, n_los)));
    f.D = D;
    /* NField (NField.cpp:7-59) */
    double n = d_nfield<EXT>(m, p, rr, lr_h, lr_l);
    f.n = m.n_plasma ? n : n / gamma;
    /* get_n_los_prime (utils.cpp:157-162) */
    double fct = (D + 1.) * (gamma / (gamma + 1.));
    for (int i = 0; i < 3; ++i) f.nl[i] = D * n_los[i] - fct * beta[i];
    f.nu = nu / D;
}

/* lab-frame transfer coefficients (sky Stokes frame for Q/U) */
struct DCoef { double kI, kQ, kU, kV, kF, kC, hQ, eI, eQ, eU, eV; };

template <int FULL, bool EXT = false>
__device__ inline void d_coefficients(const DModel &m, const double *p, const double *n_los, double nu, DCoef &o) {
    DFrame f;
    d_frame<EXT>(m, p, n_los, nu, f);
    double D = f.D, n = f.n, fnu = f.nu;
    double nup = sqrt(n * JSX_QE * JSX_QE / (JSX_PI * JSX_ME));
    if (!f.vec) {
        double bs = f.b[0];
        double nub = JSX_QE * bs / (2. * JSX_PI * JSX_ME * JSX_CL);
        double x = nub / fnu;
        double lh = 0.0, ll = 0.0;
        if (x > 0.0 && x < 1.79769313486231570815e+308) lh = jsx_log_dd(x, &ll);
        double k0 = JSX_PI * nup * nup * nub / (JSX_CL * fnu * fnu);
        double ki = k0 * d_pow_shared(x, lh, ll, m.e_ki) * m.f_ki_s;
        double e0 = JSX_PI * nup * nup * nub * JSX_ME / JSX_CL;
        double ei = e0 * d_pow_shared(x, lh, ll, m.e_ei) * m.f_ei_s;
        o.kI = ki / D;
        o.eI = ei * D * D;
        o.kQ = o.kU = o.kV = o.kF = o.kC = o.hQ = o.eQ = o.eU = o.eV = 0.0;
        return;
    }
    const double *b = f.b, *nl = f.nl;
    double cr[3];
    d_cross3(nl, b, cr);
    double nub = JSX_QE * d_norm3(cr) / (2. * JSX_PI * JSX_ME * JSX_CL);
    double x = nub / fnu;
    double lh = 0.0, ll = 0.0;
    if (x > 0.0 && x < 1.79769313486231570815e+308) lh = jsx_log_dd(x, &ll);
    double k0 = JSX_PI * nup * nup * nub / (JSX_CL * fnu * fnu);
    double ki = k0 * d_pow_shared(x, lh, ll, m.e_ki) * m.f_ki_v;
    double e0 = JSX_PI * nup * nup * nub * JSX_ME / JSX_CL;
    double ei = e0 * d_pow_shared(x, lh, ll, m.e_ei) * m.f_ei_v;
    o.kI = ki / D;
    o.eI = ei * D * D;
    if (!FULL) return;
    double kq = m.r_kq * ki;
    double eq = m.r_eq * ei;
    double bnorm = d_norm3(b);
    double cos_theta = d_dot3(b, nl) / bnorm;
    double nubv = JSX_QE * bnorm / (2. * JSX_PI * JSX_ME * JSX_CL);
    double k0v = JSX_PI * nup * nup * nubv / (JSX_CL * fnu * fnu);
    double kv = -k0v * cos_theta * d_pow_shared(x, lh, ll, m.e_kv) * m.f_kv;
    double e0v = JSX_PI * nup * nup * nubv * JSX_ME / JSX_CL;
    double ev = -e0v * cos_theta * d_pow_shared(x, lh, ll, m.e_ki) * m.f_ev;
    double kFc = 2. * JSX_PI * nup * nup * nubv * cos_theta / (JSX_CL * fnu * fnu);
    double kCc = -JSX_PI * nup * nup * nub * nub / (JSX_CL * fnu * fnu * fnu);
    double kF = m.f_kFr * kFc;
    double kC = m.kCr_pre * (m.kCr_gm - d_pow_shared(x, lh, ll, m.e_kCr)) * kCc;
    if (m.thermal_fraction != 0.0) {
        double nth = m.thermal_fraction * n;
        double nupt = sqrt(nth * JSX_QE * JSX_QE / (JSX_PI * JSX_ME));
        kF = kF + 2. * JSX_PI * nupt * nupt * nubv * cos_theta / (JSX_CL * fnu * fnu);
        kC = kC + -JSX_PI * nupt * nupt * nub * nub / (JSX_CL * fnu * fnu * fnu);
    }
    const double *a = m.axis;
    double an = d_dot3(a, nl);
    double e1[3] = {a[0] - an * nl[0], a[1] - an * nl[1], a[2] - an * nl[2]};
    double e2[3];
    d_cross3(nl, a, e2);
    double bx = d_dot3(b, e1), by = d_dot3(b, e2);
    double q2 = bx * bx + by * by;
    double c2 = 1.0, s2 = 0.0;
    if (q2 > 0.0) { c2 = -(bx * bx - by * by) / q2; s2 = -(2.0 * bx * by) / q2; }
    o.kQ = (kq * c2) / D;
    o.kU = (kq * s2) / D;
    o.kV = kv / D;
    o.kF = kF / D;
    o.kC = (kC * c2) / D;
    o.hQ = -(kC * s2) / D;
    o.eQ = (eq * c2) * D * D;
    o.eU = (eq * s2) * D * D;
    o.eV = ev * D * D;
}

#endif /* JSX_DEVICE_H */
