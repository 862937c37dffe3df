/*
 * jsx_kernels.h -- kernels for the FP64 path.
 *
 *   jsx_k_plain<NS>   : one thread per pixel, one pass per image (the plain
 *                       form of Observation::run, src/Observation.cpp:20-41).
 *   jsx_k_persist<NS> : persistent kernel; each lane that finishes (or misses)
 *                       takes the next ray from a device ray queue, so the
 *                       wavefront stays full while rays of very different
 *                       step counts drain.  Ray ids are handed out per
 *                       wavefront with one atomic per refill round.
 */
#ifndef JSX_KERNELS_H
#define JSX_KERNELS_H

#include "jsx_solver.h"

/* All per-launch constants live in device memory and are read through a
 * pointer that is re-laundered every loop trip: the compiler must then
 * re-fetch them with scalar loads instead of pinning ~150 dwords of model
 * constants in registers across the loop (which spills them into VGPR lanes
 * and caps occupancy at one wavefront per SIMD). */
struct DParams {
    DModel m;
    DImage im;
    DSolver S;
};

/* the launch-constant pointer, viewed in the constant address space so that
 * reads are scalar (s_load) and re-issued after every launder point */
#if defined(__HIP_DEVICE_COMPILE__)
#define JSX_CONST __attribute__((address_space(4)))
#define JSX_OPAQUE_S(x) asm volatile("" : "+s"(x))   /* value now unknown, kept in SGPRs */
#define JSX_OPAQUE_V(x) asm volatile("" : "+v"(x))   /* value now unknown, kept in VGPRs */
#else
#define JSX_CONST
#define JSX_OPAQUE_S(x) asm volatile("" : "+r"(x))
#define JSX_OPAQUE_V(x) asm volatile("" : "+r"(x))
#endif

__device__ __forceinline__ const JSX_CONST DParams *d_launder(const DParams *p) {
    const JSX_CONST DParams *q = (const JSX_CONST DParams *)p;
    JSX_OPAQUE_S(q);
    return q;
}

/* ray id -> pixel: 0 row-major, 1 column-major (apex columns first) */
__device__ inline int64_t d_ray_to_pixel(const DImage &im, int64_t id, int order) {
    if (order == 1) {
        int64_t col = id / im.na, row = id % im.na;
        return row * im.nl + col;
    }
    return id;
}

/* EXT: the model uses a field or geometry kind outside the workloads' set */
template <int NS, bool EXT>
__global__ __launch_bounds__(256) void jsx_k_plain(const DParams *__restrict__ P, DOut out, int64_t npix) {
    int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    Lane<NS> L;
    const JSX_CONST DParams *q = d_launder(P);
    lane_begin<NS, EXT>(L, q->m, q->im, q->S, pix);
    while (L.phase != PH_DONE) {
        q = d_launder(P);
        double r[NS];
        lane_rhs<NS, EXT>(L, q->m, q->im, q->S, r);
        q = d_launder(P);
        lane_consume(L, q->S, q->im, r);
    }
    lane_write(L, out, pix);
}

/* Observation::run_stripe (src/Observation.cpp:310-347, 58-124): optical
 * depth of the middle row's pixels n_along/2 .. n_along-1 with the fixed-step
 * integrate_tau -- classic RK4 over a const-step time range of dt = L/n (a step
 * is taken while (step+1) dt - L <= DBL_EPSILON; times are step * dt), stopping
 * at the first state with tau >= tau_max, whose value is added; later
 * segments are dropped.  One thread per stripe pixel. */
template <bool EXT>
__global__ __launch_bounds__(256) void jsx_k_tau_stripe(const DParams *__restrict__ P, double *tau, int64_t n_out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_out) return;
    const JSX_CONST DParams *c = d_launder(P);
    const int64_t pix = (c->im.na / 2) * c->im.nl + c->im.nl / 2 + q;
    double o[3];
    d_pixel_origin(c->im, pix, o);
    double dir[3] = {c->im.dir[0], c->im.dir[1], c->im.dir[2]};
    DSeg segs[2];
    int clipped = 0;
    const int ns = d_model_hit<EXT>(c->m, o, dir, segs, &clipped);
    const double n_los[3] = {-1. * dir[0], -1. * dir[1], -1. * dir[2]};
    double bg = 0.0;
    for (int si = 0; si < ns; ++si) {
        double pin[3], dv[3];
        for (int i = 0; i < 3; ++i) {
            pin[i] = si ? segs[1].pin[i] : segs[0].pin[i];
            dv[i] = (si ? segs[1].pout[i] : segs[0].pout[i]) - pin[i];
        }
        const double length = d_norm3(dv);
        const double dt = length / c->S.n;
        double x = 0.0, t = 0.0;
        long step = 0;
        int found = x >= c->S.tau_max;
        while (!found && (double)(step + 1) * dt + 0.0 - length <= JSX_DBL_EPS) {
            double k[4];
            const double ts[4] = {t, t + dt * 0.5, t + dt * 0.5, t + dt * 1.0};
            for (int s = 0; s < 4; ++s) {
                double p[3];
                for (int i = 0; i < 3; ++i) p[i] = pin[i] + ts[s] * dir[i];
                DCoef cf;
                c = d_launder(P);
                d_coefficients<0, EXT>(c->m, p, n_los, c->S.nu, cf);
                k[s] = cf.kI;
            }
            x = x + dt * (1.0 / 6.0) * k[0] + dt * (1.0 / 3.0) * k[1] + dt * (1.0 / 3.0) * k[2] +
                dt * (1.0 / 6.0) * k[3];
            step++;
            t = 0.0 + (double)step * dt;
            found = x >= c->S.tau_max;
        }
        bg += x;
        if (found) break;
    }
    tau[q] = bg;
}

/* Geometry::hit for explicit rays */
__global__ void jsx_k_hit(DModel m, const double *rays, int64_t n, int32_t *n_seg, double *segs) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double o[3] = {rays[6 * q], rays[6 * q + 1], rays[6 * q + 2]};
    double d[3] = {rays[6 * q + 3], rays[6 * q + 4], rays[6 * q + 5]};
    d_normalize3(d);
    DSeg s[2];
    int clipped;
    int ns = d_model_hit<true>(m, o, d, s, &clipped);
    n_seg[q] = ns;
    for (int k = 0; k < 3; ++k) {
        segs[12 * q + k] = ns > 0 ? s[0].pin[k] : 0.0;
        segs[12 * q + 3 + k] = ns > 0 ? s[0].pout[k] : 0.0;
        segs[12 * q + 6 + k] = ns > 1 ? s[1].pin[k] : 0.0;
        segs[12 * q + 9 + k] = ns > 1 ? s[1].pout[k] : 0.0;
    }
}

/* Field and geometry values at explicit points (the layout of
 * jsx_fields_host): every kind, FP64 */
__global__ void jsx_k_fields(DModel m, const double *pts, const double *nls, int64_t n, double nu, double *out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double p[3] = {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
    const double nl[3] = {nls[3 * q], nls[3 * q + 1], nls[3 * q + 2]};
    double *o = out + (int64_t)JSX_FIELDS_STRIDE * q;
    double v[3];
    d_vfield<true>(m, p, v);
    const double gamma = d_getG(v);
    double rr = 0.0, lh = 0.0, ll = 0.0, bl[3] = {0.0, 0.0, 0.0};
    if (d_need_r(m)) d_radius_log(p, &rr, &lh, &ll);
    const int vec = d_bfield<true>(m, p, rr, lh, ll, bl);
    const double n_lab = d_nfield<true>(m, p, rr, lh, ll);
    DFrame f;
    d_frame<true>(m, p, nl, nu, f);
    DCoef c;
    d_coefficients<1, true>(m, p, nl, nu, c);
    const double row[JSX_FIELDS_STRIDE] = {
        v[0], v[1], v[2], gamma,
        bl[0], vec ? bl[1] : 0.0, vec ? bl[2] : 0.0, (double)vec,
        f.b[0], vec ? f.b[1] : 0.0, vec ? f.b[2] : 0.0, n_lab, f.n,
        f.D, f.nl[0], f.nl[1], f.nl[2], f.nu,
        c.kI, c.kQ, c.kU, c.kV, c.kF, c.kC, c.hQ, c.eI, c.eQ, c.eU, c.eV,
        (double)d_is_within(m, p), 0.0, 0.0};
    for (int k = 0; k < JSX_FIELDS_STRIDE; ++k) o[k] = row[k];
}

#endif /* JSX_KERNELS_H */
