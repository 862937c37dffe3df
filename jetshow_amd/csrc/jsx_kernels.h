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

template <int NS>
__global__ __launch_bounds__(256) void jsx_k_plain(const DParams *__restrict__ P, DOut out, int64_t npix) {
    int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    Lane<NS> L;
    const JSX_CONST DParams *q = d_launder(P);
    lane_begin(L, q->m, q->im, q->S, pix);
    while (L.phase != PH_DONE) {
        q = d_launder(P);
        double r[NS];
        lane_rhs(L, q->m, q->im, q->S, r);
        q = d_launder(P);
        lane_consume(L, q->S, q->im, r);
    }
    lane_write(L, out, pix);
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
    int ns = d_model_hit(m, o, d, s, &clipped);
    n_seg[q] = ns;
    for (int k = 0; k < 3; ++k) {
        segs[12 * q + k] = ns > 0 ? s[0].pin[k] : 0.0;
        segs[12 * q + 3 + k] = ns > 0 ? s[0].pout[k] : 0.0;
        segs[12 * q + 6 + k] = ns > 1 ? s[1].pin[k] : 0.0;
        segs[12 * q + 9 + k] = ns > 1 ? s[1].pout[k] : 0.0;
    }
}

#endif /* JSX_KERNELS_H */
