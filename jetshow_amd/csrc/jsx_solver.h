/*
 * jsx_solver.h -- the per-ray integration as a lane state machine (FP64).
 *
 * Reference: Observation::observe_single_pixel (src/Observation.cpp:349-395),
 * integrate_tau_adaptive (:126-214), integrate_i_adaptive (:242-274), the
 * commented full-Stokes pass (:277-308), the systems of src/System.cpp:15-69,
 * and Boost.odeint's runge_kutta_dopri5 / controlled_runge_kutta /
 * dense_output_runge_kutta / integrate_adaptive as those call sites use them.
 *
 * Shape for the GPU: one ray per lane, and every trip of the caller's loop
 * evaluates exactly one right-hand side for the lane (whatever pass and stage
 * it is in), then runs the small bookkeeping transition that decides where the
 * next evaluation is.  Lanes of a wavefront that are in different passes or
 * stages therefore still execute the expensive coefficient code together; the
 * divergent part is only the cheap bookkeeping.  No state array is indexed by a
 * run-time value, so everything stays in registers.
 */
#ifndef JSX_SOLVER_H
#define JSX_SOLVER_H

#include "jsx_device.h"

/* Dormand-Prince 5(4) tableau as odeint spells it */
#define DP_A2 (1.0 / 5.0)
#define DP_A3 (3.0 / 10.0)
#define DP_A4 (4.0 / 5.0)
#define DP_A5 (8.0 / 9.0)
#define DP_B21 (1.0 / 5.0)
#define DP_B31 (3.0 / 40.0)
#define DP_B32 (9.0 / 40.0)
#define DP_B41 (44.0 / 45.0)
#define DP_B42 (-56.0 / 15.0)
#define DP_B43 (32.0 / 9.0)
#define DP_B51 (19372.0 / 6561.0)
#define DP_B52 (-25360.0 / 2187.0)
#define DP_B53 (64448.0 / 6561.0)
#define DP_B54 (-212.0 / 729.0)
#define DP_B61 (9017.0 / 3168.0)
#define DP_B62 (-355.0 / 33.0)
#define DP_B63 (46732.0 / 5247.0)
#define DP_B64 (49.0 / 176.0)
#define DP_B65 (-5103.0 / 18656.0)
#define DP_C1 (35.0 / 384.0)
#define DP_C3 (500.0 / 1113.0)
#define DP_C4 (125.0 / 192.0)
#define DP_C5 (-2187.0 / 6784.0)
#define DP_C6 (11.0 / 84.0)
#define DP_DC1 (DP_C1 - 5179.0 / 57600.0)
#define DP_DC3 (DP_C3 - 7571.0 / 16695.0)
#define DP_DC4 (DP_C4 - 393.0 / 640.0)
#define DP_DC5 (DP_C5 - -92097.0 / 339200.0)
#define DP_DC6 (DP_C6 - 187.0 / 2100.0)
#define DP_DC7 (-1.0 / 40.0)

#define JSX_BISECT_CAP 200

/* solver constants, host-filled */
struct DSolver {
    double nu, tau_max, tau_min, dt_max;
    double tau_abs, tau_rel, st_abs, st_rel, tau_prec;
    double tau_n_min, tau_n_max, sched_b, sched_k;
    double emin;                 /* pow(5, -5) */
    int n, n_max, max_fail, stokes;
};

/* image: along centres [nl], across half table [na/2][nl] */
struct DImage {
    const double *along;
    const double *across_half;
    int64_t na, nl;
    double dir[3], d2[3], zscale;
};

struct DOut {
    double *tau, *I, *Q, *U, *V, *l;
    uint32_t *flags, *n_rhs, *n_steps;
};

enum { PH_TAU = 0, PH_STOKES = 1, PH_DONE = 2 };
enum { SUB_INIT = 0 };

template <int NS>
struct Lane {
    int64_t pix;
    double s0in[3], s0out[3], s1in[3], s1out[3];
    int nseg, seg, phase, sub, tries, deriv_ok, first;
    double L, t, dt, t_old, tq;
    double x[NS], xt[NS];    /* xt: stage input; after stage 7 it holds the new state */
    double k1[NS], k2[NS], k3[NS], k4[NS], k5[NS], k6[NS], k7[NS];
    double x_old, k_old;
    double tau_bg, thick;
    uint32_t flags, n_rhs, n_steps;
};

__device__ inline int d_less_with_sign(double t1, double t2, double dt) {
    if (dt > 0) return t2 - t1 > JSX_DBL_EPS;
    return t1 - t2 > JSX_DBL_EPS;
}

__device__ inline int d_steps_schedule(const DSolver &S, double tau) {
    if (tau <= S.tau_n_min) return S.n;
    if (tau >= S.tau_n_max) return S.n_max;
    double n = S.sched_k * jsx_pow((jsx_log(tau) + S.sched_b), 2.0) + S.n;
    return (int)ceil(n);
}

/* pixel (row i, column j) -> ray origin (ImagePlane.cpp:15-30) */
__device__ inline void d_pixel_origin(const DImage &im, int64_t pix, double *o) {
    int64_t i = pix / im.nl, j = pix % im.nl;
    int64_t h = im.na / 2;
    double across = i < h ? im.across_half[(h - 1 - i) * im.nl + j] : -im.across_half[(i - h) * im.nl + j];
    double along = im.along[j];
    o[0] = 1. * 0.0;
    o[1] = 1. * across;
    o[2] = im.zscale * along;
}

/* endpoints of the current segment */
template <int NS>
__device__ inline double lane_cin(const Lane<NS> &L, int i) { return L.seg ? L.s1in[i] : L.s0in[i]; }
template <int NS>
__device__ inline double lane_cout(const Lane<NS> &L, int i) { return L.seg ? L.s1out[i] : L.s0out[i]; }

template <int NS>
__device__ inline void lane_select_seg(Lane<NS> &L, int s) { L.seg = s; }

template <int NS>
__device__ inline double lane_seg_length(const Lane<NS> &L) {
    double dv[3] = {lane_cout(L, 0) - lane_cin(L, 0), lane_cout(L, 1) - lane_cin(L, 1), lane_cout(L, 2) - lane_cin(L, 2)};
    return d_norm3(dv);
}

/* stage inputs of one Dormand-Prince try (odeint's scale_sum order) */
template <int NS>
__device__ inline void lane_prepare_stage(Lane<NS> &L, int s) {
    double dt = L.dt, t = L.t;
    switch (s) {
    case 2:
        for (int i = 0; i < NS; ++i) L.xt[i] = L.x[i] + dt * DP_B21 * L.k1[i];
        L.tq = t + dt * DP_A2;
        break;
    case 3:
        for (int i = 0; i < NS; ++i) L.xt[i] = L.x[i] + dt * DP_B31 * L.k1[i] + dt * DP_B32 * L.k2[i];
        L.tq = t + dt * DP_A3;
        break;
    case 4:
        for (int i = 0; i < NS; ++i)
            L.xt[i] = L.x[i] + dt * DP_B41 * L.k1[i] + dt * DP_B42 * L.k2[i] + dt * DP_B43 * L.k3[i];
        L.tq = t + dt * DP_A4;
        break;
    case 5:
        for (int i = 0; i < NS; ++i)
            L.xt[i] = L.x[i] + dt * DP_B51 * L.k1[i] + dt * DP_B52 * L.k2[i] + dt * DP_B53 * L.k3[i] +
                      dt * DP_B54 * L.k4[i];
        L.tq = t + dt * DP_A5;
        break;
    case 6:
        for (int i = 0; i < NS; ++i)
            L.xt[i] = L.x[i] + dt * DP_B61 * L.k1[i] + dt * DP_B62 * L.k2[i] + dt * DP_B63 * L.k3[i] +
                      dt * DP_B64 * L.k4[i] + dt * DP_B65 * L.k5[i];
        L.tq = t + dt;
        break;
    default: /* 7: the new state, derivative evaluated there */
        for (int i = 0; i < NS; ++i)
            L.xt[i] = L.x[i] + dt * DP_C1 * L.k1[i] + dt * DP_C3 * L.k3[i] + dt * DP_C4 * L.k4[i] +
                      dt * DP_C5 * L.k5[i] + dt * DP_C6 * L.k6[i];
        L.tq = t + dt;
        break;
    }
    L.sub = s;
}

/* Hairer's continuous extension (odeint runge_kutta_dopri5::calc_state) */
template <int NS>
__device__ inline double lane_dense(const Lane<NS> &L, double t) {
    double dt = (L.t - L.t_old);
    double theta = (t - L.t_old) / dt;
    double X1 = 5.0 * (2558722523.0 - 31403016.0 * theta) / 11282082432.0;
    double X3 = 100.0 * (882725551.0 - 15701508.0 * theta) / 32700410799.0;
    double X4 = 25.0 * (443332067.0 - 31403016.0 * theta) / 1880347072.0;
    double X5 = 32805.0 * (23143187.0 - 3489224.0 * theta) / 199316789632.0;
    double X6 = 55.0 * (29972135.0 - 7076736.0 * theta) / 822651844.0;
    double X7 = 10.0 * (7414447.0 - 829305.0 * theta) / 29380423.0;
    double theta_m_1 = theta - 1.0;
    double theta_sq = theta * theta;
    double A = theta_sq * (3.0 - 2.0 * theta);
    double B = theta_sq * theta_m_1;
    double C = theta_sq * theta_m_1 * theta_m_1;
    double Dd = theta * theta_m_1 * theta_m_1;
    double b1 = A * DP_C1 - C * X1 + Dd;
    double b3 = A * DP_C3 + C * X3;
    double b4 = A * DP_C4 - C * X4;
    double b5 = A * DP_C5 + C * X5;
    double b6 = A * DP_C6 - C * X6;
    double b7 = B + C * X7;
    return L.x_old + dt * b1 * L.k_old + dt * b3 * L.k3[0] + dt * b4 * L.k4[0] + dt * b5 * L.k5[0] +
           dt * b6 * L.k6[0] + dt * b7 * L.k1[0];
}

/* ---- transitions: each returns when the lane needs a RHS or is done ---- */

/* The transitions are loops, not mutual recursion: a recursive call chain
 * cannot be inlined, and a lane passed to an outlined function lives in
 * scratch memory for the whole kernel. */
template <int NS>
__device__ inline void lane_finish(Lane<NS> &L) {
    L.phase = PH_DONE;
}

template <int NS>
__device__ inline void lane_try_begin(Lane<NS> &L, const DSolver &S) {
    for (;;) {
        L.n_steps++;
        if (L.phase == PH_TAU && S.dt_max != 0.0 && !(L.dt - S.dt_max <= JSX_DBL_EPS)) {
            L.dt = S.dt_max;
            if (L.tries++ >= S.max_fail) {
                L.flags |= JSX_PIX_FAIL_MAXTRIALS;
                L.tau_bg = L.tau_bg + L.x[0];
                L.thick = L.thick + L.L;
                for (int i = 0; i < NS; ++i) L.x[i] = 0.0;
                lane_finish(L);
                return;
            }
            continue;
        }
        lane_prepare_stage(L, 2);
        return;
    }
}

template <int NS>
__device__ inline void lane_stokes_seg_setup(Lane<NS> &L, const DSolver &S) {
    lane_select_seg(L, L.seg);
    L.L = lane_seg_length(L);
    double auto_n = S.n;
    if (L.tau_bg > S.tau_n_min) auto_n = d_steps_schedule(S, L.tau_bg);
    L.dt = L.L / auto_n;
    L.t = 0.0;
    L.first = 1;
}

template <int NS>
__device__ inline void lane_stokes_loop(Lane<NS> &L, const DSolver &S) {
    for (;;) {
        if (!d_less_with_sign(L.t, L.L, L.dt)) {
            /* segment finished: move to the nearer one */
            L.seg = L.seg - 1;
            if (L.seg < 0) { lane_finish(L); return; }
            lane_stokes_seg_setup(L, S);
            continue;
        }
        if (d_less_with_sign(L.L, L.t + L.dt, L.dt)) L.dt = L.L - L.t;
        L.tries = 0;
        if (L.first) {
            for (int i = 0; i < NS; ++i) L.xt[i] = L.x[i];
            L.tq = L.t;
            L.sub = SUB_INIT;
            L.first = 0;
            return;
        }
        lane_try_begin(L, S);
        return;
    }
}

template <int NS>
__device__ inline void lane_stokes_seg_begin(Lane<NS> &L, const DSolver &S) {
    lane_stokes_seg_setup(L, S);
    lane_stokes_loop(L, S);
}

template <int NS>
__device__ inline void lane_stokes_begin(Lane<NS> &L, const DSolver &S) {
    if (!(L.tau_bg > S.tau_min)) {
        L.flags |= JSX_PIX_I_SKIPPED;
        for (int i = 0; i < NS; ++i) L.x[i] = 0.0;
        lane_finish(L);
        return;
    }
    L.phase = PH_STOKES;
    for (int i = 0; i < NS; ++i) L.x[i] = 0.0;
    L.seg = L.nseg - 1;
    lane_stokes_seg_begin(L, S);
}

/* the next tau segment: reset the range (tau restarts at 0 per segment) */
template <int NS>
__device__ inline void lane_tau_next_seg(Lane<NS> &L, const DSolver &S) {
    L.seg = L.seg + 1;
    lane_select_seg(L, L.seg);
    L.L = lane_seg_length(L);
    L.x[0] = 0.0; L.t = 0.0; L.dt = L.L / S.n; L.deriv_ok = 0;
}

template <int NS>
__device__ inline void lane_tau_range_next(Lane<NS> &L, const DSolver &S) {
    for (;;) {
        if (!d_less_with_sign(L.t, L.L, L.dt)) {
            /* segment ended below tau_max */
            L.tau_bg = L.tau_bg + L.x[0];
            L.thick = L.thick + L.L;
            if (L.seg + 1 >= L.nseg) { lane_stokes_begin(L, S); return; }
            lane_tau_next_seg(L, S);
            continue;
        }
        if (d_less_with_sign(L.L, L.t + L.dt, L.dt)) { L.dt = L.L - L.t; L.deriv_ok = 0; }
        L.t_old = L.t;
        L.tries = 0;
        if (!L.deriv_ok) {
            L.xt[0] = L.x[0];
            L.tq = L.t;
            L.sub = SUB_INIT;
            return;
        }
        lane_try_begin(L, S);
        return;
    }
}

template <int NS>
__device__ inline void lane_tau_seg_end(Lane<NS> &L, const DSolver &S, int found) {
    L.tau_bg = L.tau_bg + L.x[0];
    L.thick = L.thick + L.L;
    if (found || L.seg + 1 >= L.nseg) {
        if (found) L.nseg = L.seg + 1;
        lane_stokes_begin(L, S);
        return;
    }
    lane_tau_next_seg(L, S);
    lane_tau_range_next(L, S);
}

/* begin a ray: generate it, hit the geometry, set up the tau pass */
template <int NS, bool EXT = false>
__device__ inline void lane_begin(Lane<NS> &L, const DModel &m, const DImage &im, const DSolver &S, int64_t pix) {
    L.pix = pix;
    L.flags = 0; L.n_rhs = 0; L.n_steps = 0;
    L.tau_bg = 0.0; L.thick = 0.0;
    L.seg = 0;
    for (int i = 0; i < NS; ++i) L.x[i] = 0.0;
    double o[3];
    d_pixel_origin(im, pix, o);
    DSeg segs[2];
    int clipped = 0;
    int ns = d_model_hit<EXT>(m, o, im.dir, segs, &clipped);
    if (clipped) L.flags |= JSX_PIX_CLIPPED;
    L.nseg = ns;
    if (ns == 0) { L.phase = PH_DONE; return; }
    L.flags |= JSX_PIX_HIT;
    if (ns == 2) L.flags |= JSX_PIX_NSEG2;
    for (int i = 0; i < 3; ++i) {
        L.s0in[i] = segs[0].pin[i]; L.s0out[i] = segs[0].pout[i];
        L.s1in[i] = ns == 2 ? segs[1].pin[i] : 0.0;
        L.s1out[i] = ns == 2 ? segs[1].pout[i] : 0.0;
    }
    L.phase = PH_TAU;
    L.seg = 0;
    lane_select_seg(L, 0);
    L.L = lane_seg_length(L);
    L.x[0] = 0.0; L.t = 0.0; L.dt = L.L / S.n; L.deriv_ok = 0;
    if (L.x[0] >= S.tau_max) { /* degenerate tau_max <= 0 */
        L.t_old = 0.0;
        lane_tau_seg_end(L, S, 1);
        return;
    }
    lane_tau_range_next(L, S);
}

/* the point at which the pending right-hand side is evaluated */
template <int NS>
__device__ inline void lane_point(const Lane<NS> &L, const DImage &im, double *p) {
    if (L.phase == PH_TAU) {
        for (int i = 0; i < 3; ++i) p[i] = lane_cin(L, i) + L.tq * im.dir[i];
    } else {
        for (int i = 0; i < 3; ++i) p[i] = lane_cout(L, i) - L.tq * im.dir[i];
    }
}

/* consume the right-hand side r evaluated at (xt, tq) */
template <int NS>
__device__ inline void lane_consume(Lane<NS> &L, const DSolver &S, const DImage &im, const double *r) {
    L.n_rhs++;
    int ns_live = L.phase == PH_TAU ? 1 : NS;
    for (int i = 0; i < ns_live; ++i)
        if (r[i] != r[i]) L.flags |= JSX_PIX_NAN;
    switch (L.sub) {
    case SUB_INIT:
        for (int i = 0; i < NS; ++i) L.k1[i] = r[i];
        if (L.phase == PH_TAU) L.deriv_ok = 1;
        L.tries = 0;
        lane_try_begin(L, S);
        return;
    case 2: for (int i = 0; i < NS; ++i) L.k2[i] = r[i]; lane_prepare_stage(L, 3); return;
    case 3: for (int i = 0; i < NS; ++i) L.k3[i] = r[i]; lane_prepare_stage(L, 4); return;
    case 4: for (int i = 0; i < NS; ++i) L.k4[i] = r[i]; lane_prepare_stage(L, 5); return;
    case 5: for (int i = 0; i < NS; ++i) L.k5[i] = r[i]; lane_prepare_stage(L, 6); return;
    case 6: for (int i = 0; i < NS; ++i) L.k6[i] = r[i]; lane_prepare_stage(L, 7); return;
    default: break;
    }
    /* stage 7 done: error estimate and step control */
    for (int i = 0; i < NS; ++i) L.k7[i] = r[i];
    double dt = L.dt;
    double err;
    if (L.phase == PH_TAU || NS == 1) {
        double xe = dt * DP_DC1 * L.k1[0] + dt * DP_DC3 * L.k3[0] + dt * DP_DC4 * L.k4[0] + dt * DP_DC5 * L.k5[0] +
                    dt * DP_DC6 * L.k6[0] + dt * DP_DC7 * L.k7[0];
        double ea = L.phase == PH_TAU ? S.tau_abs : S.st_abs;
        double er = L.phase == PH_TAU ? S.tau_rel : S.st_rel;
        err = fabs(xe) / (ea + er * (fabs(L.x[0]) + dt * fabs(L.k1[0])));
    } else {
        err = 0.0;
        for (int i = 0; i < NS; ++i) {
            double xe = dt * DP_DC1 * L.k1[i] + dt * DP_DC3 * L.k3[i] + dt * DP_DC4 * L.k4[i] +
                        dt * DP_DC5 * L.k5[i] + dt * DP_DC6 * L.k6[i] + dt * DP_DC7 * L.k7[i];
            double e = fabs(xe) / (S.st_abs + S.st_rel * (fabs(L.x[i]) + dt * fabs(L.k1[i])));
            err = (err < e) ? e : err;
        }
    }
    int ok;
    double maxdt = L.phase == PH_TAU ? S.dt_max : 0.0;
    if (err > 1.0) {
        double fac = 0.9 * jsx_pow(err, -1.0 / 3);
        L.dt = L.dt * (fac > 0.2 ? fac : 0.2);
        if (maxdt != 0.0) L.dt = L.dt > 0 ? (L.dt < maxdt ? L.dt : maxdt) : (L.dt > -maxdt ? L.dt : -maxdt);
        ok = 0;
    } else {
        L.t = L.t + L.dt;
        if (err < 0.5) {
            double e2 = S.emin > err ? S.emin : err;
            L.dt = L.dt * (0.9 * jsx_pow(e2, -1.0 / 5.0));
            if (maxdt != 0.0) L.dt = L.dt > 0 ? (L.dt < maxdt ? L.dt : maxdt) : (L.dt > -maxdt ? L.dt : -maxdt);
        }
        ok = 1;
    }
    if (L.tries++ >= S.max_fail) {
        L.flags |= JSX_PIX_FAIL_MAXTRIALS;
        if (L.phase == PH_TAU) {
            L.tau_bg = L.tau_bg + L.x[0];
            L.thick = L.thick + L.L;
            for (int i = 0; i < NS; ++i) L.x[i] = 0.0;
        }
        lane_finish(L);
        return;
    }
    if (!ok) { lane_try_begin(L, S); return; }
    /* accepted */
    if (L.phase == PH_TAU) {
        L.x_old = L.x[0];
        L.k_old = L.k1[0];
        L.x[0] = L.xt[0];
        L.k1[0] = L.k7[0];
        if (L.x[0] >= S.tau_max) {
            /* bisection on the dense output (Observation.cpp:166-187) */
            double t0 = L.t_old, t1 = L.t, t_m;
            double x_m = L.x[0];
            int it = 0;
            while (fabs(x_m - S.tau_max) > S.tau_prec) {
                if (++it > JSX_BISECT_CAP) { L.flags |= JSX_PIX_FAIL_BISECT; break; }
                t_m = 0.5 * (t0 + t1);
                x_m = lane_dense(L, t_m);
                if (x_m > S.tau_max) t1 = t_m;
                else t0 = t_m;
            }
            t_m = 0.5 * (t0 + t1);
            /* Ray(point_in, direction).point(t_m), direction normalised again */
            double po[3], ci[3] = {lane_cin(L, 0), lane_cin(L, 1), lane_cin(L, 2)};
            d_ray_point(ci, im.d2, t_m, po);
            for (int i = 0; i < 3; ++i) {
                if (L.seg) L.s1out[i] = po[i];
                else L.s0out[i] = po[i];
            }
            L.flags |= JSX_PIX_TAU_CROSSED;
            lane_tau_seg_end(L, S, 1);
            return;
        }
        lane_tau_range_next(L, S);
        return;
    }
    for (int i = 0; i < NS; ++i) { L.x[i] = L.xt[i]; L.k1[i] = L.k7[i]; }
    lane_stokes_loop(L, S);
}

/* evaluate the pending right-hand side for the lane */
template <int NS, bool EXT = false>
__device__ inline void lane_rhs(const Lane<NS> &L, const DModel &m, const DImage &im, const DSolver &S, double *r) {
    double p[3];
    lane_point(L, im, p);
    double n_los[3] = {-1. * im.dir[0], -1. * im.dir[1], -1. * im.dir[2]};
    DCoef c;
    d_coefficients<NS == 4 ? 1 : 0, EXT>(m, p, n_los, S.nu, c);
    if (L.phase == PH_TAU) {
        r[0] = c.kI;
        for (int i = 1; i < NS; ++i) r[i] = 0.0;
        return;
    }
    if (NS == 1) {
        r[0] = c.eI - c.kI * L.xt[0];
        return;
    }
    const double *x = L.xt;
    r[0] = c.eI - c.kI * x[0] - c.kQ * x[1] - c.kU * x[2] - c.kV * x[NS > 3 ? 3 : 0];
    if (NS == 4) {
        r[1] = c.eQ - c.kI * x[1] - c.kQ * x[0] - c.kF * x[2] - c.hQ * x[3];
        r[2] = c.eU - c.kI * x[2] - c.kU * x[0] + c.kF * x[1] - c.kC * x[3];
        r[3] = c.eV - c.kI * x[3] - c.kV * x[0] + c.hQ * x[1] + c.kC * x[2];
    }
}

template <int NS>
__device__ inline void lane_write(const Lane<NS> &L, const DOut &o, int64_t q) {
    uint32_t fl = L.flags;
    if ((fl & JSX_PIX_HIT) && L.x[0] != L.x[0]) fl |= JSX_PIX_NAN;
    if (o.tau) o.tau[q] = L.tau_bg;
    if (o.I) o.I[q] = L.x[0];
    if (o.Q) o.Q[q] = NS == 4 ? L.x[NS > 1 ? 1 : 0] : 0.0;
    if (o.U) o.U[q] = NS == 4 ? L.x[NS > 2 ? 2 : 0] : 0.0;
    if (o.V) o.V[q] = NS == 4 ? L.x[NS > 3 ? 3 : 0] : 0.0;
    if (o.l) o.l[q] = L.thick;
    if (o.flags) o.flags[q] = fl;
    if (o.n_rhs) o.n_rhs[q] = L.n_rhs;
    if (o.n_steps) o.n_steps[q] = L.n_steps;
}

#endif /* JSX_SOLVER_H */
