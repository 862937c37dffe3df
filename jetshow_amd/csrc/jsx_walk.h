/*
 * jsx_walk.h -- mixed-precision rays through grid-defined fields, integrated
 * cell by cell (included by jsx_host.hip).
 *
 * SimulationInterpolater's fields (src/SimulationInterpolater.cpp:33-102)
 * are linear inside each triangle of the (z, r) node grid and kinked on its
 * edges: the grid lines z = z_i, r = r_j and each cell's diagonal.  An
 * adaptive Dormand-Prince step that spans a kink carries an error its
 * embedded estimate does not see, so a tolerance of 1e-5 against the
 * converged image costs about one step (6+ right-hand sides) per kink.  Here
 * a ray is cut at every kink instead: the pieces between consecutive
 * crossings lie inside one triangle, where the fields -- and so the
 * coefficients -- are smooth, and each piece takes one 4th-order exponential
 * (Magnus) step from the coefficients at its two Gauss points:
 *
 *   tau pass:    tau += h/2 (k1 + k2)                  (Gauss-Legendre)
 *   Stokes pass: I <- e^-mu I + phi1(-mu) v,  mu = h/2 (k1 + k2),
 *                v = h/2 (e1 + e2) + sqrt(3)/12 h^2 (k1 e2 - k2 e1)
 *
 * exact for piecewise-constant coefficients and 4th-order inside a piece;
 * 2 coefficient evaluations per piece.  The crossings come from the ray's
 * own geometry: z(t) is linear, rho(t)^2 = A t^2 + 2 B t + C is quadratic,
 * and a diagonal rho = r_a + (dr/dz)(z - z_a) squares to a quadratic too;
 * the walk takes the nearest crossing ahead of the current cell's five lines.
 *
 * Scope: models whose B (random-direction magnitude), N and Gamma all come
 * from regular grids of one geometry -- the c2/c4 grid workloads.  With a
 * random-direction B field every polarized coefficient is zero, so full
 * Stokes is the I equation with Q = U = V = 0.  The passes follow the
 * lane machine's (jsx_mixed.h): tau from the observer until tau_max (per
 * segment, as the reference checks it) or the path's end, the Stokes pass
 * back to the observer from the optical-depth horizon, the tau_max point or
 * the end; tau <= tau_min skips the Stokes pass.  Outputs and flags as the
 * lane's; n_rhs counts coefficient evaluations, n_steps pieces.
 */
#ifndef JSX_WALK_H
#define JSX_WALK_H

#include "jsx_mixed.h"
#include "jsx_batch.h"

#define JSX_WALK_MAX_PIECES (1 << 22)   /* per pass: beyond it the rest is one piece (flag MAXTRIALS) */
/* Stokes pass: a piece optically thicker than this is taken in equal parts
   below it -- through a thick piece the emergent intensity is the source
   function near its exit, which two Gauss points of the whole piece miss
   (a ray at the jet's base, tau 274 over 46 pieces, was 9e-4 off) */
#ifndef JSX_WALK_MU_MAX
#define JSX_WALK_MU_MAX 0.5
#endif
#define JSX_WALK_MAX_PARTS 256

/* the ray during one pass: position o + t u [pc], rho^2 = A t^2 + 2 B t + C */
struct WalkLine {
    double o[3], u[3];
    double A, B, C;
};

__device__ inline void walk_line(WalkLine &w, const double *o_pc, const double *u) {
    for (int i = 0; i < 3; ++i) { w.o[i] = o_pc[i]; w.u[i] = u[i]; }
    w.A = u[0] * u[0] + u[1] * u[1];
    w.B = o_pc[0] * u[0] + o_pc[1] * u[1];
    w.C = o_pc[0] * o_pc[0] + o_pc[1] * o_pc[1];
}

/* smallest root of a t^2 + 2 b t + c = 0 greater than t_min (cap if none) */
__device__ inline double walk_root(double a, double b, double c, double t_min, double cap) {
    double best = cap;
    if (fabs(a) > 1e-300) {
        const double disc = b * b - a * c;
        if (disc < 0.0) return best;
        const double sq = sqrt(disc);
        /* the two roots without cancellation */
        const double q = -(b + (b >= 0.0 ? sq : -sq));
        const double r1 = q / a, r2 = q != 0.0 ? c / q : r1;
        if (r1 > t_min && r1 < best) best = r1;
        if (r2 > t_min && r2 < best) best = r2;
    } else if (fabs(b) > 1e-300) {
        const double r = -c / (2.0 * b);
        if (r > t_min && r < best) best = r;
    }
    return best;
}

/* the first kink crossing after t (at most t_end): the current cell's two z
   lines, two rho lines and diagonal, the cell taken just past t */
__device__ inline double walk_next(const WalkLine &w, const DGrid &g, double t, double t_end, double nudge) {
    const double tn = t + nudge;
    const double z = w.o[2] + tn * w.u[2];
    const double rho = sqrt(fmax(w.A * tn * tn + 2.0 * w.B * tn + w.C, 0.0));
    int i = (int)floor((z - g.z0) / g.dz);
    int j = (int)floor((rho - g.r0) / g.dr);
    if (i < -1) i = -1;
    if (i > g.nz - 1) i = g.nz - 1;
    if (j < -1) j = -1;
    if (j > g.nr - 1) j = g.nr - 1;
    double best = t_end;
    const double t_min = tn;
    /* z lines bounding the cell (outside the grid: the grid's edge) */
    if (w.u[2] != 0.0) {
        const double zs[2] = {g.z0 + (double)i * g.dz, g.z0 + (double)(i + 1) * g.dz};
        for (int k = 0; k < 2; ++k) {
            if ((k == 0 && i < 0) || (k == 1 && i > g.nz - 2)) continue;
            const double tz = (zs[k] - w.o[2]) / w.u[2];
            if (tz > t_min && tz < best) best = tz;
        }
    }
    /* rho lines */
    const double rs[2] = {g.r0 + (double)j * g.dr, g.r0 + (double)(j + 1) * g.dr};
    for (int k = 0; k < 2; ++k) {
        if ((k == 0 && j < 0) || (k == 1 && j > g.nr - 2)) continue;
        best = walk_root(w.A, w.B, w.C - rs[k] * rs[k], t_min, best);
    }
    /* the cell's diagonal rho = r_a + s (z - z_a), s = dr/dz, on rho >= 0 */
    if (i >= 0 && i <= g.nz - 2 && j >= 0 && j <= g.nr - 2) {
        const double s = g.dr / g.dz;
        const double za = g.z0 + (double)i * g.dz, ra = rs[0];
        const double al = ra + s * (w.o[2] - za), be = s * w.u[2];
        const double td = walk_root(w.A - be * be, w.B - al * be, w.C - al * al, t_min, best);
        if (td < best && al + be * td >= 0.0) best = td;
    }
    return best;
}

/* coefficients at the two Gauss points of [ta, ta + h] */
template <int FULL>
__device__ inline void walk_gauss(const DModel &m, const MConsts &c, const WalkLine &w, double ta, double h, MCoef &k1,
                                  MCoef &k2) {
    const double t1 = ta + 0.21132486540518713 * h, t2 = ta + 0.78867513459481287 * h;
    double p1[3], p2[3];
    for (int i = 0; i < 3; ++i) {
        p1[i] = w.o[i] + t1 * w.u[i];
        p2[i] = w.o[i] + t2 * w.u[i];
    }
    m_coefficients<FULL>(m, c, p1, k1);
    m_coefficients<FULL>(m, c, p2, k2);
}

struct WalkOut {
    double tau, I, l;
    uint32_t flags, n_rhs, n_steps;
};

/* one ray, pixel pix */
template <int FULL>
__device__ inline void walk_ray(const DModel &m, const DImage &im, const DSolver &S, const MConsts &c, int64_t pix,
                                WalkOut &out) {
    out.tau = 0.0; out.I = 0.0; out.l = 0.0;
    out.flags = 0; out.n_rhs = 0; out.n_steps = 0;
    double o[3];
    d_pixel_origin(im, pix, o);
    DSeg segs[2];
    int clipped = 0;
    const int ns = d_model_hit<false>(m, o, im.dir, segs, &clipped);
    if (clipped) out.flags |= JSX_PIX_CLIPPED;
    if (ns == 0) return;
    out.flags |= JSX_PIX_HIT;
    if (ns == 2) out.flags |= JSX_PIX_NSEG2;
    const DGrid &g = m.gb;
    const double nudge = 1e-9 * fmin(g.dz, g.dr);
    /* tau pass: from each segment's entry along the ray */
    double tau_tot = 0.0;
    int s_end = ns - 1;            /* last segment the Stokes pass covers */
    double t_stop[2] = {0.0, 0.0}; /* where it starts (from the entry) in each segment */
    int hz = 0;
    for (int s = 0; s < ns; ++s) {
        double pin[3], pout[3], u[3];
        for (int i = 0; i < 3; ++i) { pin[i] = segs[s].pin[i] * c.inv_pc; pout[i] = segs[s].pout[i] * c.inv_pc; }
        const double d[3] = {pout[0] - pin[0], pout[1] - pin[1], pout[2] - pin[2]};
        const double L = d_norm3(d);
        for (int i = 0; i < 3; ++i) u[i] = L > 0.0 ? d[i] / L : 0.0;
        t_stop[s] = L;
        WalkLine w;
        walk_line(w, pin, u);
        double t = 0.0, tau_seg = 0.0;
        int found = 0, pieces = 0;
        while (t < L && !found) {
            double tb = ++pieces < JSX_WALK_MAX_PIECES ? walk_next(w, g, t, L, nudge) : L;
            if (pieces >= JSX_WALK_MAX_PIECES) out.flags |= JSX_PIX_FAIL_MAXTRIALS;
            const double h = tb - t;
            MCoef k1, k2;
            walk_gauss<FULL>(m, c, w, t, h, k1, k2);
            out.n_rhs += 2;
            out.n_steps++;
            const double ka = (double)k1.kI, kb = (double)k2.kI;
            if (ka != ka || kb != kb) out.flags |= JSX_PIX_NAN;
            const double dtau = 0.5 * h * (ka + kb);
            if (tau_seg + dtau >= S.tau_max) {
                /* kI linear through the Gauss points: tau(s) = a s + b s^2 / 2 */
                const double s1 = 0.21132486540518713 * h, s2 = 0.78867513459481287 * h;
                const double bb = (kb - ka) / (s2 - s1), aa = ka - bb * s1;
                const double need = S.tau_max - tau_seg;
                double sx = walk_root(0.5 * bb, 0.5 * aa, -need, 0.0, h);
                if (!(sx <= h)) sx = h;
                tb = t + sx;
                /* tau reported as the lanes report it: through the step (here
                   the piece) that crossed tau_max */
                tau_seg += dtau;
                found = 1;
                out.flags |= JSX_PIX_TAU_CROSSED;
            } else {
                tau_seg += dtau;
            }
            t = tb;
            if (c.tau_hz > 0.0 && !hz && tau_tot + tau_seg >= c.tau_hz) {
                hz = 1;
                s_end = s;
                t_stop[s] = t;
                out.flags |= JSX_PIX_HORIZON;
            }
        }
        tau_tot += tau_seg;
        out.l += L * JSX_PC;   /* the segment's length, as the lanes count it */
        if (found) {
            if (!hz) { s_end = s; t_stop[s] = t; }
            break;
        }
    }
    out.tau = tau_tot;
    if (!(tau_tot > S.tau_min)) {
        out.flags |= JSX_PIX_I_SKIPPED;
        return;
    }
    /* Stokes pass: from the stop point of the last segment back to the observer */
    double I = 0.0;
    for (int s = s_end; s >= 0; --s) {
        double pin[3], pout[3], u[3];
        for (int i = 0; i < 3; ++i) { pin[i] = segs[s].pin[i] * c.inv_pc; pout[i] = segs[s].pout[i] * c.inv_pc; }
        const double d[3] = {pout[0] - pin[0], pout[1] - pin[1], pout[2] - pin[2]};
        const double L = d_norm3(d);
        const double Ls = t_stop[s];
        double start[3];
        for (int i = 0; i < 3; ++i) {
            u[i] = L > 0.0 ? -d[i] / L : 0.0;
            start[i] = pin[i] - Ls * u[i];
        }
        WalkLine w;
        walk_line(w, start, u);
        double t = 0.0;
        int pieces = 0;
        while (t < Ls) {
            const double tb = ++pieces < JSX_WALK_MAX_PIECES ? walk_next(w, g, t, Ls, nudge) : Ls;
            if (pieces >= JSX_WALK_MAX_PIECES) out.flags |= JSX_PIX_FAIL_MAXTRIALS;
            const double h = tb - t;
            MCoef k1, k2;
            walk_gauss<FULL>(m, c, w, t, h, k1, k2);
            out.n_rhs += 2;
            out.n_steps++;
            int parts = 1;
            {
                const double mu0 = 0.5 * h * ((double)k1.kI + (double)k2.kI);
                if (mu0 > JSX_WALK_MU_MAX) {
                    const double np = ceil(mu0 / JSX_WALK_MU_MAX);
                    parts = np < (double)JSX_WALK_MAX_PARTS ? (int)np : JSX_WALK_MAX_PARTS;
                }
            }
            const double hp = h / (double)parts;
            for (int q = 0; q < parts; ++q) {
                if (parts > 1) {
                    walk_gauss<FULL>(m, c, w, t + (double)q * hp, hp, k1, k2);
                    out.n_rhs += 2;
                }
                const double ka = (double)k1.kI, kb = (double)k2.kI, ea = (double)k1.eI, eb = (double)k2.eI;
                if (ka != ka || kb != kb || ea != ea || eb != eb) out.flags |= JSX_PIX_NAN;
                const double mu = 0.5 * hp * (ka + kb);
                const double v = 0.5 * hp * (ea + eb) + 0.14433756729740643 * hp * hp * (ka * eb - kb * ea);
                const double ph = mu != 0.0 ? -expm1(-mu) / mu : 1.0;   /* phi1(-mu) */
                I = exp(-mu) * I + ph * v;
            }
            t = tb;
        }
    }
    out.I = I;
}

__global__ void jsx_k_walk_reset(unsigned long long *counter) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0ull;
}

/* one thread per ray of the launch's requests (ray ids in request order,
   off[] their prefix sums), handed out in wavefront chunks */
template <int NS>
__global__ __launch_bounds__(256) void jsx_k_walk(const DParamsM *__restrict__ reqs, const int64_t *__restrict__ off,
                                                   const DOut *__restrict__ outs, int nreq, int64_t total,
                                                   unsigned long long *counter) {
    const int lane = __lane_id();
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(counter, 64ull);
        base = __shfl(base, 0);
        if (base >= (unsigned long long)total) break;
        const int64_t id = (int64_t)base + lane;
        if (id < total) {
            const int req = d_find_request(off, nreq, id);
            const int64_t pix = id - off[req];
            const DParamsM *q = reqs + req;
            WalkOut w;
            walk_ray<NS == 4 ? 1 : 0>(q->m, q->im, q->S, q->c, pix, w);
            const DOut o = outs[req];
            if (o.tau) o.tau[pix] = w.tau;
            if (o.I) o.I[pix] = w.I;
            if (o.Q) o.Q[pix] = 0.0;
            if (o.U) o.U[pix] = 0.0;
            if (o.V) o.V[pix] = 0.0;
            if (o.l) o.l[pix] = w.l;
            if (o.flags) o.flags[pix] = w.flags;
            if (o.n_rhs) o.n_rhs[pix] = w.n_rhs;
            if (o.n_steps) o.n_steps[pix] = w.n_steps;
        }
    }
}

#endif /* JSX_WALK_H */
