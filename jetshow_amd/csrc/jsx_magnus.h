/*
 * jsx_magnus.h -- exponential (Magnus) propagators for the pieces of split
 * Faraday-thick rays.
 *
 * The Stokes system dS/dt = eps(t) - K(t) S has K = k_I 1 + G(eta, rho), where
 *
 *   G(eta, rho) = [[0,    eta_Q,  eta_U,  eta_V],
 *                  [eta_Q, 0,     rho_V, -rho_U],
 *                  [eta_U, -rho_V, 0,     rho_Q],
 *                  [eta_V, rho_U, -rho_Q, 0    ]]
 *
 * with eta = (k_Q, k_U, k_V) (dichroism) and rho = (k_C, -h_Q, k_F)
 * (birefringence: conversion and Faraday rotation).  These matrices form the
 * Lorentz algebra: [G(e1, r1), G(e2, r2)] = G(r2 x e1 - r1 x e2, e1 x e2 - r1 x r2).
 *
 * One step of length h uses the 4th-order Magnus expansion with the two
 * Gauss-Legendre points of the step (A = -K, b = eps, augmented by a constant
 * 1 so the source rides along):
 *
 *   Omega = h/2 (A1 + A2) + sqrt(3)/12 h^2 [A2, A1]
 *   [S; 1](t + h) = exp([[Omega_A, Omega_b], [0, 0]]) [S; 1](t)
 *
 * and the exponential is exact: with M = -mu 1 + G(eta, rho), G's eigenvalues
 * are +-a, +-i b where a^2, b^2 solve x^2 - (|eta|^2 - |rho|^2) x - (eta.rho)^2
 * = 0, so G^4 = p G^2 + q 1 (p = a^2 - b^2, q = a^2 b^2) and any function of G
 * is a cubic in G whose coefficients interpolate the function at the
 * eigenvalues:
 *
 *   exp(M)         = e^-mu (c0 + c1 G + c2 G^2 + c3 G^3)
 *   phi1(M) v      = (d0 + d1 G + d2 G^2 + d3 G^3) v,  phi1(M) = int_0^1 e^{sM} ds
 *
 * c_k, d_k in closed form (hyperbolic functions of a, trigonometric of b) or,
 * when a^2 + b^2 <= 1, as Taylor series in G reduced with G^4 = p G^2 + q.
 * A whole step spans any number of Faraday rotations (the rotation is inside
 * the exponential), so on the rays that bound Dormand-Prince by |k_F| the step
 * follows how fast the coefficients change instead, ~10^2 times fewer
 * coefficient evaluations.  The step error is estimated by the 2nd-order
 * exponential of the same points (Omega without the commutator), which
 * controls the step size with exponent 1/3.
 *
 * All of this runs in FP64 (FP64 vector issue is full rate on gfx950), with
 * multiply-adds fused (JSX_MIXED_FMA: nothing here has to reproduce the
 * oracle bit for bit); the coefficients come from the FP32 coefficient code
 * of the mixed path.
 */
#ifndef JSX_MAGNUS_H
#define JSX_MAGNUS_H

#include "jsx_mixed.h"

#define JSX_MG_NSER 20          /* Taylor terms of the small-generator branch */
#define JSX_MG_GK_START 60      /* backward-recurrence start of g_k           */

struct MgGen { double e[3], r[3]; };

__device__ __forceinline__ void mg_apply(const MgGen &g, const double *x, double *y) {
    JSX_MIXED_FMA
    y[0] = g.e[0] * x[1] + g.e[1] * x[2] + g.e[2] * x[3];
    y[1] = g.e[0] * x[0] + g.r[2] * x[2] - g.r[1] * x[3];
    y[2] = g.e[1] * x[0] - g.r[2] * x[1] + g.r[0] * x[3];
    y[3] = g.e[2] * x[0] + g.r[1] * x[1] - g.r[0] * x[2];
}

/* [G(a), G(b)] */
__device__ __forceinline__ void mg_comm(const MgGen &a, const MgGen &b, MgGen &c) {
    JSX_MIXED_FMA
    c.e[0] = (b.r[1] * a.e[2] - b.r[2] * a.e[1]) - (a.r[1] * b.e[2] - a.r[2] * b.e[1]);
    c.e[1] = (b.r[2] * a.e[0] - b.r[0] * a.e[2]) - (a.r[2] * b.e[0] - a.r[0] * b.e[2]);
    c.e[2] = (b.r[0] * a.e[1] - b.r[1] * a.e[0]) - (a.r[0] * b.e[1] - a.r[1] * b.e[0]);
    c.r[0] = (a.e[1] * b.e[2] - a.e[2] * b.e[1]) - (a.r[1] * b.r[2] - a.r[2] * b.r[1]);
    c.r[1] = (a.e[2] * b.e[0] - a.e[0] * b.e[2]) - (a.r[2] * b.r[0] - a.r[0] * b.r[2]);
    c.r[2] = (a.e[0] * b.e[1] - a.e[1] * b.e[0]) - (a.r[0] * b.r[1] - a.r[1] * b.r[0]);
}

/* 1/k (k <= JSX_MG_GK_START) */
__device__ __forceinline__ double mg_inv(int k) {
    static const double t[JSX_MG_GK_START + 1] = {
        0.0, 1.0, 1.0 / 2, 1.0 / 3, 1.0 / 4, 1.0 / 5, 1.0 / 6, 1.0 / 7, 1.0 / 8, 1.0 / 9, 1.0 / 10,
        1.0 / 11, 1.0 / 12, 1.0 / 13, 1.0 / 14, 1.0 / 15, 1.0 / 16, 1.0 / 17, 1.0 / 18, 1.0 / 19, 1.0 / 20,
        1.0 / 21, 1.0 / 22, 1.0 / 23, 1.0 / 24, 1.0 / 25, 1.0 / 26, 1.0 / 27, 1.0 / 28, 1.0 / 29, 1.0 / 30,
        1.0 / 31, 1.0 / 32, 1.0 / 33, 1.0 / 34, 1.0 / 35, 1.0 / 36, 1.0 / 37, 1.0 / 38, 1.0 / 39, 1.0 / 40,
        1.0 / 41, 1.0 / 42, 1.0 / 43, 1.0 / 44, 1.0 / 45, 1.0 / 46, 1.0 / 47, 1.0 / 48, 1.0 / 49, 1.0 / 50,
        1.0 / 51, 1.0 / 52, 1.0 / 53, 1.0 / 54, 1.0 / 55, 1.0 / 56, 1.0 / 57, 1.0 / 58, 1.0 / 59, 1.0 / 60};
    return t[k];
}

/* 1/k! (k < JSX_MG_NSER) */
__device__ __forceinline__ double mg_ifact(int k) {
    static const double t[JSX_MG_NSER] = {
        1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0, 1.0 / 5040.0, 1.0 / 40320.0,
        1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0, 1.0 / 479001600.0, 1.0 / 6227020800.0,
        1.0 / 87178291200.0, 1.0 / 1307674368000.0, 1.0 / 20922789888000.0, 1.0 / 355687428096000.0,
        1.0 / 6402373705728000.0, 1.0 / 121645100408832000.0};
    return t[k];
}

/* Multiplication by G in the basis {1, G, G^2, G^3}: G^4 = p G^2 + q. */
__device__ __forceinline__ void mg_times_g(double *r, double p, double q) {
    JSX_MIXED_FMA
    const double n0 = r[3] * q, n2 = r[1] + r[3] * p;
    r[3] = r[2];
    r[1] = r[0];
    r[0] = n0;
    r[2] = n2;
}

/* The series of g_k(z) = int_0^1 s^k e^{sz} ds, k < JSX_MG_NSER, z <= 0 in
 * use, fed to `use(k, g_k)` in descending k when |z| < JSX_MG_NSER + 5
 * (backward recurrence g_{k-1} = (e^z - z g_k) / k from k = JSX_MG_GK_START
 * with g_N ~ e^z / (N + 1 + z); errors shrink by |z|/k per step), otherwise in
 * ascending k (forward recurrence g_k = (e^z - k g_{k-1}) / z from
 * g_0 = expm1(z) / z; errors shrink by k/|z| per step). */
template <class F>
__device__ __forceinline__ void mg_g_series(double z, double ez, F use) {
    JSX_MIXED_FMA
    if (fabs(z) >= (double)(JSX_MG_NSER + 5)) {
        const double iz = 1.0 / z;
        double g = expm1(z) * iz;
        use(0, g);
#pragma unroll
        for (int k = 1; k < JSX_MG_NSER; ++k) {
            g = (ez - (double)k * g) * iz;
            use(k, g);
        }
        return;
    }
    /* start of the backward recurrence: for z <= 0 each step down scales the
       relative error of the start value by at most |z| / k, so the start
       index K with prod_{k=NSER+1..K} |z| / k <= 1e-17 gives g_0..g_NSER-1 to
       rounding; K is the largest any active lane of the wavefront needs (a
       uniform loop keeps 1/k a scalar load) */
    const double az = fabs(z);
    const int K = __ballot(az > 12.0) ? JSX_MG_GK_START : __ballot(az > 8.0) ? 56 : __ballot(az > 4.0) ? 48
                : __ballot(az > 2.0) ? 40 : __ballot(az > 1.0) ? 35 : __ballot(az > 0.1) ? 32 : 28;
    double g = ez / ((double)(K + 1) + z);
    for (int k = K; k > JSX_MG_NSER; --k) g = (ez - z * g) * mg_inv(k);
#pragma unroll
    for (int k = JSX_MG_NSER; k > 0; --k) {
        g = (ez - z * g) * mg_inv(k);
        use(k - 1, g);
    }
}

/* sinh(a)/a - 1 and cosh(a) - 1 for a < 0.5, 1 - sin(b)/b for b < 0.5 (Taylor) */
__device__ __forceinline__ double mg_shm1(double a2) {
    JSX_MIXED_FMA
    return a2 * (1.0 / 6) * (1 + a2 * (1.0 / 20) * (1 + a2 * (1.0 / 42) * (1 + a2 * (1.0 / 72) *
           (1 + a2 * (1.0 / 110) * (1 + a2 * (1.0 / 156) * (1 + a2 * (1.0 / 210)))))));
}
__device__ __forceinline__ double mg_chm1(double a2) {
    JSX_MIXED_FMA
    return a2 * 0.5 * (1 + a2 * (1.0 / 12) * (1 + a2 * (1.0 / 30) * (1 + a2 * (1.0 / 56) *
           (1 + a2 * (1.0 / 90) * (1 + a2 * (1.0 / 132) * (1 + a2 * (1.0 / 182)))))));
}
__device__ __forceinline__ double mg_scm1(double b2) {
    JSX_MIXED_FMA
    return b2 * (1.0 / 6) * (1 - b2 * (1.0 / 20) * (1 - b2 * (1.0 / 42) * (1 - b2 * (1.0 / 72) *
           (1 - b2 * (1.0 / 110) * (1 - b2 * (1.0 / 156) * (1 - b2 * (1.0 / 210)))))));
}

/* (e^z - 1) / z */
__device__ __forceinline__ double mg_phi1(double z) { return z != 0.0 ? expm1(z) / z : 1.0; }

/* Coefficients of exp(-mu 1 + G) = sum_k c_k G^k and phi1(-mu 1 + G) =
 * sum_k d_k G^k (k < 4).  Returns 0 when the step is unusable (growth beyond
 * e^40, or a non-finite coefficient), which the caller treats as a rejected
 * step.
 *
 * Two forms: the Taylor series in G (a^2 + b^2 <= 1 only) and the closed form
 * from the eigenvalues (any G: its small-argument pieces -- cosh a - 1,
 * sinh(a)/a - 1, 1 - cos b, 1 - e^-mu cos b, and phi1's even and odd parts at
 * a < 1/2 -- are evaluated without cancellation, and G with all eigenvalues
 * zero takes the limits).  MODE 0 (the kernels): the series when every active
 * lane of the wavefront can take it, else the closed form for all, so a
 * wavefront never runs both; MODE 1: the closed form; MODE 2: per lane, the
 * series exactly when a^2 + b^2 <= 1. */
template <int MODE>
__device__ inline int mg_coef_t(double mu, const MgGen &g, double *c, double *d) {
    JSX_MIXED_FMA
    const double ee = g.e[0] * g.e[0] + g.e[1] * g.e[1] + g.e[2] * g.e[2];
    const double rr = g.r[0] * g.r[0] + g.r[1] * g.r[1] + g.r[2] * g.r[2];
    const double er = g.e[0] * g.r[0] + g.e[1] * g.r[1] + g.e[2] * g.r[2];
    const double p = ee - rr, q = er * er;
    const double s = sqrt(p * p + 4.0 * q);
    double a2, b2;
    if (p >= 0.0) {
        a2 = 0.5 * (s + p);
        b2 = a2 > 0.0 ? q / a2 : 0.0;
    } else {
        b2 = 0.5 * (s - p);
        a2 = b2 > 0.0 ? q / b2 : 0.0;
    }
    const double s2 = a2 + b2;
    const double em = exp(-mu);
    const bool series = MODE == 1 ? false : MODE == 2 ? s2 <= 1.0 : __ballot(!(s2 <= 1.0)) == 0ull;
    if (series) {
        /* Taylor series in G, powers reduced with G^4 = p G^2 + q */
        double rc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = JSX_MG_NSER - 1; k >= 0; --k) {
            mg_times_g(rc, p, q);
            rc[0] += mg_ifact(k);
        }
        for (int i = 0; i < 4; ++i) c[i] = em * rc[i];
        /* d: sum_k g_k / k! G^k, accumulated in the order the g_k come */
        double rd[4] = {0.0, 0.0, 0.0, 0.0}, ba[4] = {1.0, 0.0, 0.0, 0.0};
        if (fabs(mu) >= (double)(JSX_MG_NSER + 5)) {
            mg_g_series(-mu, em, [&](int k, double gk) {
                const double f = gk * mg_ifact(k);
                for (int i = 0; i < 4; ++i) rd[i] += f * ba[i];
                mg_times_g(ba, p, q);
            });
        } else {
            mg_g_series(-mu, em, [&](int k, double gk) {
                mg_times_g(rd, p, q);
                rd[0] += gk * mg_ifact(k);
            });
        }
        for (int i = 0; i < 4; ++i) d[i] = rd[i];
    } else {
        const double a = sqrt(a2), b = sqrt(b2);
        if (a - mu > 40.0) return 0;
        /* e^-mu (cosh a - 1), e^-mu (sinh a / a - 1) */
        double chm1, shm1;
        if (a < 0.5) {
            chm1 = em * mg_chm1(a2);
            shm1 = em * mg_shm1(a2);
        } else {
            /* e^(a - mu) and e^(-a - mu) from one exponential */
            const double ea = exp(a);
            const double ep = em * ea, en = em / ea;
            chm1 = 0.5 * (ep + en) - em;
            shm1 = 0.5 * (ep - en) / a - em;
        }
        const double ch = em + chm1, sh = em + shm1;     /* e^-mu cosh a, e^-mu sinh(a)/a */
        double hs, hc;
        sincos(0.5 * b, &hs, &hc);
        const double sb = 2.0 * hs * hc;                 /* sin b */
        const double omc = 2.0 * hs * hs;                /* 1 - cos b */
        const double cb = 1.0 - omc;
        const double scm1 = b < 0.5 ? mg_scm1(b2) : 1.0 - sb / b;      /* 1 - sin(b)/b */
        const double sc = 1.0 - scm1;
        const double is2 = 1.0 / s2;
        c[0] = (b2 * ch + a2 * em * cb) * is2;
        c[1] = (b2 * sh + a2 * em * sc) * is2;
        c[2] = (chm1 + em * omc) * is2;
        c[3] = (shm1 + em * scm1) * is2;
        /* phi1's even (E) and odd (O) parts at lambda = a and lambda = i b */
        double Ea = 0.0, Oa = 0.0, g03[4] = {0.0, 0.0, 0.0, 0.0};
        if (a < 0.5) {
            /* sum_k a^{2k} g_{2k} / (2k)!, sum_k a^{2k} g_{2k+1} / (2k+1)! */
            if (fabs(mu) >= (double)(JSX_MG_NSER + 5)) {
                double ak = 1.0;
                mg_g_series(-mu, em, [&](int k, double gk) {
                    if (k < 4) g03[k] = gk;
                    if (k & 1) {
                        Oa += ak * gk * mg_ifact(k);
                        ak *= a2;
                    } else {
                        Ea += ak * gk * mg_ifact(k);
                    }
                });
            } else {
                mg_g_series(-mu, em, [&](int k, double gk) {
                    if (k < 4) g03[k] = gk;
                    if (k & 1) Oa = Oa * a2 + gk * mg_ifact(k);
                    else Ea = Ea * a2 + gk * mg_ifact(k);
                });
            }
        } else {
            const double pp = mg_phi1(a - mu), pm = mg_phi1(-a - mu);
            Ea = 0.5 * (pp + pm);
            Oa = 0.5 * (pp - pm) / a;
        }
        /* phi1(i b - mu): real part and imaginary part / b */
        double Eb = 1.0, Ob = 0.5;
        if (mu * mu + b2 > 1e-300) {
            const double dz = 1.0 / (mu * mu + b2);
            /* 1 - e^-mu cos b, both terms non-negative for mu >= 0 */
            const double one_m = -expm1(-mu) + em * omc;
            Eb = (mu * one_m + b * em * sb) * dz;
            Ob = (one_m - mu * em * sc) * dz;
        }
        if (s2 < 1e-200) {
            /* all eigenvalues zero (G^4 = 0): the series' first four terms */
            c[0] = em; c[1] = em; c[2] = 0.5 * em; c[3] = em * (1.0 / 6.0);
            d[0] = g03[0]; d[1] = g03[1]; d[2] = 0.5 * g03[2]; d[3] = g03[3] * (1.0 / 6.0);
        } else {
            d[0] = (b2 * Ea + a2 * Eb) * is2;
            d[1] = (b2 * Oa + a2 * Ob) * is2;
            d[2] = (Ea - Eb) * is2;
            d[3] = (Oa - Ob) * is2;
        }
    }
    const double chk = c[0] + c[1] + c[2] + c[3] + d[0] + d[1] + d[2] + d[3];
    return chk == chk && fabs(chk) < 1e300;
}

__device__ inline int mg_coef(double mu, const MgGen &g, double *c, double *d) { return mg_coef_t<0>(mu, g, c, d); }

/* y = (k0 + k1 G + k2 G^2 + k3 G^3) x */
__device__ __forceinline__ void mg_poly(const MgGen &g, const double *k, const double *x, double *y) {
    JSX_MIXED_FMA
    double x1[4], x2[4], x3[4];
    mg_apply(g, x, x1);
    mg_apply(g, x1, x2);
    mg_apply(g, x2, x3);
    for (int i = 0; i < 4; ++i) y[i] = k[0] * x[i] + k[1] * x1[i] + k[2] * x2[i] + k[3] * x3[i];
}

/* E = exp(-mu 1 + G) (column-major, E[4 j + i]) and w = phi1(-mu 1 + G) v,
 * with mg_coef_t<MODE> */
template <int MODE = 0>
__device__ inline int mg_expphi(double mu, const MgGen &g, const double *v, double *E, double *w) {
    JSX_MIXED_FMA
    double c[4], d[4];
    if (!mg_coef_t<MODE>(mu, g, c, d)) return 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double x[4] = {j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0, j == 3 ? 1.0 : 0.0};
        mg_poly(g, c, x, E + 4 * j);
    }
    mg_poly(g, d, v, w);
    return 1;
}

/* The transfer terms of one coefficient evaluation: A x = -kI x - G x, source s */
__device__ __forceinline__ void mg_gen(const MCoef &k, MgGen &g) {
    g.e[0] = (double)k.kQ; g.e[1] = (double)k.kU; g.e[2] = (double)k.kV;
    g.r[0] = (double)k.kC; g.r[1] = -(double)k.hQ; g.r[2] = (double)k.kF;
}

/* One Magnus step over [t, t + h] from the coefficients at the two Gauss
 * points: E4, w4 (4th order) and an estimate of the error of the 2nd-order
 * step of the same points, err = max(|E4 - E2| / rel, |w4 - w2| / (rel sP))
 * with sP the larger of p_floor and the particular solution's scale.
 * est 1: E2 and w2 formed (a second exponential, compared column by column).
 * est 0: their first-order difference, the commutator terms the 4th-order
 * step adds -- dG = sqrt(3)/12 h^2 [A2, A1] in the generator and
 * dv = v4 - v2 in the source -- carried through exp(-mu + G), whose norm
 * e^{-mu + |eta|} bounds: |E4 - E2| ~ e^{-mu+|eta|} max|dG|,
 * |w4 - w2| ~ e^{-mu+|eta|} (max|dv| + max|dG| max|w4|).  Returns 0 if the
 * step is unusable. */
__device__ inline int mg_step(const MCoef &k1, const MCoef &k2, double h, double rel, double p_floor,
                              const double *P, double *E4, double *w4, double *err, int est = 0) {
    JSX_MIXED_FMA
    const double hh = 0.5 * h, cc = 0.14433756729740643 * h * h;   /* sqrt(3)/12 h^2 */
    const double mu = hh * ((double)k1.kI + (double)k2.kI);
    MgGen G1, G2, g2, g4, cm;
    mg_gen(k1, G1);
    mg_gen(k2, G2);
    mg_comm(G2, G1, cm);              /* [G2, G1] = [A2, A1] */
    for (int i = 0; i < 3; ++i) {
        g2.e[i] = -hh * (G1.e[i] + G2.e[i]);
        g2.r[i] = -hh * (G1.r[i] + G2.r[i]);
        g4.e[i] = g2.e[i] + cc * cm.e[i];
        g4.r[i] = g2.r[i] + cc * cm.r[i];
    }
    /* v2 = h/2 (s1 + s2), v4 = v2 + cc (A2 s1 - A1 s2) */
    const double s1[4] = {(double)k1.eI, (double)k1.eQ, (double)k1.eU, (double)k1.eV};
    const double s2[4] = {(double)k2.eI, (double)k2.eQ, (double)k2.eU, (double)k2.eV};
    double v2[4], v4[4], t1[4], t2[4];
    mg_apply(G2, s1, t1);
    mg_apply(G1, s2, t2);
    for (int i = 0; i < 4; ++i) {
        v2[i] = hh * (s1[i] + s2[i]);
        const double a2s1 = -(double)k2.kI * s1[i] - t1[i];
        const double a1s2 = -(double)k1.kI * s2[i] - t2[i];
        v4[i] = v2[i] + cc * (a2s1 - a1s2);
    }
    if (est == 0) {
        double c4[4], d4[4];
        if (!mg_coef(mu, g4, c4, d4)) {
            *err = __builtin_inf();
            return 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double x[4] = {j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0, j == 3 ? 1.0 : 0.0};
            mg_poly(g4, c4, x, E4 + 4 * j);
        }
        mg_poly(g4, d4, v4, w4);
        double dG = 0.0, dv = 0.0, ww = 0.0, sP = p_floor;
        for (int i = 0; i < 3; ++i) dG = fmax(dG, fmax(fabs(cc * cm.e[i]), fabs(cc * cm.r[i])));
        for (int i = 0; i < 4; ++i) {
            dv = fmax(dv, fabs(v4[i] - v2[i]));
            ww = fmax(ww, fabs(w4[i]));
            sP = fmax(sP, fmax(fabs(P[i]), fabs(w4[i])));
        }
        const double eta = sqrt(g4.e[0] * g4.e[0] + g4.e[1] * g4.e[1] + g4.e[2] * g4.e[2]);
        /* e^z and phi1(z) = (e^z - 1) / z from one expm1 */
        const double zg = eta - mu, em1 = expm1(zg);
        const double grow = em1 + 1.0;
        const double e1 = grow * dG / rel;
        const double eP = (zg != 0.0 ? em1 / zg : 1.0) * (dv + dG * ww);
        const double e2 = sP > 0.0 ? eP / (rel * sP) : (eP > 0.0 ? __builtin_inf() : 0.0);
        *err = e1 > e2 ? e1 : e2;
        return (*err == *err) ? 1 : 0;
    }
    double c4[4], d4[4], c2[4], d2[4];
    if (!mg_coef(mu, g4, c4, d4) || !mg_coef(mu, g2, c2, d2)) {
        *err = __builtin_inf();
        return 0;
    }
    double eE = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double x[4] = {j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0, j == 3 ? 1.0 : 0.0};
        double y2[4];
        mg_poly(g4, c4, x, E4 + 4 * j);
        mg_poly(g2, c2, x, y2);
        for (int i = 0; i < 4; ++i) {
            const double dv = fabs(E4[4 * j + i] - y2[i]);
            eE = dv > eE ? dv : eE;
        }
    }
    double w2[4];
    mg_poly(g4, d4, v4, w4);
    mg_poly(g2, d2, v2, w2);
    double eP = 0.0, sP = p_floor;
    for (int i = 0; i < 4; ++i) {
        const double dv = fabs(w4[i] - w2[i]);
        eP = dv > eP ? dv : eP;
        const double a = fabs(P[i]), b = fabs(w4[i]);
        sP = a > sP ? a : sP;
        sP = b > sP ? b : sP;
    }
    const double e1 = eE / rel;
    const double e2 = sP > 0.0 ? eP / (rel * sP) : (eP > 0.0 ? __builtin_inf() : 0.0);
    *err = e1 > e2 ? e1 : e2;
    return (*err == *err) ? 1 : 0;
}

#endif /* JSX_MAGNUS_H */
