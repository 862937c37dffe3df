/*
 * jsx_math.h -- deterministic FP64 elementary functions shared by the HIP
 * kernels and by the host.
 *
 * Why this exists: the per-pixel solver decides discrete things from
 * transcendental results (accept/reject of a Dormand-Prince step, whether tau
 * crossed tau_max, whether I is integrated at all).  glibc's libm and the
 * GPU's device library give results that differ in the last ulp, so a CPU
 * restatement and a GPU kernel built on different libraries can take
 * different discrete paths on marginal rays.  Every function here uses only
 * IEEE-754 operations that are correctly rounded on both x86-64 and gfx950
 * (+, -, *, /, sqrt, fma, floor) in a fixed order, so the same inputs give the
 * same bits on both sides.  Compile every user with -ffp-contract=off.
 *
 * Methods (no vendor library code is reproduced):
 *   exp : Cody & Waite argument reduction x = k ln2 + r with ln2 split into a
 *         32-bit head and a tail (Cody & Waite, "Software Manual for the
 *         Elementary Functions", 1980), Taylor polynomial of degree 13 on
 *         |r| <= ln2/2 evaluated by Horner's rule with fma, exact scaling by 2^k.
 *   log : x = 2^k m, m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
 *         s = f/(2+f), log(1+f) = f - f^2/2 + s (f^2/2 + R(s^2)) with
 *         R(z) = sum_{j>=1} 2 z^j/(2j+1) (the atanh series), result kept as a
 *         double-double head/tail using Knuth's two-sum and an fma two-product.
 *   pow : exp(y * log(x)) with the product formed in double-double, which
 *         keeps the result within ~1 ulp over the ranges the solver uses.
 *   sin(atan(u)), cos(atan(u)) : the algebraic identities u/sqrt(1+u^2),
 *         1/sqrt(1+u^2) (scaled for |u| > 1), which is what the reference's
 *         ToroidalBField / SimulationBField compute through atan/sin/cos.
 *   sin, cos, tan : x = k pi/2 + r with pi/2 in three parts (Cody & Waite),
 *         |r| <= pi/4, Taylor polynomials of sin and cos (to x^27 / x^26),
 *         quadrant from k; for |x| < 2^20 (the field models call it with
 *         |x| <= a pitch angle).
 *   acos : asin's Maclaurin series sum (2k)! / (4^k k!^2 (2k+1)) y^(2k+1) on
 *         |y| <= 1/2, with acos(u) = pi/2 - asin(u) for |u| <= 1/2 and the
 *         half-angle forms 2 asin(sqrt((1-u)/2)), pi - 2 asin(sqrt((1+u)/2))
 *         beyond; NaN outside [-1, 1].
 *   I0, I1 (modified Bessel functions, parameter constants only) : their
 *         power series sum (x/2)^(2k+n) / (k! (k+n)!) to the last term that
 *         changes the sum.
 */
#ifndef JSX_MATH_H
#define JSX_MATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define JSX_HD __host__ __device__ inline
#else
#define JSX_HD static inline
#endif

#define JSX_LN2_HI 6.93147180369123816490e-01   /* ln 2 head, 32 significant bits */
#define JSX_LN2_LO 1.90821492927058770002e-10   /* ln 2 - JSX_LN2_HI */
#define JSX_INV_LN2 1.44269504088896338700e+00
#define JSX_SQRT2 1.41421356237309514547e+00
#define JSX_EXP_OVF 7.09782712893383973096e+02  /* ln(DBL_MAX) */
#define JSX_EXP_UNF -7.45133219101941108420e+02 /* ln(smallest subnormal / 2) */

typedef union { double d; uint64_t u; } jsx_bits;

JSX_HD double jsx_from_bits(uint64_t u) { jsx_bits b; b.u = u; return b.d; }
JSX_HD uint64_t jsx_to_bits(double d) { jsx_bits b; b.d = d; return b.u; }

/* 2^k for -1022 <= k <= 1023, exact. */
JSX_HD double jsx_pow2i(int k) { return jsx_from_bits((uint64_t)(k + 1023) << 52); }

/* exp(hi + lo) for |lo| <= ulp(hi). */
JSX_HD double jsx_exp_dd(double hi, double lo) {
    if (hi != hi) return hi;
    if (hi > JSX_EXP_OVF) return jsx_from_bits(0x7ff0000000000000ULL);
    if (hi < JSX_EXP_UNF) return 0.0;
    double kd = floor(hi * JSX_INV_LN2 + 0.5);
    double r = (hi - kd * JSX_LN2_HI) - kd * JSX_LN2_LO;
    r = r + lo;
    /* Taylor coefficients 1/j!, folded by the compiler with correct rounding */
    double p = 1.0 / 6227020800.0;                 /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);               /* 1/12! */
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    int k = (int)kd;
    if (k >= -1021 && k <= 1023) return p * jsx_pow2i(k);
    if (k > 1023) return (p * jsx_pow2i(1023)) * jsx_pow2i(k - 1023);
    /* result is subnormal or close to it: scale in two exact-or-once-rounded steps */
    return (p * jsx_pow2i(k + 1000)) * jsx_pow2i(-1000);
}

JSX_HD double jsx_exp(double x) { return jsx_exp_dd(x, 0.0); }

/* log(x) as head + tail; *lo receives the tail.  x must be finite and > 0. */
JSX_HD double jsx_log_dd(double x, double *lo) {
    int kadj = 0;
    if (x < 2.2250738585072014e-308) { x = x * 18014398509481984.0; kadj = -54; } /* 2^54 */
    uint64_t u = jsx_to_bits(x);
    int e = (int)((u >> 52) & 0x7ff) - 1023 + kadj;
    double m = jsx_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > JSX_SQRT2) { m = m * 0.5; e = e + 1; }
    double f = m - 1.0;                       /* exact (Sterbenz) */
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 2.0 / 21.0;
    R = fma(R, z, 2.0 / 19.0);
    R = fma(R, z, 2.0 / 17.0);
    R = fma(R, z, 2.0 / 15.0);
    R = fma(R, z, 2.0 / 13.0);
    R = fma(R, z, 2.0 / 11.0);
    R = fma(R, z, 2.0 / 9.0);
    R = fma(R, z, 2.0 / 7.0);
    R = fma(R, z, 2.0 / 5.0);
    R = fma(R, z, 2.0 / 3.0);
    R = R * z;
    double fsq = f * f;
    double fsq_lo = fma(f, f, -fsq);
    double hfsq = 0.5 * fsq;
    double hfsq_lo = 0.5 * fsq_lo;
    double t = s * (hfsq + R);
    double kd = (double)e;
    double a = kd * JSX_LN2_HI;               /* exact: 32-bit head times |k| < 2^11 */
    /* two-sum(a, f) */
    double h1 = a + f;
    double bb = h1 - a;
    double l1 = (a - (h1 - bb)) + (f - bb);
    /* two-sum(h1, -hfsq) */
    double h2 = h1 - hfsq;
    double bb2 = h2 - h1;
    double l2 = (h1 - (h2 - bb2)) + (-hfsq - bb2);
    double tail = ((l1 + l2) - hfsq_lo) + (t + kd * JSX_LN2_LO);
    double hi = h2 + tail;
    *lo = tail - (hi - h2);
    return hi;
}

JSX_HD double jsx_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return jsx_from_bits(0x7ff8000000000000ULL);
    if (x == 0.0) return -jsx_from_bits(0x7ff0000000000000ULL);
    if (x == jsx_from_bits(0x7ff0000000000000ULL)) return x;
    double lo;
    double hi = jsx_log_dd(x, &lo);
    return hi + lo;
}

/* exp(y * (lh + ll)) where (lh, ll) = jsx_log_dd(x): lets callers that raise one
 * base to several exponents share the logarithm and still get exactly the
 * bits jsx_pow(x, y) returns. */
JSX_HD double jsx_pow_from_log(double lh, double ll, double y) {
    double ph = y * lh;
    double pl = fma(y, lh, -ph) + y * ll;
    return jsx_exp_dd(ph, pl);
}

/* True when jsx_pow(x, y) takes the generic exp(y log x) route, i.e. when
 * jsx_pow_from_log gives the same bits. */
JSX_HD int jsx_pow_is_generic(double x, double y) {
    return x > 0.0 && x < 1.79769313486231570815e+308 && x == x && y == y &&
           y != 0.0 && y != 1.0 && y != 2.0 && y != -1.0 && x != 1.0;
}

JSX_HD double jsx_pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (y == 1.0) return x;
    if (y == 2.0) return x * x;
    if (y == -1.0) return 1.0 / x;
    const double inf = jsx_from_bits(0x7ff0000000000000ULL);
    if (x == 0.0) return y < 0.0 ? inf : 0.0;
    double sign = 1.0;
    if (x < 0.0) {
        /* negative base: defined for integer exponents, sign from parity */
        if (floor(y) != y) return jsx_from_bits(0x7ff8000000000000ULL);
        double h = 0.5 * y;
        int odd = (floor(h) != h) && (y < 9007199254740992.0) && (y > -9007199254740992.0);
        sign = odd ? -1.0 : 1.0;
        x = -x;
        if (x == 1.0) return sign;
    }
    if (x == inf) return sign * (y > 0.0 ? inf : 0.0);
    double ll;
    double lh = jsx_log_dd(x, &ll);
    return sign * jsx_pow_from_log(lh, ll, y);
}

/* sin(atan(u)) and cos(atan(u)) without the trigonometric detour. */
JSX_HD void jsx_sincos_atan(double u, double *s, double *c) {
    if (u != u) { *s = u; *c = u; return; }
    double au = u < 0.0 ? -u : u;
    if (au <= 1.0) {
        double q = sqrt(1.0 + u * u);
        *s = u / q;
        *c = 1.0 / q;
    } else {
        double v = 1.0 / au;
        double q = sqrt(1.0 + v * v);
        *s = (u < 0.0 ? -1.0 : 1.0) / q;
        *c = v / q;
    }
}

/* sin and cos for |x| < 2^20 (see header) */
JSX_HD void jsx_sincos(double x, double *s, double *c) {
    if (x != x) { *s = x; *c = x; return; }
    const double inv_hpi = 0.63661977236758134308;
    const double hp1 = 1.5707963267341256e+00;   /* pi/2 = hp1 + hp2 + hp3, hp1 and hp2 with 33 bits */
    const double hp2 = 6.0771005065061922e-11;
    const double hp3 = 2.0222662487959506e-21;
    double k = floor(x * inv_hpi + 0.5);
    double r = ((x - k * hp1) - k * hp2) - k * hp3;
    double r2 = r * r;
    /* sin r = r (1 - r^2/3! + r^4/5! - ...), cos r = 1 - r^2/2! + ... */
    double ps = 0.0, pc = 0.0;
    double fs = 1.0, fc = 1.0;
    double ts[14], tc[14];
    for (int n = 0; n < 14; ++n) {
        ts[n] = fs; tc[n] = fc;
        fs = fs / ((double)(2 * n + 2) * (double)(2 * n + 3));
        fc = fc / ((double)(2 * n + 1) * (double)(2 * n + 2));
    }
    for (int n = 13; n >= 0; --n) {
        double sg = (n & 1) ? -1.0 : 1.0;
        ps = fma(ps, r2, sg * ts[n]);
        pc = fma(pc, r2, sg * tc[n]);
    }
    double sr = r * ps, cr = pc;
    long long q = ((long long)k) & 3;
    if (q == 0) { *s = sr; *c = cr; }
    else if (q == 1) { *s = cr; *c = -sr; }
    else if (q == 2) { *s = -sr; *c = -cr; }
    else { *s = -cr; *c = sr; }
}

JSX_HD double jsx_tan(double x) {
    double s, c;
    jsx_sincos(x, &s, &c);
    return s / c;
}

/* asin(y) for |y| <= 1/2 by its Maclaurin series y sum_k c_k y^2k,
 * c_k = (2k)! / (4^k (k!)^2 (2k + 1)), 30 terms (c_29 (1/4)^29 < 1e-19)
 * evaluated by Horner's rule from the smallest term */
JSX_HD double jsx_asin_small(double y) {
    double c[30];
    double a = 1.0;                      /* (2k)! / (4^k (k!)^2) */
    for (int k = 0; k < 30; ++k) {
        c[k] = a / (double)(2 * k + 1);
        a = a * ((double)(2 * k + 1) / (double)(2 * k + 2));
    }
    double y2 = y * y, p = c[29];
    for (int k = 28; k >= 0; --k) p = p * y2 + c[k];
    return y * p;
}

JSX_HD double jsx_acos(double u) {
    const double pi = 3.141592653589793238462643383279502884;
    if (!(u >= -1.0 && u <= 1.0)) return jsx_from_bits(0x7ff8000000000000ULL);
    if (u >= -0.5 && u <= 0.5) return pi / 2.0 - jsx_asin_small(u);
    if (u > 0.5) return 2.0 * jsx_asin_small(sqrt((1.0 - u) / 2.0));
    return pi - 2.0 * jsx_asin_small(sqrt((1.0 + u) / 2.0));
}

/* modified Bessel functions I0(x), I1(x) by their power series */
JSX_HD void jsx_bessel_i01(double x, double *i0, double *i1) {
    double h = 0.5 * x, h2 = h * h;
    double t0 = 1.0, t1 = h, s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < 500; ++k) {
        double n0 = s0 + t0, n1 = s1 + t1;
        if (n0 == s0 && n1 == s1) break;
        s0 = n0; s1 = n1;
        t0 = t0 * h2 / ((double)(k + 1) * (double)(k + 1));
        t1 = t1 * h2 / ((double)(k + 1) * (double)(k + 2));
    }
    *i0 = s0;
    *i1 = s1;
}

#endif /* JSX_MATH_H */
