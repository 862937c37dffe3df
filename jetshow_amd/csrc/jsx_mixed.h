/*
 * jsx_mixed.h -- the mixed-precision solver path.
 *
 * Same algorithm and state machine as the FP64 path (jsx_solver.h), with the
 * precision split chosen for gfx950:
 *
 *   FP64 : ray/surface roots and the segment end points (same code as the FP64
 *          path, so hit masks are identical), the integration variable t, the
 *          step size, the error norm and accept/reject decision, and the
 *          dense-output bisection that places tau = tau_max.
 *   FP32 : every transfer coefficient and every Runge-Kutta stage.  The state
 *          (tau, or I/Q/U/V) is carried as an FP32 value plus an FP32
 *          compensation term (Kahan summation of the step increments).
 *
 * The integration variable is measured in parsec, so FP32 stage values and
 * per-parsec coefficients stay far from the FP32 range limits; results are in
 * the same units as the FP64 path.
 *
 * Coefficients are rebuilt from cancellation-free kinematics:
 *   1 - beta mu = (1 - beta) + beta (1 - mu),  1 - beta = 1 / (Gamma^2 (1 + beta)),
 *   n' = D n_perp + (mu - beta)/(1 - beta mu) v_hat,
 *   b' = b / Gamma + (1 - 1/Gamma)(v_hat . b) v_hat,
 * which equal getD / get_n_los_prime / getBprime (src/utils.cpp:147-168)
 * algebraically.  Powers x^e go through x = m 2^k: log2(x) = k + log2f(m) and
 * the product e log2(x) are formed in FP64, then 2^frac in FP32, which keeps
 * them to ~1e-7 relative instead of the ~1e-6 of a plain FP32 pow.
 */
#ifndef JSX_MIXED_H
#define JSX_MIXED_H

#include "jsx_kernels.h"

/* Multiply-adds of the FP32 coefficient and stage arithmetic may be fused
 * (the translation unit is compiled with -ffp-contract=off so that the FP64
 * path reproduces the oracle bit for bit; the mixed path has no such
 * constraint, and fused FP32 multiply-adds cut its VALU instruction count).
 * The pragma is lexical: the shared geometry and the FP64 step control that
 * these functions call stay uncontracted. */
#if defined(__clang__)
#define JSX_MIXED_FMA _Pragma("clang fp contract(fast)")
#else
#define JSX_MIXED_FMA
#endif

/* lane-internal flag bit, cleared before the flags are written: the optical-
 * depth horizon lies in the second segment of a two-segment path */
#define JSX_PIX_HZ_SEG1 0x80000000u

/* host-derived FP32 constants of the mixed path */
struct MConsts {
    float A_k, A_e, A_kv, A_ev, A_F, A_C;  /* per-parsec coefficient prefactors           */
    float r_kq, r_eq, f_kFr, kC_a, kC_lin;
    float e_ki, e_ei, e_kv;
    float n0, b0, n_negexp, b_negexp;
    float beta_flat, omb_flat, gam_flat;  /* ConstFlat beta, 1 - beta, Gamma              */
    float one_minus_mu;                   /* 1 - n.z for v along z                        */
    float n_los[3], axis[3], sin_pitch, cos_pitch, thermal_fraction;
    double inv_pc;
    double kC_y0, kC_y1;                  /* k_C_r's exponent y = y0 + y1 log2(x)          */
    double tau_abs, tau_rel, st_abs, st_rel, dt_max_pc;
    double st_rel_x;                      /* the caller's Stokes tolerance on |x|          */
    double tau_hz;                        /* optical-depth horizon of the Stokes pass, 0 off */
    int mag_floor;                        /* floor the error scale at the RHS terms' rounding */
    double st_dt_max_pc;                  /* longest Stokes-pass step [pc] (grid fields), 0 none */
    int split_ok;                         /* long rays may go to the time-parallel pieces  */
};

/* x^e for x > 0 (see header) */
__device__ __forceinline__ float m_powf(float x, double e) {
    JSX_MIXED_FMA
    int k;
    float m = frexpf(x, &k);                 /* x = m 2^k, m in [0.5, 1) */
    double l2 = (double)k + (double)__builtin_amdgcn_logf(m);
    double y = e * l2;
    double n = floor(y);
    float f = (float)(y - n);
    return ldexpf(__builtin_amdgcn_exp2f(f), (int)n);
}

__device__ __forceinline__ float m_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

/* grid lookup in FP32 after the FP64 index computation */
template <bool EXT = false>
__device__ inline float m_grid(const DGrid &G, const double *p_cm) { return (float)d_grid_interp<EXT>(G, p_cm); }

struct MCoef { float kI, kQ, kU, kV, kF, kC, hQ, eI, eQ, eU, eV; };

/* p_pc: point in parsec (FP64); p_cm needed only for grid lookups */
template <int FULL, bool EXT = false>
__device__ inline void m_coefficients(const DModel &m, const MConsts &c, const double *p_pc, MCoef &o) {
    JSX_MIXED_FMA
    float px = (float)p_pc[0], py = (float)p_pc[1], pz = (float)p_pc[2];
    double p_cm[3];
    int need_cm = (m.bfield == JSX_B_GRID_SCALAR || m.bfield == JSX_B_GRID_VECTOR || m.nfield == JSX_N_GRID ||
                   m.vfield == JSX_V_GRID || (EXT && m.ext));
    if (need_cm)
        for (int i = 0; i < 3; ++i) p_cm[i] = p_pc[i] * JSX_PC;
    /* ---- bulk velocity: direction v_hat, Gamma, beta, 1 - beta ---- */
    float vh[3] = {0.f, 0.f, 1.f};
    float gam, beta, omb;
    if (m.vfield == JSX_V_CONST_FLAT) {
        gam = c.gam_flat; beta = c.beta_flat; omb = c.omb_flat;
    } else {
        if (m.vfield == JSX_V_LINEAR_GAMMA_Z) {
            double z = p_pc[2] * JSX_PC;
            double g;
            if (z <= m.gz0) g = m.gamma0;
            else if (z >= m.gz1) g = m.gamma1;
            else g = m.gamma0 + (m.gamma1 - m.gamma0) * ((z - m.gz0) / (m.gz1 - m.gz0));
            gam = (float)g;
        } else if (m.vfield == JSX_V_GRID) {
            gam = m.has_gv ? m_grid<EXT>(m.gv, p_cm) : 1.0f;
        } else if (EXT && m.vfield >= JSX_V_SHEARED_FLAT) {
            /* the further reference V fields: velocity from the FP64 field code */
            double v[3];
            d_vfield_ext(m, p_cm, v);
            double vv = d_norm3(v);
            gam = (float)d_getG(v);
            if (vv > 0.0) { vh[0] = (float)(v[0] / vv); vh[1] = (float)(v[1] / vv); vh[2] = (float)(v[2] / vv); }
        } else { /* ConstCentral: radial from -v_origin */
            gam = c.gam_flat;
            float q[3] = {px + (float)(m.v_origin[0] * c.inv_pc), py + (float)(m.v_origin[1] * c.inv_pc),
                          pz + (float)(m.v_origin[2] * c.inv_pc)};
            float rq = __builtin_amdgcn_rsqf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            vh[0] = q[0] * rq; vh[1] = q[1] * rq; vh[2] = q[2] * rq;
        }
        float ig2 = m_rcp(gam * gam);
        beta = __builtin_sqrtf(1.0f - ig2);
        omb = ig2 * m_rcp(1.0f + beta);
        if (m.vfield == JSX_V_CONST_CENTRAL) { beta = c.beta_flat; omb = c.omb_flat; }
    }
    const float *n = c.n_los;
    float mu, one_mu;
    if (m.vfield == JSX_V_CONST_CENTRAL || (EXT && m.vfield >= JSX_V_SHEARED_FLAT)) {
        float d0 = vh[0] - n[0], d1 = vh[1] - n[1], d2 = vh[2] - n[2];
        one_mu = 0.5f * (d0 * d0 + d1 * d1 + d2 * d2);
        mu = 1.0f - one_mu;
    } else {
        one_mu = c.one_minus_mu;
        mu = n[2];
    }
    float ombm = omb + beta * one_mu;          /* 1 - beta mu */
    float D = m_rcp(gam * ombm);
    float npar = ((omb - one_mu)) * m_rcp(ombm); /* (mu - beta)/(1 - beta mu) */
    float nl[3];
    for (int i = 0; i < 3; ++i) nl[i] = D * (n[i] - mu * vh[i]) + npar * vh[i];
    /* ---- density (plasma frame) ---- */
    float rr = 0.f, irr = 0.f;
    int need_r = (m.bfield == JSX_B_RANDOM_SCALAR || m.bfield == JSX_B_TOROIDAL_R || m.bfield == JSX_B_HELICAL_R ||
                  m.nfield == JSX_N_BK);
    if (need_r) {
        rr = __builtin_sqrtf(px * px + py * py + pz * pz);
        irr = m_rcp(rr);
    }
    float nd;
    switch (m.nfield) {
    case JSX_N_CONST: nd = c.n0; break;
    case JSX_N_BK: nd = c.n0 * (m.n_negexp == -2.0 ? irr * irr : (m.n_negexp == -1.0 ? irr : m_powf(rr, m.n_negexp))); break;
    case JSX_N_BK_Z: { float az = fabsf(pz); nd = c.n0 * (m.n_negexp == -2.0 ? m_rcp(az * az) : m_powf(az, m.n_negexp)); break; }
    case JSX_N_GRID: nd = m.has_gn ? m_grid<EXT>(m.gn, p_cm) : 0.f; break;
    default:
        nd = 0.f;
        if constexpr (EXT) {   /* the further reference N fields, from the FP64 field code */
            double r64, lh, ll;
            d_radius_log(p_cm, &r64, &lh, &ll);
            nd = (float)d_nfield<true>(m, p_cm, r64, lh, ll);
        }
    }
    if (!m.n_plasma) nd = nd * m_rcp(gam);
    /* ---- magnetic field (plasma frame) ---- */
    float b[3];
    int vec = 1;
    float bs = 0.f;
    switch (m.bfield) {
    case JSX_B_RANDOM_SCALAR:
        bs = c.b0 * (m.b_negexp == -1.0 ? irr : m_powf(rr, m.b_negexp)); vec = 0; break;
    case JSX_B_RANDOM_SCALAR_Z: {
        float az = fabsf(pz);
        bs = c.b0 * (m.b_negexp == -1.0 ? m_rcp(az) : m_powf(az, m.b_negexp)); vec = 0; break;
    }
    case JSX_B_GRID_SCALAR: bs = m.has_gb ? m_grid<EXT>(m.gb, p_cm) : 0.f; vec = 0; break;
    case JSX_B_TOROIDAL: {
        /* (-sin fi, cos fi) with fi = atan(y/x): sign of x folds into the azimuth */
        float u = py * m_rcp(px);
        float q = __builtin_amdgcn_rsqf(1.0f + u * u);
        float sf = u * q, cf = q;
        float bb = m.b_negexp == -1.0 ? c.b0 * m_rcp(pz) : c.b0 * m_powf(fabsf(pz), m.b_negexp);
        b[0] = -sf * bb; b[1] = cf * bb; b[2] = 0.f;
        break;
    }
    case JSX_B_TOROIDAL_R:
    case JSX_B_HELICAL_R: {
        float bb = c.b0 * (m.b_negexp == -1.0 ? irr : m_powf(rr, m.b_negexp));
        float irho = __builtin_amdgcn_rsqf(px * px + py * py);
        float ex = -py * irho, ey = px * irho;
        if (m.bfield == JSX_B_TOROIDAL_R) { b[0] = bb * ex; b[1] = bb * ey; b[2] = 0.f; }
        else { float bphi = bb * c.sin_pitch; b[0] = bphi * ex; b[1] = bphi * ey; b[2] = bb * c.cos_pitch; }
        break;
    }
    case JSX_B_CONST_CYLINDER_Z: {
        float az = fabsf(pz);
        b[0] = 0.f; b[1] = 0.f; b[2] = c.b0 * (m.b_negexp == 0.0 ? 1.0f : m_powf(az, m.b_negexp));
        break;
    }
    case JSX_B_GRID_VECTOR: {
        float u = py * m_rcp(px);
        float q = __builtin_amdgcn_rsqf(1.0f + u * u);
        float sf = u * q, cf = q;
        float vp = m.has_gb ? m_grid<EXT>(m.gb, p_cm) : 0.f;
        float vf = m.has_gb2 ? m_grid<EXT>(m.gb2, p_cm) : 0.f;
        b[0] = -sf * vf; b[1] = cf * vf; b[2] = vp;
        break;
    }
    default:
        bs = 0.f; vec = 0;
        if constexpr (EXT) {   /* the further reference B fields, from the FP64 field code */
            double r64 = 0.0, lh = 0.0, ll = 0.0, bl[3];
            if (d_need_r(m)) d_radius_log(p_cm, &r64, &lh, &ll);
            vec = d_bfield<true>(m, p_cm, r64, lh, ll, bl);
            if (vec) { b[0] = (float)bl[0]; b[1] = (float)bl[1]; b[2] = (float)bl[2]; }
            else bs = (float)bl[0];
        }
    }
    if (vec && !m.b_plasma) {
        float ig = m_rcp(gam);
        float vb = vh[0] * b[0] + vh[1] * b[1] + vh[2] * b[2];
        float f = (1.0f - ig) * vb;
        for (int i = 0; i < 3; ++i) b[i] = b[i] * ig + f * vh[i];
    }
    /* ---- coefficients ---- */
    float bp, bn = 0.f, cth = 0.f;
    if (vec) {
        float cx = nl[1] * b[2] - nl[2] * b[1];
        float cy = nl[2] * b[0] - nl[0] * b[2];
        float cz = nl[0] * b[1] - nl[1] * b[0];
        bp = __builtin_sqrtf(cx * cx + cy * cy + cz * cz);
        if (FULL) {
            bn = __builtin_sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
            cth = (b[0] * nl[0] + b[1] * nl[1] + b[2] * nl[2]) * m_rcp(bn);
        }
    } else {
        bp = bs;
    }
    /* x = nu_B_perp / nu' = C_B bp D / nu, folded as x = xk * bp * D (xk in A constants) */
    float x = bp * D;                       /* in units of nu / C_B */
    float kI, eI;
    float f_ks = vec ? 1.0f : 0.0f;         /* tangled field: A constants carry the random-B factor */
    (void)f_ks;
    float xs = x > 0.f ? x : 1e-30f;
    int kx;
    float mx = frexpf(xs, &kx);
    double l2 = (double)kx + (double)__builtin_amdgcn_logf(mx);
    /* x^e = 2^(e l2) with the product in FP64 */
    auto pw = [&](double e) -> float {
        double y = e * l2;
        double nn = floor(y);
        return ldexpf(__builtin_amdgcn_exp2f((float)(y - nn)), (int)nn);
    };
    float p_ki = pw(c.e_ki), p_ei = pw(c.e_ei);
    kI = c.A_k * p_ki * nd * x;             /* A_k n bp D x^(s/2) with bp D = x */
    eI = c.A_e * p_ei * nd * bp * D * D;
    if (x <= 0.f) { kI = 0.f; eI = 0.f; }
    o.kI = kI; o.eI = eI;
    if (!FULL || !vec) {
        o.kQ = o.kU = o.kV = o.kF = o.kC = o.hQ = o.eQ = o.eU = o.eV = 0.f;
        return;
    }
    float kq = c.r_kq * kI, eq = c.r_eq * eI;
    float p_kv = pw(c.e_kv);
    float kv = -c.A_kv * nd * bn * D * cth * p_kv;
    float ev = -c.A_ev * nd * bn * D * D * cth * p_ki;
    float kFc = c.A_F * nd * bn * cth * D;
    float kCc = -c.A_C * nd * (bp * D) * (bp * D);
    float kF = c.f_kFr * kFc;
    /* k_C_r = 2/(s-2) (gmin^(2-s) - X^((s-2)/2)) k_C_c with X = nu_B_perp / nu'
       (utils.cpp:93-98), without the cancellation of its two terms for s near
       2: = -(2/(s-2)) gmin^(2-s) expm1(y) k_C_c, y = (s-2)(ln gmin + ln(X)/2)
       formed in FP64 (kC_a = -(2/(s-2)) gmin^(2-s); at s = 2 the limit -2 y') */
    float ykc = (float)(c.kC_y0 + c.kC_y1 * l2);
    float kC = c.kC_a * (c.kC_lin != 0.f ? ykc : expm1f(ykc)) * kCc;
    if (c.thermal_fraction != 0.f) {
        float nth = c.thermal_fraction;
        kF = kF + nth * c.A_F * nd * bn * cth * D;
        kC = kC + nth * (-c.A_C) * nd * (bp * D) * (bp * D);
    }
    const float *a = c.axis;
    float an = a[0] * nl[0] + a[1] * nl[1] + a[2] * nl[2];
    float e1[3] = {a[0] - an * nl[0], a[1] - an * nl[1], a[2] - an * nl[2]};
    float e2[3] = {nl[1] * a[2] - nl[2] * a[1], nl[2] * a[0] - nl[0] * a[2], nl[0] * a[1] - nl[1] * a[0]};
    float bx = b[0] * e1[0] + b[1] * e1[1] + b[2] * e1[2];
    float by = b[0] * e2[0] + b[1] * e2[1] + b[2] * e2[2];
    float q2 = bx * bx + by * by;
    float c2 = 1.f, s2 = 0.f;
    if (q2 > 0.f) { float iq = m_rcp(q2); c2 = -(bx * bx - by * by) * iq; s2 = -(2.f * bx * by) * iq; }
    o.kQ = kq * c2; o.kU = kq * s2; o.kV = kv; o.kF = kF; o.kC = kC * c2; o.hQ = -(kC * s2);
    o.eQ = eq * c2; o.eU = eq * s2; o.eV = ev;
}

/* ---- the mixed lane ---- */
template <int NS>
struct MLane {
    int64_t pix;
    double s0in[3], s0out[3], s1in[3], s1out[3];
    double cin[3], cout[3];                     /* copies of the current segment's ends */
    int nseg, seg, phase, sub, tries, deriv_ok, first;
    double L, t, dt, t_old, tq;                 /* parsec */
    float x[NS], xlo[NS], xt[NS];
    float k1[NS], k2[NS], k3[NS], k4[NS], k5[NS], k6[NS], k7[NS];
    double x_old, k_old;
    double tau_bg, thick;
    uint32_t flags, n_rhs, n_steps;
    float mag, magp;                            /* rounding scales of the I and Q/U/V rows of the RHS at
                                                   the step's start (ml_rhs) */
};

/* floats ml_rhs writes: the NS derivatives and the rounding scale of the I
 * row (and of the Q/U/V rows) */
template <int NS> constexpr int ml_nr() { return NS == 4 ? 6 : 2; }

/* compensated accumulation: (hi, lo) += inc with Knuth's two-sum */
__device__ __forceinline__ void m_kahan_add(float &hi, float &lo, float inc) {
    float y = inc + lo;
    float s = hi + y;
    float bb = s - hi;
    lo = (hi - (s - bb)) + (y - bb);
    hi = s;
}

template <int NS> __device__ inline double ml_cin(const MLane<NS> &L, int i) { return L.cin[i]; }
template <int NS> __device__ inline double ml_cout(const MLane<NS> &L, int i) { return L.cout[i]; }
template <int NS>
__device__ inline void ml_set_seg(MLane<NS> &L, int s) {
    L.seg = s;
    if (s) {
        for (int i = 0; i < 3; ++i) { L.cin[i] = L.s1in[i]; L.cout[i] = L.s1out[i]; }
    } else {
        for (int i = 0; i < 3; ++i) { L.cin[i] = L.s0in[i]; L.cout[i] = L.s0out[i]; }
    }
}
template <int NS>
__device__ inline double ml_seg_length(const MLane<NS> &L) {
    double dv[3] = {ml_cout(L, 0) - ml_cin(L, 0), ml_cout(L, 1) - ml_cin(L, 1), ml_cout(L, 2) - ml_cin(L, 2)};
    return d_norm3(dv);
}

/* h * sum_j c_j k_j, the 5th-order increment of a step of size h */
template <int NS>
__device__ inline float ml_increment(const MLane<NS> &L, int i, double h) {
    JSX_MIXED_FMA
    return (float)h * ((float)DP_C1 * L.k1[i] + (float)DP_C3 * L.k3[i] + (float)DP_C4 * L.k4[i] +
                       (float)DP_C5 * L.k5[i] + (float)DP_C6 * L.k6[i]);
}

template <int NS>
__device__ inline void ml_prepare_stage(MLane<NS> &L, int s) {
    JSX_MIXED_FMA
    float dt = (float)L.dt;
    double t = L.t, dtd = L.dt;
    switch (s) {
    case 2:
        for (int i = 0; i < NS; ++i) L.xt[i] = L.x[i] + dt * ((float)DP_B21 * L.k1[i]);
        L.tq = t + dtd * DP_A2; break;
    case 3:
        for (int i = 0; i < NS; ++i) L.xt[i] = L.x[i] + dt * ((float)DP_B31 * L.k1[i] + (float)DP_B32 * L.k2[i]);
        L.tq = t + dtd * DP_A3; break;
    case 4:
        for (int i = 0; i < NS; ++i)
            L.xt[i] = L.x[i] + dt * ((float)DP_B41 * L.k1[i] + (float)DP_B42 * L.k2[i] + (float)DP_B43 * L.k3[i]);
        L.tq = t + dtd * DP_A4; break;
    case 5:
        for (int i = 0; i < NS; ++i)
            L.xt[i] = L.x[i] + dt * ((float)DP_B51 * L.k1[i] + (float)DP_B52 * L.k2[i] + (float)DP_B53 * L.k3[i] +
                                     (float)DP_B54 * L.k4[i]);
        L.tq = t + dtd * DP_A5; break;
    case 6:
        for (int i = 0; i < NS; ++i)
            L.xt[i] = L.x[i] + dt * ((float)DP_B61 * L.k1[i] + (float)DP_B62 * L.k2[i] + (float)DP_B63 * L.k3[i] +
                                     (float)DP_B64 * L.k4[i] + (float)DP_B65 * L.k5[i]);
        L.tq = t + dtd; break;
    default:
        for (int i = 0; i < NS; ++i) L.xt[i] = L.x[i] + ml_increment(L, i, L.dt);
        L.tq = t + dtd; break;
    }
    L.sub = s;
}

template <int NS>
__device__ inline double ml_dense(const MLane<NS> &L, double t) {
    double dt = (L.t - L.t_old);
    double theta = (t - L.t_old) / dt;
    double X1 = 5.0 * (2558722523.0 - 31403016.0 * theta) / 11282082432.0;
    double X3 = 100.0 * (882725551.0 - 15701508.0 * theta) / 32700410799.0;
    double X4 = 25.0 * (443332067.0 - 31403016.0 * theta) / 1880347072.0;
    double X5 = 32805.0 * (23143187.0 - 3489224.0 * theta) / 199316789632.0;
    double X6 = 55.0 * (29972135.0 - 7076736.0 * theta) / 822651844.0;
    double X7 = 10.0 * (7414447.0 - 829305.0 * theta) / 29380423.0;
    double tm1 = theta - 1.0, tsq = theta * theta;
    double A = tsq * (3.0 - 2.0 * theta), B = tsq * tm1, C = tsq * tm1 * tm1, Dd = theta * tm1 * tm1;
    double b1 = A * DP_C1 - C * X1 + Dd, b3 = A * DP_C3 + C * X3, b4 = A * DP_C4 - C * X4;
    double b5 = A * DP_C5 + C * X5, b6 = A * DP_C6 - C * X6, b7 = B + C * X7;
    return L.x_old + dt * (b1 * L.k_old + b3 * (double)L.k3[0] + b4 * (double)L.k4[0] + b5 * (double)L.k5[0] +
                           b6 * (double)L.k6[0] + b7 * (double)L.k1[0]);
}

template <int NS> __device__ void ml_stokes_seg_begin(MLane<NS> &L, const DSolver &S, const MConsts &c);
template <int NS> __device__ void ml_tau_range_next(MLane<NS> &L, const DSolver &S, const MConsts &c);

template <int NS>
__device__ inline double ml_xval(const MLane<NS> &L, int i) { return (double)L.x[i] + (double)L.xlo[i]; }

template <int NS>
__device__ inline void ml_fail(MLane<NS> &L) {
    L.flags |= JSX_PIX_FAIL_MAXTRIALS;
    if (L.phase == PH_TAU) {
        L.tau_bg = L.tau_bg + ml_xval(L, 0);
        L.thick = L.thick + L.L * JSX_PC;
        for (int i = 0; i < NS; ++i) { L.x[i] = 0.f; L.xlo[i] = 0.f; }
    }
    L.phase = PH_DONE;
}

template <int NS>
__device__ inline void ml_try_begin(MLane<NS> &L, const DSolver &S, const MConsts &c) {
    for (;;) {
        L.n_steps++;
        if (L.phase == PH_TAU && c.dt_max_pc != 0.0 && !(L.dt - c.dt_max_pc <= JSX_DBL_EPS)) {
            L.dt = c.dt_max_pc;
            if (L.tries++ >= S.max_fail) { ml_fail(L); return; }
            continue;
        }
        ml_prepare_stage(L, 2);
        return;
    }
}

template <int NS>
__device__ inline void ml_stokes_begin(MLane<NS> &L, const DSolver &S, const MConsts &c) {
    for (int i = 0; i < NS; ++i) { L.x[i] = 0.f; L.xlo[i] = 0.f; }
    if (!(L.tau_bg > S.tau_min)) {
        L.flags |= JSX_PIX_I_SKIPPED;
        L.phase = PH_DONE;
        return;
    }
    L.phase = PH_STOKES;
    ml_set_seg(L, L.nseg - 1);
    ml_stokes_seg_begin(L, S, c);
}

template <int NS>
__device__ inline void ml_tau_seg_end(MLane<NS> &L, const DSolver &S, const MConsts &c, int found) {
    L.tau_bg = L.tau_bg + ml_xval(L, 0);
    L.thick = L.thick + L.L * JSX_PC;
    if (found || L.seg + 1 >= L.nseg) {
        if (found) L.nseg = L.seg + 1;
        /* the Stokes pass ends (starts, going outwards) at the horizon */
        if (L.flags & JSX_PIX_HORIZON) L.nseg = (L.flags & JSX_PIX_HZ_SEG1) ? 2 : 1;
        ml_stokes_begin(L, S, c);
        return;
    }
    ml_set_seg(L, L.seg + 1);
    L.L = ml_seg_length(L) * c.inv_pc;
    L.x[0] = 0.f; L.xlo[0] = 0.f; L.t = 0.0; L.dt = L.L / S.n; L.deriv_ok = 0;
    ml_tau_range_next(L, S, c);
}

template <int NS>
__device__ void ml_tau_range_next(MLane<NS> &L, const DSolver &S, const MConsts &c) {
    if (!d_less_with_sign(L.t, L.L, L.dt)) { ml_tau_seg_end(L, S, c, 0); return; }
    if (d_less_with_sign(L.L, L.t + L.dt, L.dt)) { L.dt = L.L - L.t; L.deriv_ok = 0; }
    L.t_old = L.t;
    L.tries = 0;
    if (!L.deriv_ok) { L.xt[0] = L.x[0]; L.tq = L.t; L.sub = SUB_INIT; return; }
    ml_try_begin(L, S, c);
}

template <int NS>
__device__ void ml_stokes_loop(MLane<NS> &L, const DSolver &S, const MConsts &c) {
    if (!d_less_with_sign(L.t, L.L, L.dt)) {
        if (L.seg - 1 < 0) { L.phase = PH_DONE; return; }
        ml_set_seg(L, L.seg - 1);
        ml_stokes_seg_begin(L, S, c);
        return;
    }
    if (d_less_with_sign(L.L, L.t + L.dt, L.dt)) L.dt = L.L - L.t;
    L.tries = 0;
    if (L.first) {
        for (int i = 0; i < NS; ++i) L.xt[i] = L.x[i];
        L.tq = L.t; L.sub = SUB_INIT; L.first = 0;
        return;
    }
    ml_try_begin(L, S, c);
}

template <int NS>
__device__ void ml_stokes_seg_begin(MLane<NS> &L, const DSolver &S, const MConsts &c) {
    L.L = ml_seg_length(L) * c.inv_pc;
    double auto_n = S.n;
    /* the first step from the optical depth the pass covers */
    const double tau_pass = (L.flags & JSX_PIX_HORIZON) && c.tau_hz < L.tau_bg ? c.tau_hz : L.tau_bg;
    if (tau_pass > S.tau_n_min) auto_n = d_steps_schedule(S, tau_pass);
    L.dt = L.L / auto_n;
    if (c.st_dt_max_pc > 0.0 && L.dt > c.st_dt_max_pc) L.dt = c.st_dt_max_pc;
    L.t = 0.0;
    L.first = 1;
    ml_stokes_loop(L, S, c);
}

template <int NS, bool EXT = false>
__device__ inline void ml_begin(MLane<NS> &L, const DModel &m, const DImage &im, const DSolver &S, const MConsts &c,
                                int64_t pix) {
    L.pix = pix;
    L.flags = 0; L.n_rhs = 0; L.n_steps = 0;
    L.tau_bg = 0.0; L.thick = 0.0; L.seg = 0; L.mag = 0.f; L.magp = 0.f;
    for (int i = 0; i < NS; ++i) { L.x[i] = 0.f; L.xlo[i] = 0.f; }
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
    ml_set_seg(L, 0);
    L.L = ml_seg_length(L) * c.inv_pc;
    L.t = 0.0; L.dt = L.L / S.n; L.deriv_ok = 0;
    ml_tau_range_next(L, S, c);
}

template <int NS>
__device__ inline void ml_point_pc(const MLane<NS> &L, const DImage &im, const MConsts &c, double *p) {
    if (L.phase == PH_TAU) {
        for (int i = 0; i < 3; ++i) p[i] = ml_cin(L, i) * c.inv_pc + L.tq * im.dir[i];
    } else {
        for (int i = 0; i < 3; ++i) p[i] = ml_cout(L, i) * c.inv_pc - L.tq * im.dir[i];
    }
}

/* the derivative at the stage point: r[0..NS-1]; then the size of the terms
 * the I row is the sum of (|e_I| + sum_j |K_Ij x_j|), and for full Stokes the
 * largest such size of the Q, U, V rows: the scales of their FP32 rounding.
 * Where emission and absorption balance (the Stokes vector near the source
 * function inside an optically thick region) a derivative is small while its
 * rounding is not (see ml_step_error).  r holds ml_nr<NS>() floats. */
template <int NS, bool EXT = false>
__device__ inline void ml_rhs(const MLane<NS> &L, const DModel &m, const DImage &im, const MConsts &c, float *r) {
    JSX_MIXED_FMA
    double p[3];
    ml_point_pc(L, im, c, p);
    MCoef k;
    m_coefficients<NS == 4 ? 1 : 0, EXT>(m, c, p, k);
    if (L.phase == PH_TAU) {
        r[0] = k.kI;
        for (int i = 1; i < ml_nr<NS>(); ++i) r[i] = 0.f;
        /* full Stokes: the decay rate bound of the transfer operator's norm,
           k_I - |eta| (eta = (k_Q, k_U, k_V); rotation does not change the
           norm), integrated beside tau for the optical-depth horizon */
        if constexpr (NS == 4) r[1] = k.kI - sqrtf(k.kQ * k.kQ + k.kU * k.kU + k.kV * k.kV);
        return;
    }
    const float *x = L.xt;
    if (NS == 1) {
        r[0] = k.eI - k.kI * x[0];
        r[1] = c.mag_floor ? fabsf(k.eI) + fabsf(k.kI * x[0]) : 0.f;
        return;
    }
    r[0] = k.eI - k.kI * x[0] - k.kQ * x[1] - k.kU * x[2] - k.kV * x[NS > 3 ? 3 : 0];
    if (NS == 4) {
        r[1] = k.eQ - k.kI * x[1] - k.kQ * x[0] - k.kF * x[2] - k.hQ * x[3];
        r[2] = k.eU - k.kI * x[2] - k.kU * x[0] + k.kF * x[1] - k.kC * x[3];
        r[3] = k.eV - k.kI * x[3] - k.kV * x[0] + k.hQ * x[1] + k.kC * x[2];
        r[4] = r[5] = 0.f;
        if (!c.mag_floor) return;
        const float a0 = fabsf(x[0]), a1 = fabsf(x[1]), a2 = fabsf(x[2]), a3 = fabsf(x[3]);
        const float kI = fabsf(k.kI), kQ = fabsf(k.kQ), kU = fabsf(k.kU), kV = fabsf(k.kV);
        const float kF = fabsf(k.kF), kC = fabsf(k.kC), hQ = fabsf(k.hQ);
        r[4] = fabsf(k.eI) + kI * a0 + kQ * a1 + kU * a2 + kV * a3;
        const float m1 = fabsf(k.eQ) + kI * a1 + kQ * a0 + kF * a2 + hQ * a3;
        const float m2 = fabsf(k.eU) + kI * a2 + kU * a0 + kF * a1 + kC * a3;
        const float m3 = fabsf(k.eV) + kI * a3 + kV * a0 + hQ * a1 + kC * a2;
        r[5] = fmaxf(m1, fmaxf(m2, m3));
    }
}

/* FP64 error norm of the finished step from the FP32 stage derivatives.
 * The scale |x_i| + dt |k1_i| of each of Q, U, V is floored at the largest of
 * the three: a component passing through zero is then controlled relative to
 * the polarized intensity, not to itself, where the FP32 rounding of the large
 * terms that drive it (k_F U, k_C V, ...) would dominate and reject sound
 * steps (a third of all steps on Faraday-thick rays).  `floor` (0 for a
 * whole ray) bounds every scale from below; the pieces of a split ray use it
 * (jsx_split.h): a propagator column starts at a unit vector and multiplies
 * the Stokes vector accumulated before it, so its entries' errors matter in
 * absolute terms (floor 1); a piece's particular solution starts at zero but
 * its error matters relative to the ray's intensity (floor |I| where the ray
 * was split).  Without the floors both would keep resolving decays and
 * build-ups to 1e-6 of their own small size, at a fraction of the step the
 * whole ray takes there.  `rel_min` raises the relative tolerance (pieces).
 * `mag` and `magp` (0 for pieces) are the sizes of the terms the
 * derivatives of the I and of the Q/U/V rows are sums of at the step's start
 * (ml_rhs): the FP32 floor applies to dt * max(|dx/dt|, mag), since the
 * stages' rounding scales with the terms, not with their sum.  Inside an
 * optically thick region the intensity sits at the source function, the
 * derivative is a small difference of emission and absorption and, floored
 * at dt |dx/dt| alone, the rounding held each step to a few hundredths of an
 * optical depth (590k derivative evaluations for a ray through tau 1e4, where
 * FP64 takes 41k). */
template <int NS>
__device__ inline double ml_step_error(const MLane<NS> &L, const MConsts &c, double floor = 0.0, double rel_min = 0.0,
                                       float mag = 0.f, float magp = 0.f) {
    double dt = L.dt, err = 0.0;
    int nl = L.phase == PH_TAU ? 1 : NS;
    /* odeint's norm |xerr| / (abs + rel (|x| + dt |dx/dt|)) with the relative
       tolerance split: the caller's on |x| (st_rel_x), the FP32 floor on the
       step's increment dt |dx/dt| (st_rel), which is what the stages' rounding
       scales with -- so on long optically thin paths, where |x| >> dt |dx/dt|,
       the state is held to the caller's tolerance as in FP64 */
    double ea = L.phase == PH_TAU ? c.tau_abs : c.st_abs, er = L.phase == PH_TAU ? c.tau_rel : c.st_rel;
    double erx = L.phase == PH_TAU ? er : c.st_rel_x;
    er = er > rel_min ? er : rel_min;
    erx = erx > rel_min ? erx : rel_min;
    double xe[NS], sx[NS], sk[NS], px = 0.0, pk = 0.0;
    for (int i = 0; i < nl; ++i) {
        xe[i] = dt * (DP_DC1 * (double)L.k1[i] + DP_DC3 * (double)L.k3[i] + DP_DC4 * (double)L.k4[i] +
                      DP_DC5 * (double)L.k5[i] + DP_DC6 * (double)L.k6[i] + DP_DC7 * (double)L.k7[i]);
        sx[i] = fabs(ml_xval(L, i));
        sk[i] = dt * fabs((double)L.k1[i]);
        if (i > 0) {
            px = sx[i] > px ? sx[i] : px;
            pk = sk[i] > pk ? sk[i] : pk;
        }
    }
    if (L.phase != PH_TAU) {
        const double dm = dt * (double)mag, dmp = dt * (double)magp;
        sk[0] = sk[0] > dm ? sk[0] : dm;
        pk = pk > dmp ? pk : dmp;
    }
    for (int i = 0; i < nl; ++i) {
        double s = i > 0 ? erx * px + er * pk : erx * sx[i] + er * sk[i];
        double sf = er * floor;
        s = s > sf ? s : sf;
        double e = fabs(xe[i]) / (ea + s);
        err = (err < e) ? e : err;
    }
    return err;
}

/* error of the finished step: one lane per ray */
struct MlNoReduce {
    template <int NS>
    __device__ double operator()(const MLane<NS> &L, const MConsts &c) const {
        return ml_step_error(L, c, 0.0, 0.0, L.mag, L.magp);
    }
};

/* whether a step-error functor reads the lane's RHS rounding scales (whole
 * rays: their callers pass ml_rhs's ml_nr<NS>() floats to ml_consume; the
 * pieces of a split ray pass NS) */
template <class Red> constexpr bool ml_red_mag() { return false; }
template <> constexpr bool ml_red_mag<MlNoReduce>() { return true; }

/* step-size controller on the finished step's error norm (odeint's
 * controlled_runge_kutta with the default error checker): reject and shrink
 * when err > 1, else advance t and grow when err < 0.5; returns 1 when the
 * step is accepted */
template <int NS>
__device__ __forceinline__ int ml_control(MLane<NS> &L, const DSolver &S, double maxdt, double err) {
    if (err > 1.0) {
        float fac = 0.9f * __builtin_amdgcn_exp2f(-(1.0f / 3.0f) * __builtin_amdgcn_logf((float)err));
        L.dt = L.dt * (double)(fac > 0.2f ? fac : 0.2f);
        if (maxdt != 0.0 && L.dt > maxdt) L.dt = maxdt;
        return 0;
    }
    L.t = L.t + L.dt;
    if (err < 0.5) {
        double e2 = S.emin > err ? S.emin : err;
        L.dt = L.dt * (double)(0.9f * __builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf((float)e2)));
        if (maxdt != 0.0 && L.dt > maxdt) L.dt = maxdt;
    }
    return 1;
}

template <int NS, class Red = MlNoReduce>
__device__ inline void ml_consume(MLane<NS> &L, const DSolver &S, const DImage &im, const MConsts &c, const float *r,
                                  Red red = Red()) {
    constexpr bool MAG = ml_red_mag<Red>();
    L.n_rhs++;
    int ns_live = L.phase == PH_TAU ? 1 : NS;
    for (int i = 0; i < ns_live; ++i)
        if (r[i] != r[i]) L.flags |= JSX_PIX_NAN;
    switch (L.sub) {
    case SUB_INIT:
        for (int i = 0; i < NS; ++i) L.k1[i] = r[i];
        if constexpr (MAG) {
            L.mag = r[NS];
            if constexpr (NS == 4) L.magp = r[NS + 1];
        }
        if (L.phase == PH_TAU) L.deriv_ok = 1;
        L.tries = 0;
        ml_try_begin(L, S, c);
        return;
    case 2: for (int i = 0; i < NS; ++i) L.k2[i] = r[i]; ml_prepare_stage(L, 3); return;
    case 3: for (int i = 0; i < NS; ++i) L.k3[i] = r[i]; ml_prepare_stage(L, 4); return;
    case 4: for (int i = 0; i < NS; ++i) L.k4[i] = r[i]; ml_prepare_stage(L, 5); return;
    case 5: for (int i = 0; i < NS; ++i) L.k5[i] = r[i]; ml_prepare_stage(L, 6); return;
    case 6: for (int i = 0; i < NS; ++i) L.k6[i] = r[i]; ml_prepare_stage(L, 7); return;
    default: break;
    }
    for (int i = 0; i < NS; ++i) L.k7[i] = r[i];
    /* the step's rounding scale is that at its start (as the floor's
       dt |dx/dt| takes k1): the end's, where a field's kink makes the
       derivative jump, would loosen the control exactly there */
    double err = red(L, c);
    const double h = L.dt;   /* the step just taken; the controller rescales L.dt */
    const int ok = ml_control(L, S, L.phase == PH_TAU ? c.dt_max_pc : c.st_dt_max_pc, err);
    if constexpr (MAG) {
        if (ok) {
            L.mag = r[NS];
            if constexpr (NS == 4) L.magp = r[NS + 1];
        }
    }
    if (L.tries++ >= S.max_fail) { ml_fail(L); return; }
    if (!ok) { ml_try_begin(L, S, c); return; }
    /* accepted: Kahan-compensated state update with the increment x_new - x */
    if (L.phase == PH_TAU) {
        L.x_old = ml_xval(L, 0);
        L.k_old = (double)L.k1[0];
        m_kahan_add(L.x[0], L.xlo[0], ml_increment(L, 0, h));
        L.k1[0] = L.k7[0];
        if constexpr (NS == 4) {
            /* component 1: the decay exponent along the whole path so far
               (not reset between segments) */
            m_kahan_add(L.x[1], L.xlo[1], ml_increment(L, 1, h));
            L.k1[1] = L.k7[1];
        }
        double xv = ml_xval(L, 0);
        if (xv >= S.tau_max) {
            double t0 = L.t_old, t1 = L.t, t_m, x_m = xv;
            int it = 0;
            while (fabs(x_m - S.tau_max) > S.tau_prec) {
                if (++it > JSX_BISECT_CAP) { L.flags |= JSX_PIX_FAIL_BISECT; break; }
                t_m = 0.5 * (t0 + t1);
                x_m = ml_dense(L, t_m);
                if (x_m > S.tau_max) t1 = t_m;
                else t0 = t_m;
            }
            t_m = 0.5 * (t0 + t1);
            /* a horizon placed before this step lies nearer the observer:
               the Stokes pass keeps it */
            if (!(L.flags & JSX_PIX_HORIZON)) {
                double po[3], ci[3] = {ml_cin(L, 0), ml_cin(L, 1), ml_cin(L, 2)};
                d_ray_point(ci, im.d2, t_m * JSX_PC, po);
                for (int i = 0; i < 3; ++i) {
                    if (L.seg) L.s1out[i] = po[i];
                    else L.s0out[i] = po[i];
                    L.cout[i] = po[i];
                }
            }
            L.flags |= JSX_PIX_TAU_CROSSED;
            ml_tau_seg_end(L, S, c, 1);
            return;
        }
        /* optical-depth horizon: the first step end past which everything
           deeper reaches the observer attenuated by e^-tau_hz ends the
           Stokes pass's path -- Stokes I: tau itself; full Stokes: the
           integral of k_I - |eta|, the bound on the transfer operator's
           norm decay, which a growing polarized mode (|eta| > k_I, as the
           reference's coefficients allow near a field along the line of
           sight) holds back.  The tau pass runs on, so tau and l stay those
           of the whole path. */
        const double decay = NS == 4 ? ml_xval(L, NS > 1 ? 1 : 0) : L.tau_bg + xv;
        if (c.tau_hz > 0.0 && !(L.flags & JSX_PIX_HORIZON) && decay >= c.tau_hz) {
            double po[3], ci[3] = {ml_cin(L, 0), ml_cin(L, 1), ml_cin(L, 2)};
            d_ray_point(ci, im.d2, L.t * JSX_PC, po);
            for (int i = 0; i < 3; ++i) {
                if (L.seg) L.s1out[i] = po[i];
                else L.s0out[i] = po[i];
            }
            L.flags |= JSX_PIX_HORIZON | (L.seg ? JSX_PIX_HZ_SEG1 : 0u);
        }
        ml_tau_range_next(L, S, c);
        return;
    }
    for (int i = 0; i < NS; ++i) {
        m_kahan_add(L.x[i], L.xlo[i], ml_increment(L, i, h));
        L.k1[i] = L.k7[i];
    }
    ml_stokes_loop(L, S, c);
}

template <int NS>
__device__ inline void ml_write(const MLane<NS> &L, const DOut &o, int64_t q) {
    uint32_t fl = L.flags & ~JSX_PIX_HZ_SEG1;
    double x0 = ml_xval(L, 0);
    if ((fl & JSX_PIX_HIT) && x0 != x0) fl |= JSX_PIX_NAN;
    if (o.tau) o.tau[q] = L.tau_bg;
    if (o.I) o.I[q] = x0;
    if (o.Q) o.Q[q] = NS == 4 ? ml_xval(L, NS > 1 ? 1 : 0) : 0.0;
    if (o.U) o.U[q] = NS == 4 ? ml_xval(L, NS > 2 ? 2 : 0) : 0.0;
    if (o.V) o.V[q] = NS == 4 ? ml_xval(L, NS > 3 ? 3 : 0) : 0.0;
    if (o.l) o.l[q] = L.thick;
    if (o.flags) o.flags[q] = fl;
    if (o.n_rhs) o.n_rhs[q] = L.n_rhs;
    if (o.n_steps) o.n_steps[q] = L.n_steps;
}

/* ---- kernels ---- */
struct DParamsM {
    DModel m;
    DImage im;
    DSolver S;
    MConsts c;
};

__device__ __forceinline__ const JSX_CONST DParamsM *d_launder_m(const DParamsM *p) {
    const JSX_CONST DParamsM *q = (const JSX_CONST DParamsM *)p;
    JSX_OPAQUE_S(q);
    return q;
}

/* EXT: the model uses a field or geometry kind outside the workloads' set */
template <int NS, bool EXT>
__global__ __launch_bounds__(256) void jsx_km_plain(const DParamsM *__restrict__ P, DOut out, int64_t npix) {
    int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    MLane<NS> L;
    const JSX_CONST DParamsM *q = d_launder_m(P);
    ml_begin<NS, EXT>(L, q->m, q->im, q->S, q->c, pix);
    while (L.phase != PH_DONE) {
        float r[ml_nr<NS>()];
        q = d_launder_m(P);
        ml_rhs<NS, EXT>(L, q->m, q->im, q->c, r);
        q = d_launder_m(P);
        ml_consume(L, q->S, q->im, q->c, r);
    }
    ml_write(L, out, pix);
}

#endif /* JSX_MIXED_H */
