/*
 * lane_host.cpp -- TEST INFRASTRUCTURE ONLY: the solver's device code built
 * for the CPU (with the stand-in hip/hip_runtime.h beside this file) and
 * stepped one lane at a time.  Used by the CPU tests to check the
 * mixed-precision lane and the time-parallel split of long rays against the
 * oracle without a GPU.  Nothing in the product loads this library.
 */
#include "../../jetshow_amd/csrc/jsx_host.hip"

namespace {

std::vector<double> g_costs;   /* (Dormand-Prince, exponential) cost estimates of split rays */

struct HostCtx {
    jsx_ctx *ctx = nullptr;
    HostCtx(int ngrids, const jsx_grid_desc *gd, const double *const *gv) {
        jsx_ctx_create(0, 0, &ctx);
        for (int i = 0; i < ngrids; ++i) {
            int32_t h;
            jsx_grid_upload(ctx, &gd[i], gv[i], &h);   /* handles 1..n in order */
        }
    }
    ~HostCtx() { jsx_ctx_destroy(ctx); }
};

struct Out1 {
    double tau, I, Q, U, V, l;
    uint32_t flags, n_rhs, n_steps;
    DOut view() {
        DOut o;
        o.tau = &tau; o.I = &I; o.Q = &Q; o.U = &U; o.V = &V; o.l = &l;
        o.flags = &flags; o.n_rhs = &n_rhs; o.n_steps = &n_steps;
        return o;
    }
};

void store(const Out1 &s, int64_t i, const jsx_outputs &o) {
    if (o.tau) o.tau[i] = s.tau;
    if (o.I) o.I[i] = s.I;
    if (o.Q) o.Q[i] = s.Q;
    if (o.U) o.U[i] = s.U;
    if (o.V) o.V[i] = s.V;
    if (o.l) o.l[i] = s.l;
    if (o.flags) o.flags[i] = s.flags;
    if (o.n_rhs) o.n_rhs[i] = s.n_rhs;
    if (o.n_steps) o.n_steps[i] = s.n_steps;
}

template <int NS>
void run_fp64(const DParams &q, int64_t pix, Out1 &o) {
    Lane<NS> L;
    lane_begin(L, q.m, q.im, q.S, pix);
    while (L.phase != PH_DONE) {
        double r[NS];
        lane_rhs(L, q.m, q.im, q.S, r);
        lane_consume(L, q.S, q.im, r);
    }
    lane_write(L, o.view(), 0);
}

template <int NS>
void run_mixed(const DParamsM &q, int64_t pix, Out1 &o) {
    MLane<NS> L;
    ml_begin(L, q.m, q.im, q.S, q.c, pix);
    while (L.phase != PH_DONE) {
        float r[ml_nr<NS>()];
        ml_rhs(L, q.m, q.im, q.c, r);
        ml_consume(L, q.S, q.im, q.c, r);
    }
    ml_write(L, o.view(), 0);
}

struct ConstReduce {
    double v;
    double operator()(const MLane<4> &, const MConsts &) const { return v; }
};

/* one piece, its five lanes in lock step (the GPU runs them as a 5-lane group) */
void run_piece(const DParamsM &q, const LongRay &R, const Piece &pc, PieceOut &po) {
    MLane<4> L[JSX_SPLIT_GROUP];
    for (int role = 0; role < JSX_SPLIT_GROUP; ++role) {
        MLane<4> &a = L[role];
        a.pix = 0;
        a.phase = PH_STOKES;
        a.nseg = 1;
        a.seg = 0;
        for (int i = 0; i < 3; ++i) {
            double o = pc.interval ? R.s0out[i] : R.cout[i];
            a.cout[i] = o;
            a.cin[i] = o;
        }
        a.flags = 0; a.n_rhs = 0; a.n_steps = 0;
        a.tau_bg = R.tau_bg; a.thick = 0.0;
        for (int i = 0; i < 4; ++i) {
            a.x[i] = (role == i + 1) ? 1.0f : 0.0f;
            a.xlo[i] = 0.f;
        }
        a.t = pc.ta; a.L = pc.tb; a.dt = pc.dt0; a.first = 1;
        ml_stokes_loop(a, q.S, q.c);
    }
    po.flags = 0;
    while (L[0].phase != PH_DONE) {
        float r[JSX_SPLIT_GROUP][4];
        for (int role = 0; role < JSX_SPLIT_GROUP; ++role)
            split_piece_rhs(L[role], q.m, q.im, q.c, role == 0 ? 1.0f : 0.0f, r[role]);
        double emax = 0.0;
        if (L[0].sub != SUB_INIT && L[0].sub >= 7) {
            for (int role = 0; role < JSX_SPLIT_GROUP; ++role) {
                MLane<4> tmp = L[role];
                for (int i = 0; i < 4; ++i) tmp.k7[i] = r[role][i];
                double e = ml_step_error(tmp, q.c, split_scale_floor(R, role, 1.0f), JSX_SPLIT_REL_MIN);
                e = (e == e) ? e : __builtin_inf();
                emax = e > emax ? e : emax;
            }
        }
        for (int role = 0; role < JSX_SPLIT_GROUP; ++role) ml_consume(L[role], q.S, q.im, q.c, r[role], ConstReduce{emax});
    }
    for (int i = 0; i < 4; ++i) po.P[i] = ml_xval(L[0], i);
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) po.Phi[4 * j + i] = ml_xval(L[j + 1], i);
    for (int role = 0; role < JSX_SPLIT_GROUP; ++role) po.flags |= L[role].flags;
    po.n_rhs = L[0].n_rhs;
    po.n_steps = L[0].n_steps;
}

/* one piece by exponential propagators (the GPU's jsx_k_split_magnus lane) */
void run_piece_magnus(const DParamsM &q, const LongRay &R, const Piece &pc, PieceOut &po) {
    MgLane L;
    static double ustore[JSX_MG_NU * JSX_MG_USTRIDE];
    L.U.p = ustore;
    mg_lane_begin(L, R, pc, JSX_SPLIT_REL_MIN);
    while (!L.done) mg_lane_step(L, q.m, q.im, q.c, q.S);
    for (int i = 0; i < 4; ++i) po.P[i] = L.U[16 + i];
    for (int i = 0; i < 16; ++i) po.Phi[i] = L.U[i];
    po.flags = L.flags;
    po.n_rhs = L.n_rhs;
    po.n_steps = L.n_steps;
}

/* exponential pieces in budgeted rounds, as the launch sequence of
 * jsx_k_split_magnus / jsx_k_split_round runs them (one lane after another) */
void run_pieces_magnus_rounds(const DParamsM &q, const LongRay &R, SplitWS &ws, unsigned budget,
                              bool single = false) {
    /* single: the single-launch mode's decomposition -- every piece, spawned
       ones included, hands on its rest past `budget` right-hand sides, until
       no piece spawns (the GPU runs them in one launch, in any order) */
    unsigned head = 0, end = *ws.n_pieces;
    const int rounds = single ? 1 << 20 : budget ? 4 : 1;
    for (int r = 0; r < rounds && head < end; ++r) {
        const unsigned b = single || r + 1 < rounds ? budget : 0u;
        for (unsigned id = head; id < end; ++id) {
            if (ws.pieces[id].ray < 0) continue;
            MgLane L;
            static double ustore[JSX_MG_NU * JSX_MG_USTRIDE];
            L.U.p = ustore;
            mg_lane_begin(L, R, ws.pieces[id], JSX_SPLIT_REL_MIN);
            int spawn_ok = b != 0u;
            while (!L.done) {
                mg_lane_step(L, q.m, q.im, q.c, q.S);
                const int m = (!L.done && spawn_ok && L.n_rhs >= b) ? mg_spawn_parts(L, b) : 0;
                if (m > 0) {
                    const unsigned mine = atomicAdd(ws.n_pieces, (unsigned)m);
                    if (mine + (unsigned)m <= (unsigned)ws.piece_cap) {
                        mg_spawn_at(ws, (int)id, L, mine, m);
                        L.U[JSX_MG_TB] = L.t;
                        L.done = 1;
                    } else {
                        mg_spawn_void(ws, mine, m);
                        spawn_ok = 0;
                    }
                }
            }
            PieceOut &po = ws.pout[id];
            for (int i = 0; i < 4; ++i) po.P[i] = L.U[16 + i];
            for (int i = 0; i < 16; ++i) po.Phi[i] = L.U[i];
            po.flags = L.flags;
            po.n_rhs = L.n_rhs;
            po.n_steps = L.n_steps;
        }
        head = end;
        end = *ws.n_pieces < (unsigned)ws.piece_cap ? *ws.n_pieces : (unsigned)ws.piece_cap;
    }
}

/* the split pipeline for one ray, as the GPU kernels do it (sampling, piece
 * counts for a given per-piece stiffness, cuts at quantiles, pieces, combine);
 * method 0: Dormand-Prince pieces (5 lanes), 2: exponential propagators,
 * 3: exponential propagators in budgeted rounds (64 right-hand sides),
 * 4: exponential propagators split as the single launch does (48) */
void run_mixed_split(const DParamsM &q, int64_t pix, unsigned defer_rhs, float w_piece, int method, Out1 &o,
                     int *n_pieces) {
    MLane<4> L;
    ml_begin(L, q.m, q.im, q.S, q.c, pix);
    *n_pieces = 0;
    while (L.phase != PH_DONE) {
        float r[ml_nr<4>()];
        ml_rhs(L, q.m, q.im, q.c, r);
        ml_consume(L, q.S, q.im, q.c, r);
        if (L.phase != PH_DONE && split_should_park(L, defer_rhs)) break;
    }
    if (L.phase == PH_DONE) {
        ml_write(L, o.view(), 0);
        return;
    }
    /* park */
    LongRay R;
    unsigned nl = 0;
    std::vector<float> prof(JSX_SPLIT_NSAMP);
    SplitWS ws{};
    ws.n_long = &nl;
    ws.rays = &R;
    ws.cap = 1;
    ws.prof = prof.data();
    split_park(L, 0, ws);
    R.pix = 0;   /* the harness's output view holds this one pixel */
    /* sample (the GPU spreads these over a wavefront's 64 lanes) */
    const double total = (R.L - R.t) + (R.seg ? R.L0 : 0.0);
    const double du = total / JSX_SPLIT_NSAMP;
    const float floor_w = 16.0f / JSX_SPLIT_NSAMP;
    float acc = 0.f, accm = 0.f, acc_rot = 0.f, acc_tau = 0.f;
    const double rel = q.c.st_rel > JSX_SPLIT_REL_MIN ? q.c.st_rel : JSX_SPLIT_REL_MIN;
    MCoef prev;
    for (int kk = 0; kk < JSX_SPLIT_NSAMP; ++kk) {
        double p[3];
        split_point(R, q.im, q.c.inv_pc, (kk + 0.5) * du, p);
        MCoef k;
        m_coefficients<1>(q.m, q.c, p, k);
        float c = split_stiffness(k) * (float)du;
        if (!(c == c) || c > 1e30f) c = 1e30f;
        acc += c + floor_w;
        prof[kk] = acc;
        if (kk > 0) accm += split_mg_density(prev, k, du, rel, fabs(R.x[0])) + floor_w;
        prev = k;
        acc_tau += (fabsf(k.kI) + sqrtf(k.kQ * k.kQ + k.kU * k.kU + k.kV * k.kV)) * (float)du;
        acc_rot += (fabsf(k.kF) + fabsf(k.kC) + fabsf(k.hQ)) * (float)du;
    }
    g_costs.push_back(acc);
    g_costs.push_back(accm);
    g_costs.push_back(acc_rot);
    g_costs.push_back(acc_tau);
    R.w_total = acc;
    if (R.seg) {
        int kb = (int)floor((R.L - R.t) / du);
        if (kb > JSX_SPLIT_NSAMP) kb = JSX_SPLIT_NSAMP;
        R.w_split = kb > 0 ? prof[kb - 1] : 0.f;
    } else {
        R.w_split = acc;
    }
    float np = ceilf(R.w_total / w_piece);
    int k = np < 1.f ? 1 : (np > 65536.f ? 65536 : (int)np);
    if (R.seg) k = k < 2 ? 2 : k;
    R.piece0 = 0;
    R.npieces = k;
    const int cap = method >= 3 ? k + 65536 : k;
    std::vector<Piece> pcs(cap);
    std::vector<PieceOut> pout(cap);
    unsigned n_pcs = (unsigned)k;
    ws.pieces = pcs.data();
    ws.pout = pout.data();
    ws.piece_cap = cap;
    ws.n_pieces = &n_pcs;
    /* cut: same code as jsx_k_split_cut, one "thread" per piece */
    {
        const double restA = R.L - R.t;
        int nA = k;
        if (R.seg) {
            float fa = R.w_total > 0.f ? R.w_split / R.w_total : 0.5f;
            nA = (int)rintf(fa * (float)k);
            nA = nA < 1 ? 1 : (nA > k - 1 ? k - 1 : nA);
        }
        for (int j = 0; j < k; ++j) {
            Piece &pc = pcs[j];
            pc.ray = 0;
            double ua, ub;
            if (j < nA) {
                float wa0 = 0.f, wa1 = R.seg ? R.w_split : R.w_total;
                float wl = wa0 + (wa1 - wa0) * (float)j / (float)nA;
                float wh = wa0 + (wa1 - wa0) * (float)(j + 1) / (float)nA;
                ua = j == 0 ? 0.0 : split_u_of_w(prof.data(), du, wl);
                ub = j + 1 == nA ? restA : split_u_of_w(prof.data(), du, wh);
                if (ua > restA) ua = restA;
                if (ub > restA) ub = restA;
                pc.interval = 0;
                pc.ta = R.t + ua;
                pc.tb = j + 1 == nA ? R.L : R.t + ub;
            } else {
                int kb = j - nA, nB = k - nA;
                float wb0 = R.w_split, wb1 = R.w_total;
                float wl = wb0 + (wb1 - wb0) * (float)kb / (float)nB;
                float wh = wb0 + (wb1 - wb0) * (float)(kb + 1) / (float)nB;
                ua = kb == 0 ? restA : split_u_of_w(prof.data(), du, wl);
                ub = kb + 1 == nB ? total : split_u_of_w(prof.data(), du, wh);
                if (ua < restA) ua = restA;
                if (ub < restA) ub = restA;
                pc.interval = 1;
                pc.ta = kb == 0 ? 0.0 : ua - restA;
                pc.tb = kb + 1 == nB ? R.L0 : ub - restA;
            }
            if (pc.tb < pc.ta) pc.tb = pc.ta;
            pc.dt0 = (j == 0) ? R.dt : (pc.tb - pc.ta) / 100.0;
            if (!(pc.dt0 > 0.0)) pc.dt0 = (pc.tb - pc.ta) > 0.0 ? (pc.tb - pc.ta) : 1e-30;
            pc.next = j + 1 < k ? j + 1 : -1;
            pc.pad = 0;
        }
    }
    if (method >= 3) {
        if (method == 3) run_pieces_magnus_rounds(q, R, ws, 64u);
        else run_pieces_magnus_rounds(q, R, ws, 48u, true);
    } else {
        for (int j = 0; j < k; ++j) {
            if (method == 2) run_piece_magnus(q, R, pcs[j], pout[j]);
            else run_piece(q, R, pcs[j], pout[j]);
        }
    }
    /* combine: the GPU's jsx_k_split_combine for this one ray, after the four
       chain-jumping rounds of jsx_k_split_jump when the rounds made chains */
    std::vector<PieceOut> pout2(cap);
    std::vector<int> na(cap), nb(cap);
    ws.pout2 = pout2.data();
    ws.next_a = na.data();
    ws.next_b = nb.data();
    const int jumped = method >= 3;
    if (jumped)
        for (int r = 0; r < 4; ++r) jsx_k_split_jump(ws, r);
    DOut ov = o.view();
    DOut outs[1] = {ov};
    nl = 1;
    SplitWS w2 = ws;
    unsigned one = 1;
    w2.n_long = &one;
    jsx_k_split_combine(outs, w2, jumped);
    *n_pieces = (int)(n_pcs < (unsigned)cap ? n_pcs : (unsigned)cap);
}

int prepare(HostCtx &H, const jsx_jet_desc *jet, const jsx_image_desc *im, const jsx_solver_params *p, DParams &d,
            DParamsM &dm) {
    int rc;
    if ((rc = lower_model(H.ctx, *jet, p->flags, d.m))) return rc;
    if ((rc = lower_solver(H.ctx, *p, d.S))) return rc;
    if ((rc = lower_image(H.ctx, *im, d.im, nullptr))) return rc;
    dm.m = d.m; dm.im = d.im; dm.S = d.S;
    lower_mixed(*jet, *p, d.m, d.im, dm.c);
    return JSX_OK;
}

}  // namespace

/* Integrate the listed pixels one lane at a time with the kernels' own lane
 * code (FP64 or mixed, per params->precision).  defer_rhs > 0 (mixed full
 * Stokes only) sends rays past that many right-hand sides through the split
 * pipeline with w_piece stiffness per piece and piece method `method` (0:
 * Dormand-Prince, 2: exponential propagators); n_pieces (optional) receives
 * the piece count per pixel (0 = not split). */
extern "C" int lh_observe(const jsx_jet_desc *jet, const jsx_image_desc *im, const jsx_solver_params *p,
                          const int64_t *pix, int64_t n, int32_t ngrids, const jsx_grid_desc *gd,
                          const double *const *gv, unsigned defer_rhs, float w_piece, int32_t method,
                          const jsx_outputs *out, int32_t *n_pieces, int32_t grid_walk) {
    HostCtx H(ngrids, gd, gv);
    H.ctx->grid_walk = grid_walk;
    DParams d;
    DParamsM dm;
    int rc = prepare(H, jet, im, p, d, dm);
    if (rc) return rc;
    g_costs.clear();
    const bool full = p->stokes == JSX_STOKES_FULL;
    for (int64_t i = 0; i < n; ++i) {
        Out1 o{};
        int np = 0;
        if (p->precision == JSX_PREC_MIXED && walk_eligible(H.ctx, JSX_PREC_MIXED, dm.m)) {
            WalkOut w;
            if (full) walk_ray<1>(dm.m, dm.im, dm.S, dm.c, pix[i], w);
            else walk_ray<0>(dm.m, dm.im, dm.S, dm.c, pix[i], w);
            o.tau = w.tau; o.I = w.I; o.Q = o.U = o.V = 0.0; o.l = w.l;
            o.flags = w.flags; o.n_rhs = w.n_rhs; o.n_steps = w.n_steps;
        } else if (p->precision == JSX_PREC_MIXED) {
            if (full && defer_rhs) run_mixed_split(dm, pix[i], defer_rhs, w_piece, method, o, &np);
            else if (full) run_mixed<4>(dm, pix[i], o);
            else run_mixed<1>(dm, pix[i], o);
        } else {
            if (full) run_fp64<4>(d, pix[i], o);
            else run_fp64<1>(d, pix[i], o);
        }
        store(o, i, *out);
        if (n_pieces) n_pieces[i] = np;
    }
    return JSX_OK;
}

/* The mixed path's 11 coefficients (MCoef order kI kQ kU kV kF kC hQ eI eQ eU eV)
 * at points given in cm, through the same lowering as the kernels. */
extern "C" int lh_coefficients_mixed(const jsx_jet_desc *jet, const jsx_image_desc *im, const jsx_solver_params *p,
                                     const double *pts_cm, int64_t n, double *out, int32_t ngrids,
                                     const jsx_grid_desc *gd, const double *const *gv) {
    HostCtx H(ngrids, gd, gv);
    DParams d;
    DParamsM dm;
    int rc = prepare(H, jet, im, p, d, dm);
    if (rc) return rc;
    for (int64_t i = 0; i < n; ++i) {
        double pp[3] = {pts_cm[3 * i] / JSX_PC, pts_cm[3 * i + 1] / JSX_PC, pts_cm[3 * i + 2] / JSX_PC};
        MCoef k;
        m_coefficients<1>(dm.m, dm.c, pp, k);
        const float v[11] = {k.kI, k.kQ, k.kU, k.kV, k.kF, k.kC, k.hQ, k.eI, k.eQ, k.eU, k.eV};
        for (int j = 0; j < 11; ++j) out[11 * i + j] = (double)v[j] / JSX_PC;   /* per parsec -> per cm */
    }
    return JSX_OK;
}

/* exp(-mu 1 + G(eta, rho)) (column-major) and phi1(...) v: the closed forms
 * of the exponential pieces, for comparison with a general matrix exponential. */
extern "C" int lh_magnus_expphi(double mu, const double *eta, const double *rho, const double *v, double *E,
                                double *w, int32_t mode) {
    MgGen g;
    for (int i = 0; i < 3; ++i) { g.e[i] = eta[i]; g.r[i] = rho[i]; }
    /* mode 1: the closed form for every generator; 2: the series where a^2 + b^2 <= 1 */
    return mode == 1 ? mg_expphi<1>(mu, g, v, E, w) : mg_expphi<2>(mu, g, v, E, w);
}

/* [G(e1, r1), G(e2, r2)] as (eta', rho') */
extern "C" void lh_magnus_comm(const double *e1, const double *r1, const double *e2, const double *r2, double *eo,
                               double *ro) {
    MgGen a, b, c;
    for (int i = 0; i < 3; ++i) { a.e[i] = e1[i]; a.r[i] = r1[i]; b.e[i] = e2[i]; b.r[i] = r2[i]; }
    mg_comm(a, b, c);
    for (int i = 0; i < 3; ++i) { eo[i] = c.e[i]; ro[i] = c.r[i]; }
}

/* cost estimates (Dormand-Prince stiffness, exponential steps, rotation,
 * optical depth) of the rays the last lh_observe call split, in order;
 * returns how many rays */
extern "C" int64_t lh_split_costs(double *out, int64_t cap) {
    int64_t n = (int64_t)g_costs.size() / 4;
    for (int64_t i = 0; i < 4 * n && i < 4 * cap; ++i) out[i] = g_costs[i];
    return n;
}

/* the mixed visibility kernel's sine and cosine of 2 pi ph (ph in revolutions) */
extern "C" void lh_vis_sincos(const double *ph, int64_t n, float *s, float *c) {
    for (int64_t i = 0; i < n; ++i) vis_sincos_rev_mixed(ph[i], s[i], c[i]);
}

/* one lane of the mixed visibility kernel, one chunk holding every pixel:
 * FP32 partial sums over JSX_VIS_SUB pixels, FP64 beyond */
extern "C" void lh_vis_mixed(const double *x, const double *y, const double *cre, int64_t n_src, const double *u,
                             const double *v, int64_t n_vis, int32_t isign, double *vis) {
    const double sgn = isign >= 0 ? 1.0 : -1.0;
    for (int64_t k = 0; k < n_vis; ++k) {
        double A = 0.0, B = 0.0;
        for (int64_t jb = 0; jb < n_src; jb += JSX_VIS_SUB) {
            const int64_t je = jb + JSX_VIS_SUB < n_src ? jb + JSX_VIS_SUB : n_src;
            float a = 0.0f, b = 0.0f;
            for (int64_t j = jb; j < je; ++j) {
                const double ph = fma(u[k], x[j], v[k] * y[j]);
                float s, c;
                vis_sincos_rev_mixed(ph, s, c);
                a = fmaf((float)cre[j], c, a);
                b = fmaf((float)cre[j], s, b);
            }
            A += (double)a;
            B += (double)b;
        }
        vis[2 * k] = A;
        vis[2 * k + 1] = sgn * B;
    }
}

/* The request server's scheduler with the stand-in runtime (kernels do not
 * run; events report jsx_hb_event_status): one request completes normally,
 * then the launch's completion event reports a HIP error.  rc[0] = status of
 * the first request, rc[1] = status of the request in the failing launch,
 * rc[2] = a request submitted afterwards (must fail at once), rc[3] = a
 * request queued behind the failure's launch. */
extern "C" int lh_server_error_path(const jsx_jet_desc *jet, const jsx_image_desc *im, const jsx_solver_params *p,
                                    int32_t *rc) {
    jsx_ctx *ctx = nullptr;
    int r = jsx_ctx_create(0, 0, &ctx);
    if (r) return r;
    jsx_server *S = nullptr;
    if ((r = jsx_server_create(ctx, p->precision, p->stokes, 4, 0, 0, &S))) return r;
    const int64_t n = im->n_across * im->n_along;
    std::vector<double> I(n);
    jsx_request q{};
    q.jet = *jet;
    q.image = *im;
    q.params = *p;
    q.out_host.I = I.data();
    uint64_t id = 0;
    jsx_hb_event_status = hipSuccess;
    jsx_server_submit(S, &q, 0, nullptr, &id);
    rc[0] = jsx_server_wait(S, id, nullptr);
    jsx_hb_event_status = hipErrorUnknown;
    uint64_t id1 = 0, id2 = 0;
    jsx_server_submit(S, &q, 0, nullptr, &id1);
    rc[1] = jsx_server_wait(S, id1, nullptr);
    rc[2] = jsx_server_submit(S, &q, 0, nullptr, &id2);
    rc[3] = rc[2] == JSX_OK ? jsx_server_wait(S, id2, nullptr) : rc[2];
    jsx_hb_event_status = hipSuccess;
    jsx_server_destroy(S);
    jsx_ctx_destroy(ctx);
    return JSX_OK;
}
