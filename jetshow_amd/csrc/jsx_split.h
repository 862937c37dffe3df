/*
 * jsx_split.h -- time-parallel integration of Faraday-thick rays (mixed
 * full-Stokes path).
 *
 * The Stokes pass integrates the linear system dS/dt = eps(t) - K(t) S.  Over
 * [a, b] its solution is S(b) = Phi(b, a) S(a) + P(b, a), where P is the
 * solution started from S(a) = 0 and the columns of the 4x4 propagator Phi are
 * the solutions of dS/dt = -K S started from the unit vectors.  Integrating
 * pieces of one ray independently and composing
 *     S_k = Phi_k S_{k-1} + P_k
 * in order gives the same S(end) up to the integration tolerance, so one long
 * ray becomes many short, concurrent ones.  In the image this matters for the
 * rays that graze the apex of a Faraday-thick jet: the explicit Dormand-Prince
 * step is bounded by the rotation rate |k_F| there, and a single ray can need
 * 10^7 right-hand sides -- one lane for minutes while the rest of the GPU idles.
 *
 * Pipeline (all on one stream, no host round trip; counts live in HBM):
 *  1. the batch kernel parks a ray whose right-hand-side count passes
 *     `defer_rhs` at the start of a step of its Stokes pass: state, step
 *     size, position and the remaining path go to a LongRay record;
 *  2. jsx_k_split_sample: one wavefront per parked ray evaluates the transfer
 *     coefficients at NSAMP points of the remaining path and stores the
 *     cumulative stiffness  w(u) = int (k_I + |k_F| + |k_C| + |h_Q| + |k_P|) du
 *     plus a floor, the estimate of the step count;
 *  3. jsx_k_split_alloc (one workgroup): pieces per ray proportional to w
 *     (so every piece costs about the same), prefix offsets;
 *  4. jsx_k_split_cut: piece boundaries at equal-w quantiles of each of the
 *     ray's (one or two) remaining path intervals;
 *  5. pieces, by Dormand-Prince or exponential propagators (option
 *     "piece_layout"; with layout 3 the planner picks per ray from two cost
 *     estimates sampled along the rest of the ray, see jsx_k_split_sample):
 *     jsx_k_split_piece: groups of 5 lanes (particular solution and the four
 *     propagator columns) integrate one piece each by Dormand-Prince, with one
 *     shared step sequence (the group's largest error norm decides);
 *     jsx_k_split_piece20: one lane holds all 20 components, same steps;
 *     jsx_k_split_magnus: one lane accumulates the piece's propagator from
 *     exact exponentials of 4th-order Magnus steps (jsx_magnus.h);
 *  6. jsx_k_split_combine: one lane per parked ray composes its pieces in
 *     order and writes the pixel.
 */
#ifndef JSX_SPLIT_H
#define JSX_SPLIT_H

#include "jsx_magnus.h"

/* coefficient samples per parked ray, 2 per lane (a lane's exponential-step
 * density needs two).  Against 512 samples on the 512^2 paraboloid image: the
 * same method for every ray, 0.3% fewer pieces, 0.03% more coefficient
 * evaluations (the budgeted rounds even out coarser cuts), a quarter of the
 * sampling time; on the 256-model sweep more rays choose exponential pieces
 * (4060 vs 3466) at the same total time. */
#define JSX_SPLIT_NSAMP 128
/* Relative tolerance of the pieces: at least 1e-6.  The pieces' error scales
 * already hold I to ~2e-7 of FP64 on the rays that get split (Faraday-thick,
 * stability-bound steps), where a tighter tolerance would only add steps. */
#define JSX_SPLIT_REL_MIN 1e-6
#define JSX_SPLIT_GROUP 5       /* lanes per piece                                 */

struct LongRay {
    int slot, seg;                /* request index; geometric segment being integrated */
    int64_t pix;
    uint32_t flags, n_rhs, n_steps, pad;
    double tau_bg, thick;
    double t, L, dt;              /* parsec: position, segment length, proposed step */
    double cout[3];               /* origin of the current segment's Stokes pass [cm] */
    double s0out[3], L0;          /* segment 0 (still to do when seg == 1)           */
    double x[4];                  /* Stokes vector at t                              */
    int piece0, npieces;          /* filled by the planner                           */
    float w_total, w_split;       /* cumulative stiffness: total, at the interval end */
    float w_dp, w_mg;             /* cost estimates: Dormand-Prince, exponential      */
    int method;                   /* piece method chosen by the planner (1 or 2)     */
};

struct Piece {
    int ray, interval;            /* interval 0: rest of the current segment; 1: segment 0 */
    double ta, tb, dt0;
    int next, pad;                /* the ray's following piece (-1: last) */
};

static_assert(sizeof(Piece) == 40, "Piece is read as five 8-byte words");

struct PieceOut {
    double P[4];
    double Phi[16];               /* column-major: Phi[4 * j + i] = row i of column j */
    uint32_t flags, n_rhs, n_steps, pad;
};

/* what the piece counter needs of a parked ray, compact (8 B instead of the
 * ray's record): its cost total and method, and whether it spans two segments */
struct RaySum {
    float w;                      /* cumulative cost over the remaining path      */
    int kind;                     /* piece method | 16 if the path has two segments */
};

struct SplitWS {
    unsigned *n_long;             /* parked rays (may exceed cap: then clamp)        */
    LongRay *rays;
    RaySum *rsum;                 /* cap summaries written by the sampler           */
    int cap;
    float *prof;                  /* cap x NSAMP cumulative stiffness               */
    unsigned *n_pieces;           /* pieces planned                                 */
    unsigned *piece_head;         /* next Dormand-Prince piece to integrate         */
    unsigned *dp_end;             /* Dormand-Prince pieces are [0, dp_end)           */
    unsigned *mg_head;            /* next exponential piece to integrate            */
    unsigned *round_end;          /* exponential pieces [.., round_end) are this round's */
    unsigned *n_mg;               /* rays given exponential pieces                  */
    unsigned long long *rhs_pieces; /* right-hand sides of the Dormand-Prince pieces  */
    unsigned long long *rhs_mg;   /* coefficient evaluations of the exponential pieces */
    Piece *pieces;
    PieceOut *pout;
    PieceOut *pout2;              /* second buffer of the chain-jumping rounds      */
    int *next_a, *next_b;         /* chain links of the jumping rounds              */
    int piece_cap;
    int target_pieces;            /* Dormand-Prince pieces the planner aims at      */
    int target_mg;                /* exponential pieces of the first round          */
    int layout;                   /* 0/1: Dormand-Prince only; 2: exponential only;
                                     3: per ray by its rotation                       */
    float mg_cost;                /* layout 3: cost ratio below which a ray's pieces
                                     are exponential                                 */
    int mg_profile;               /* cuts of exponential rays: 0 stiffness, 1 their
                                     own step estimate                               */
    float mg_rel;                 /* smallest relative tolerance of exponential steps */
    int mg_err;                   /* their error estimate: 0 first order, 1 second exponential */
    unsigned *mg_pending;         /* single-launch mode: exponential pieces reserved
                                     and not yet finished                            */
    unsigned *mg_epoch;           /* launch number (jsx_k_split_reset counts it)     */
    unsigned *mg_ready;           /* piece_cap marks: slot written in launch mg_epoch */
    unsigned long long *mg_stat;  /* exponential launches: wavefront loop trips */
    float w_min;                  /* smallest stiffness per piece worth a split     */
    float floor_max;              /* cap of the column error-scale estimate (1: off) */
};

/* sum of a 64-bit value over the wavefront, returned in every lane */
__device__ inline unsigned long long d_wave_sum_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffull), off);
        unsigned hi = __shfl_xor((unsigned)(v >> 32), off);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

/* ---- 1. parking (called by the batch kernel for PolMixed<4>) ---- */
__device__ inline bool split_should_park(const MLane<4> &L, unsigned defer_rhs) {
    return L.phase == PH_STOKES && L.sub == 2 && L.tries == 0 && L.n_rhs > defer_rhs;
}

/* returns false (ray continues in place) when the record table is full */
__device__ inline bool split_park(const MLane<4> &L, int slot, const SplitWS &ws) {
    unsigned k = atomicAdd(ws.n_long, 1u);
    if (k >= (unsigned)ws.cap) return false;
    LongRay &R = ws.rays[k];
    R.slot = slot;
    R.seg = L.seg;
    R.pix = L.pix;
    R.flags = L.flags;
    R.n_rhs = L.n_rhs;
    R.n_steps = L.n_steps - 1;    /* the step just begun is not taken here */
    R.tau_bg = L.tau_bg;
    R.thick = L.thick;
    R.t = L.t;
    R.L = L.L;
    R.dt = L.dt;
    for (int i = 0; i < 3; ++i) {
        R.cout[i] = L.cout[i];
        R.s0out[i] = L.s0out[i];
    }
    double d[3] = {L.s0out[0] - L.s0in[0], L.s0out[1] - L.s0in[1], L.s0out[2] - L.s0in[2]};
    R.L0 = L.seg ? d_norm3(d) * (1.0 / JSX_PC) : 0.0;
    for (int i = 0; i < 4; ++i) R.x[i] = ml_xval(L, i);
    return true;
}

/* u in [0, total) of the concatenated remaining path -> point [pc] */
__device__ inline void split_point(const LongRay &R, const DImage &im, double inv_pc, double u, double *p) {
    double restA = R.L - R.t;
    if (u < restA || !R.seg) {
        double t = R.t + u;
        for (int i = 0; i < 3; ++i) p[i] = R.cout[i] * inv_pc - t * im.dir[i];
    } else {
        double t = u - restA;
        for (int i = 0; i < 3; ++i) p[i] = R.s0out[i] * inv_pc - t * im.dir[i];
    }
}

__device__ inline float split_stiffness(const MCoef &k) {
    float pol = __builtin_sqrtf(k.kQ * k.kQ + k.kU * k.kU + k.kV * k.kV);
    return fabsf(k.kI) + fabsf(k.kF) + fabsf(k.kC) + fabsf(k.hQ) + pol;
}

/* ---- 2. sampling: one wavefront per parked ray ---- */
/* Exponential steps a sample interval of length du would take: a 4th-order
 * Magnus step's error is led by its commutator terms, sqrt(3)/12 h^2 times
 * ([G2, G1], A2 s1 - A1 s2) for the generators and sources at its Gauss
 * points, so over du the step count scales as the cube root of that term for
 * the interval's end samples over the tolerance (sources relative to sP). */
__device__ inline float split_mg_density(const MCoef &a, const MCoef &b, double du, double rel, double sP) {
    JSX_MIXED_FMA
    MgGen ga, gb, cm;
    mg_gen(a, ga);
    mg_gen(b, gb);
    mg_comm(gb, ga, cm);
    double e = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double x = fabs(cm.e[i]), y = fabs(cm.r[i]);
        e = x > e ? x : e;
        e = y > e ? y : e;
    }
    if (sP > 0.0) {
        const double sa[4] = {(double)a.eI, (double)a.eQ, (double)a.eU, (double)a.eV};
        const double sb[4] = {(double)b.eI, (double)b.eQ, (double)b.eU, (double)b.eV};
        double t1[4], t2[4];
        mg_apply(gb, sa, t1);
        mg_apply(ga, sb, t2);
        for (int i = 0; i < 4; ++i) {
            const double d = fabs((-(double)b.kI * sa[i] - t1[i]) - (-(double)a.kI * sb[i] - t2[i])) / sP;
            e = d > e ? d : e;
        }
    }
    const double err = 0.14433756729740643 * du * du * e / rel;
    /* With n steps over du, each step is damped by its own optical depth
     * (errors shrink by e^-mu) and, when it spans several rotations, the
     * commutator term averages over them (~1/(rotation per step)); one
     * fixed-point pass from the undamped count. */
    const double n0 = cbrt(err);
    double n1 = n0;
    if (n0 > 0.0) {
        const double tau = 0.5 * ((double)a.kI + (double)b.kI) * du;
        const double rot = 0.5 * (fabs((double)a.kF) + fabs((double)b.kF) + fabs((double)a.kC) + fabs((double)b.kC) +
                                  fabs((double)a.hQ) + fabs((double)b.hQ)) * du;
        const double mu = tau / n0, bs = rot / n0;
        const double damp = exp(-(mu < 50.0 ? mu : 50.0)) / (bs > 1.0 ? bs : 1.0);
        n1 = cbrt(err * damp);
    }
    const float n = (float)n1;
    return (n == n && n < 1e30f) ? n : 1e30f;
}

/* ---- 2. sampling and method: one wavefront per parked ray ----
 * Two cost densities per sample: the Dormand-Prince stiffness (step count
 * bound by rotation and absorption) and the exponential steps
 * (split_mg_density); the profile of the chosen method's density places the
 * cuts.  Layout 3 chooses per ray.  Dormand-Prince needs about 7 right-hand
 * sides per radian of birefringent rotation (k_F, k_C, h_Q) and 2.6 per unit
 * of absorption depth (k_I + |eta|) on the rays that get split (CPU lane runs
 * of the bench images); an exponential step takes 2 and costs about 1.6
 * Dormand-Prince piece right-hand sides.  Rotation-dominated rays (rotation
 * >= 4 x depth, where the sampled estimate of exponential steps is
 * pessimistic by up to 10x: the steps average over many rotations) always
 * take exponential pieces; the others when 2 w_mg < mg_cost (7 rot + 2.6 tau). */
__global__ __launch_bounds__(256) void jsx_k_split_sample(const DParamsM *__restrict__ reqs, SplitWS ws) {
    const int lane = __lane_id();
    const unsigned r = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    unsigned n = *ws.n_long;
    if (n > (unsigned)ws.cap) n = ws.cap;
    if (r >= n) return;
    LongRay &R = ws.rays[r];
    const DParamsM *q = reqs + R.slot;
    const double total = (R.L - R.t) + (R.seg ? R.L0 : 0.0);
    const int per = JSX_SPLIT_NSAMP / 64;
    const double du = total / JSX_SPLIT_NSAMP;
    const double rel = q->c.st_rel > JSX_SPLIT_REL_MIN ? q->c.st_rel : JSX_SPLIT_REL_MIN;
    const double sP = fabs(R.x[0]);
    float w[JSX_SPLIT_NSAMP / 64], wm[JSX_SPLIT_NSAMP / 64];
    float acc = 0.f, accm = 0.f, rot = 0.f, tau = 0.f;
    /* floor: a piece's step count never drops below what its length needs when
       the coefficients vanish (~16 steps over the whole remaining path) */
    const float floor_w = 16.0f / JSX_SPLIT_NSAMP;
    MCoef prev;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        int kk = lane * per + j;
        double p[3];
        split_point(R, q->im, q->c.inv_pc, (kk + 0.5) * du, p);
        MCoef k;
        m_coefficients<1>(q->m, q->c, p, k);
        float c = split_stiffness(k) * (float)du;
        if (!(c == c) || c > 1e30f) c = 1e30f;
        acc += c + floor_w;
        w[j] = acc;
        const float rr = (fabsf(k.kF) + fabsf(k.kC) + fabsf(k.hQ)) * (float)du;
        const float tt = (fabsf(k.kI) + __builtin_sqrtf(k.kQ * k.kQ + k.kU * k.kU + k.kV * k.kV)) * (float)du;
        rot += (rr == rr && rr < 1e30f) ? rr : 1e30f;
        tau += (tt == tt && tt < 1e30f) ? tt : 1e30f;
        if (j > 0) {
            const float m = split_mg_density(prev, k, du, rel, sP) + floor_w;
            if (j == 1) {
                accm += m;          /* the lane's first interval: its neighbour's density */
                wm[0] = accm;
            }
            accm += m;
            wm[j] = accm;
        }
        prev = k;
    }
    /* the method: both totals over the wavefront */
    float tw = acc, tm = accm;
    for (int off = 32; off > 0; off >>= 1) {
        tw += __shfl_xor(tw, off);
        tm += __shfl_xor(tm, off);
        rot += __shfl_xor(rot, off);
        tau += __shfl_xor(tau, off);
    }
    const bool mg_better = rot >= 4.0f * tau || 2.0f * tm < ws.mg_cost * (7.0f * rot + 2.6f * tau);
    const int method = ws.layout == 2 ? 2 : (ws.layout == 3 ? (mg_better ? 2 : 1) : 1);
    if (lane == 0) {
        R.method = method;
        R.w_dp = tw;
        R.w_mg = tm;
    }
    if (method == 2 && ws.mg_profile) {
#pragma unroll
        for (int j = 0; j < per; ++j) w[j] = wm[j];
        acc = accm;
    }
    /* inclusive scan of the lane totals across the wavefront */
    float tot = acc;
    for (int off = 1; off < 64; off <<= 1) {
        float v = __shfl_up(tot, off);
        if (lane >= off) tot += v;
    }
    float before = tot - acc;
    float *prof = ws.prof + (size_t)r * JSX_SPLIT_NSAMP;
#pragma unroll
    for (int j = 0; j < per; ++j) prof[lane * per + j] = before + w[j];
    float wt = __shfl(tot, 63);
    if (lane == 0) {
        R.w_total = wt;
        ws.rsum[r] = RaySum{wt, method | (R.seg ? 16 : 0)};
    }
    /* stiffness up to the end of the current segment (interval 0): the
       cumulative value of the last sample inside it */
    if (R.seg) {
        int kb = (int)floor((R.L - R.t) / du);
        if (kb > JSX_SPLIT_NSAMP) kb = JSX_SPLIT_NSAMP;
        if (kb <= 0) {
            if (lane == 0) R.w_split = 0.f;
        } else {
#pragma unroll
            for (int j = 0; j < per; ++j)
                if (lane * per + j == kb - 1) R.w_split = before + w[j];
        }
    } else if (lane == 0) {
        R.w_split = wt;
    }
}

/* ---- 3. piece counts (one workgroup) ---- */
__device__ inline int split_count(const RaySum &R, float wp) {
    float np = ceilf(R.w / wp);
    int k = np < 1.f ? 1 : (np > 65536.f ? 65536 : (int)np);
    if (R.kind & 16) k = k < 2 ? 2 : k;
    return k;
}

/* Pieces per ray in proportion to its stiffness, for the Dormand-Prince rays
 * (aiming at target_pieces) and the exponential rays (target_mg) separately;
 * Dormand-Prince pieces take slots [0, dp_end), exponential ones follow. */
__global__ __launch_bounds__(1024) void jsx_k_split_alloc(SplitWS ws) {
    __shared__ float s_sum[2][1024];
    __shared__ float s_wp[2];
    __shared__ int s_tot[1024];
    __shared__ int s_base, s_nmg;
    const int tid = threadIdx.x;
    unsigned n = *ws.n_long;
    if (n > (unsigned)ws.cap) n = ws.cap;
    if (tid == 0) s_nmg = 0;
    __syncthreads();
    float local[2] = {0.f, 0.f};
    int nmg = 0;
    for (unsigned r = tid; r < n; r += blockDim.x) {
        const RaySum R = ws.rsum[r];
        const int m = R.kind & 15;
        local[m == 2] += R.w;
        nmg += m == 2;
    }
    if (nmg) atomicAdd(&s_nmg, nmg);
    s_sum[0][tid] = local[0];
    s_sum[1][tid] = local[1];
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (tid < h) {
            s_sum[0][tid] += s_sum[0][tid + h];
            s_sum[1][tid] += s_sum[1][tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int g = 0; g < 2; ++g) {
            const int tgt = g ? ws.target_mg : ws.target_pieces;
            float wp = s_sum[g][0] / (float)(tgt > 0 ? tgt : 1);
            s_wp[g] = wp > ws.w_min ? wp : ws.w_min;
        }
    }
    __syncthreads();
    /* grow the per-piece stiffness until the pieces fit the table */
    for (int iter = 0; iter < 32; ++iter) {
        int cnt = 0;
        for (unsigned r = tid; r < n; r += blockDim.x) {
            const RaySum R = ws.rsum[r];
            cnt += split_count(R, s_wp[(R.kind & 15) == 2]);
        }
        s_tot[tid] = cnt;
        __syncthreads();
        for (int h = 512; h > 0; h >>= 1) {
            if (tid < h) s_tot[tid] += s_tot[tid + h];
            __syncthreads();
        }
        int total = s_tot[0];
        __syncthreads();
        if (total <= ws.piece_cap) break;
        if (tid == 0) {
            s_wp[0] *= 2.0f;
            s_wp[1] *= 2.0f;
        }
        __syncthreads();
    }
    /* exclusive prefix of the piece counts in ray order, Dormand-Prince rays
       first, then exponential rays: each thread takes a contiguous run of
       rays, one block scan of the run totals (wavefront scans and a scan of
       the 16 wavefront totals) places the runs */
    __shared__ int s_wtot[16];
    __shared__ int s_gtot;
    if (tid == 0) s_base = 0;
    const unsigned per = (n + blockDim.x - 1) / blockDim.x;
    const unsigned ra = tid * per < n ? tid * per : n;
    const unsigned rb = ra + per < n ? ra + per : n;
    const int lane = tid & 63, wv = tid >> 6, nwv = (int)(blockDim.x >> 6);
    __syncthreads();
    for (int g = 0; g < 2; ++g) {
        int run = 0;
        for (unsigned r = ra; r < rb; ++r) {
            const RaySum R = ws.rsum[r];
            if (((R.kind & 15) == 2) == (g == 1)) run += split_count(R, s_wp[g]);
        }
        int incl = run;
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (lane == 63) s_wtot[wv] = incl;
        __syncthreads();
        if (wv == 0) {
            const int t = lane < nwv ? s_wtot[lane] : 0;
            int ti = t;
            for (int off = 1; off < 64; off <<= 1) {
                const int v = __shfl_up(ti, off);
                if (lane >= off) ti += v;
            }
            if (lane < nwv) s_wtot[lane] = ti - t;
            if (lane == nwv - 1) s_gtot = ti;
        }
        __syncthreads();
        int pos = s_base + s_wtot[wv] + incl - run;
        for (unsigned r = ra; r < rb; ++r) {
            const RaySum R = ws.rsum[r];
            if (((R.kind & 15) == 2) == (g == 1)) {
                const int k = split_count(R, s_wp[g]);
                ws.rays[r].piece0 = pos;
                ws.rays[r].npieces = k;
                pos += k;
            }
        }
        __syncthreads();
        if (tid == 0) {
            s_base += s_gtot;
            if (g == 0) *ws.dp_end = (unsigned)s_base;
        }
        __syncthreads();
    }
    if (tid == 0) {
        int tot = s_base;
        const unsigned np = (unsigned)(tot < ws.piece_cap ? tot : ws.piece_cap);
        unsigned de = *ws.dp_end;
        de = de < np ? de : np;
        *ws.dp_end = de;
        *ws.n_pieces = np;
        *ws.mg_head = de;
        *ws.round_end = np;
        *ws.mg_pending = np - de;
        *ws.n_mg = (unsigned)s_nmg;
    }
}

/* u of the cumulative stiffness level w (linear between samples) */
__device__ inline double split_u_of_w(const float *prof, double du, float w) {
    int lo = 0, hi = JSX_SPLIT_NSAMP - 1;
    if (w <= prof[0]) return du * (double)(w / (prof[0] > 0.f ? prof[0] : 1.f));
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (prof[mid] < w) lo = mid + 1;
        else hi = mid;
    }
    float w0 = prof[lo - 1], w1 = prof[lo];
    double f = w1 > w0 ? (double)((w - w0) / (w1 - w0)) : 0.0;
    f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
    return du * ((double)lo + f);
}

/* ---- 4. piece boundaries: one thread per piece slot of a ray ---- */
__global__ __launch_bounds__(256) void jsx_k_split_cut(SplitWS ws) {
    const unsigned r = blockIdx.x;
    unsigned n = *ws.n_long;
    if (n > (unsigned)ws.cap) n = ws.cap;
    if (r >= n) return;
    const LongRay &R = ws.rays[r];
    const int np = R.npieces, p0 = R.piece0;
    if (p0 + np > ws.piece_cap) return;
    const double restA = R.L - R.t;
    const double total = restA + (R.seg ? R.L0 : 0.0);
    const double du = total / JSX_SPLIT_NSAMP;
    const float *prof = ws.prof + (size_t)r * JSX_SPLIT_NSAMP;
    /* pieces of interval 0 in proportion to its stiffness (>= 1 each when seg == 1) */
    int nA = np;
    if (R.seg) {
        float fa = R.w_total > 0.f ? R.w_split / R.w_total : 0.5f;
        nA = (int)rintf(fa * (float)np);
        nA = nA < 1 ? 1 : (nA > np - 1 ? np - 1 : nA);
    }
    for (int k = threadIdx.x; k < np; k += blockDim.x) {
        Piece pc;
        pc.ray = (int)r;
        double ua, ub;
        if (k < nA) {
            float wa0 = 0.f, wa1 = R.seg ? R.w_split : R.w_total;
            float wl = wa0 + (wa1 - wa0) * (float)k / (float)nA;
            float wh = wa0 + (wa1 - wa0) * (float)(k + 1) / (float)nA;
            ua = k == 0 ? 0.0 : split_u_of_w(prof, du, wl);
            ub = k + 1 == nA ? restA : split_u_of_w(prof, du, wh);
            if (ua > restA) ua = restA;
            if (ub > restA) ub = restA;
            pc.interval = 0;
            pc.ta = R.t + ua;
            pc.tb = k + 1 == nA ? R.L : R.t + ub;
        } else {
            int kb = k - nA, nB = np - nA;
            float wb0 = R.w_split, wb1 = R.w_total;
            float wl = wb0 + (wb1 - wb0) * (float)kb / (float)nB;
            float wh = wb0 + (wb1 - wb0) * (float)(kb + 1) / (float)nB;
            ua = kb == 0 ? restA : split_u_of_w(prof, du, wl);
            ub = kb + 1 == nB ? total : split_u_of_w(prof, du, wh);
            if (ua < restA) ua = restA;
            if (ub < restA) ub = restA;
            pc.interval = 1;
            pc.ta = kb == 0 ? 0.0 : ua - restA;
            pc.tb = kb + 1 == nB ? R.L0 : ub - restA;
        }
        if (pc.tb < pc.ta) pc.tb = pc.ta;
        /* first piece keeps the parked step size; the others start from a
           fraction of their length and adapt */
        pc.dt0 = (k == 0) ? R.dt : (pc.tb - pc.ta) / 100.0;
        if (!(pc.dt0 > 0.0)) pc.dt0 = (pc.tb - pc.ta) > 0.0 ? (pc.tb - pc.ta) : 1e-30;
        pc.next = k + 1 < np ? p0 + k + 1 : -1;
        pc.pad = 0;
        ws.pieces[p0 + k] = pc;
        ws.pout[p0 + k].flags = 0u;
    }
}

/* ---- 5. pieces: groups of 5 lanes share one step sequence ---- */

struct SplitGroupMax {
    int lead;
    double floor;
    __device__ double operator()(const MLane<4> &L, const MConsts &c) const {
        double e = ml_step_error(L, c, floor, JSX_SPLIT_REL_MIN);
        double v = (e == e) ? e : __builtin_inf();
        double m = v;
#pragma unroll
        for (int i = 0; i < JSX_SPLIT_GROUP; ++i) {
            double o = __shfl(v, lead + i);
            m = o > m ? o : m;
        }
        return m;
    }
};

/* Error-scale floor of a piece lane (see ml_step_error).  A propagator
 * column starts at a unit vector and multiplies the Stokes vector entering
 * its piece, so its entries are held to an absolute error (floor 1); the
 * particular solution is held relative to the ray's intensity where it was
 * parked.  Column j could carry |I| / |S_j| times more error (it reaches the
 * result scaled by S_j), but S_j entering a later piece is unknown when the
 * pieces run: estimated from the parked state, the loosening breaks where
 * polarization builds up after the park point (Q/U/V errors of 4e-4 of peak
 * I at a cap of 1e3, against 6e-6 unloosened).  `floor_max` caps the ratio;
 * 1 (the default) disables the estimate. */
__device__ inline double split_scale_floor(const LongRay &R, int role, float floor_max) {
    const double I = fabs(R.x[0]);
    if (role == 0) return I;
    if (role == 1 || floor_max <= 1.0f) return 1.0;
    const double sj = fabs(R.x[role - 1]);
    double f = sj > 0.0 ? I / sj : (double)floor_max;
    return f < 1.0 ? 1.0 : (f > floor_max ? (double)floor_max : f);
}

__device__ inline void split_piece_rhs(const MLane<4> &L, const DModel &m, const DImage &im, const MConsts &c,
                                       float emit, float *r) {
    JSX_MIXED_FMA
    double p[3];
    ml_point_pc(L, im, c, p);
    MCoef k;
    m_coefficients<1>(m, c, p, k);
    const float *x = L.xt;
    r[0] = emit * k.eI - k.kI * x[0] - k.kQ * x[1] - k.kU * x[2] - k.kV * x[3];
    r[1] = emit * k.eQ - k.kI * x[1] - k.kQ * x[0] - k.kF * x[2] - k.hQ * x[3];
    r[2] = emit * k.eU - k.kI * x[2] - k.kU * x[0] + k.kF * x[1] - k.kC * x[3];
    r[3] = emit * k.eV - k.kI * x[3] - k.kV * x[0] + k.hQ * x[1] + k.kC * x[2];
}

__global__ __launch_bounds__(256) void jsx_k_split_piece(const DParamsM *__restrict__ reqs, SplitWS ws) {
    const int lane = __lane_id();
    const int grp = lane / JSX_SPLIT_GROUP;
    const int role = lane - grp * JSX_SPLIT_GROUP;
    const int lead = grp * JSX_SPLIT_GROUP;
    const bool in_group = lane < (64 / JSX_SPLIT_GROUP) * JSX_SPLIT_GROUP;
    const unsigned npieces = *ws.dp_end;
    unsigned nrays = *ws.n_long;
    if (nrays > (unsigned)ws.cap) nrays = ws.cap;
    MLane<4> L;
    int have = 0, exhausted = in_group ? 0 : 1, slot = 0;
    unsigned long long rhs_acc = 0;
    int pid = 0;
    double sfloor = 0.0;   /* error-scale floor of this lane's role (see ml_step_error) */
    for (;;) {
        /* group leaders without a piece take the next ones, one atomic per wavefront */
        int want = in_group && role == 0 && !have && !exhausted;
        unsigned long long mask = __ballot(want);
        if (mask) {
            int cnt = __popcll(mask);
            int leader = __ffsll((long long)mask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(ws.piece_head, (unsigned)cnt);
            base = __shfl(base, leader);
            unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
            unsigned id = base + (unsigned)__popcll(below);
            int gwant = __shfl(want, in_group ? lead : lane);
            unsigned gid = __shfl(id, in_group ? lead : lane);
            if (gwant) {
                if (gid >= npieces) {
                    exhausted = 1;
                } else if ((unsigned)ws.pieces[gid].ray >= nrays) {
                    /* never planned (cannot happen with the planner's sizing): skip */
                    have = 0;
                } else {
                    pid = (int)gid;
                    const Piece pc = ws.pieces[pid];
                    const LongRay &R = ws.rays[pc.ray];
                    slot = R.slot;
                    const DParamsM *q = reqs + slot;
                    L.pix = pid;
                    L.phase = PH_STOKES;
                    L.nseg = 1;
                    L.seg = 0;
                    for (int i = 0; i < 3; ++i) {
                        double o = pc.interval ? R.s0out[i] : R.cout[i];
                        L.cout[i] = o;
                        L.cin[i] = o;
                    }
                    L.flags = 0; L.n_rhs = 0; L.n_steps = 0;
                    L.tau_bg = R.tau_bg; L.thick = 0.0;
                    for (int i = 0; i < 4; ++i) {
                        L.x[i] = (role == i + 1) ? 1.0f : 0.0f;
                        L.xlo[i] = 0.f;
                    }
                    L.t = pc.ta;
                    L.L = pc.tb;
                    L.dt = pc.dt0;
                    L.first = 1;
                    sfloor = split_scale_floor(R, role, ws.floor_max);
                    ml_stokes_loop(L, q->S, q->c);
                    have = 1;
                    if (L.phase == PH_DONE) have = 2;   /* empty piece */
                }
            }
        }
        if (!__any(have) && __all(exhausted)) break;
        if (have == 1) {
            float r[4];
            SplitGroupMax red;
            red.lead = lead;
            red.floor = sfloor;
            const float emit = role == 0 ? 1.0f : 0.0f;
            /* pieces are queued ray by ray, so a wavefront's groups nearly
               always serve one request: then read its constants with scalar
               loads, else with vector loads */
            const int s0 = __builtin_amdgcn_readfirstlane(slot);
            if (__all(slot == s0)) {
                const JSX_CONST DParamsM *q = (const JSX_CONST DParamsM *)(reqs + s0);
                JSX_OPAQUE_S(q);
                split_piece_rhs(L, q->m, q->im, q->c, emit, r);
                q = (const JSX_CONST DParamsM *)(reqs + s0);
                JSX_OPAQUE_S(q);
                ml_consume(L, q->S, q->im, q->c, r, red);
            } else {
                const DParamsM *q = reqs + slot;
                JSX_OPAQUE_V(q);
                split_piece_rhs(L, q->m, q->im, q->c, emit, r);
                q = reqs + slot;
                JSX_OPAQUE_V(q);
                ml_consume(L, q->S, q->im, q->c, r, red);
            }
            if (L.phase == PH_DONE) have = 2;
        }
        if (have == 2) {
            PieceOut &o = ws.pout[pid];
            if (role == 0) {
                for (int i = 0; i < 4; ++i) o.P[i] = ml_xval(L, i);
                o.n_rhs = L.n_rhs;
                o.n_steps = L.n_steps;
                rhs_acc += L.n_rhs;
            } else {
                for (int i = 0; i < 4; ++i) o.Phi[4 * (role - 1) + i] = ml_xval(L, i);
            }
            if (L.flags) atomicOr(&o.flags, L.flags);
            have = 0;
        }
    }
    unsigned long long tot = rhs_acc;
    for (int off = 32; off > 0; off >>= 1) {
        unsigned lo = __shfl_xor((unsigned)(tot & 0xffffffffull), off);
        unsigned hi = __shfl_xor((unsigned)(tot >> 32), off);
        tot += ((unsigned long long)hi << 32) | lo;
    }
    if (lane == 0 && tot) atomicAdd(ws.rhs_pieces, tot);
}

/* ---- 5'. pieces, one lane each: P and the four propagator columns as one
 *      20-component state (the other Stokes-state layout).  One coefficient
 *      evaluation per stage serves all five vectors, at the cost of ~5x the
 *      stage storage per lane (held in VGPRs and AGPRs at one wavefront per
 *      SIMD); each piece keeps its own step sequence. ---- */
#define JSX_PIECE_NS 20

/* error norm of a 20-component piece lane: per vector v (components 4v..4v+3)
 * the scales of ml_step_error (Q/U/V floored at the vector's polarized
 * maximum) and the vector's floor (split_scale_floor of role v) */
struct PieceErr20 {
    double floor[5];
    __device__ double operator()(const MLane<JSX_PIECE_NS> &L, const MConsts &c) const {
        const double dt = L.dt;
        double er = c.st_rel > JSX_SPLIT_REL_MIN ? c.st_rel : JSX_SPLIT_REL_MIN;
        const double ea = c.st_abs;
        double err = 0.0;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            double xe[4], sc[4], pmax = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = 4 * v + q;
                xe[q] = dt * (DP_DC1 * (double)L.k1[i] + DP_DC3 * (double)L.k3[i] + DP_DC4 * (double)L.k4[i] +
                              DP_DC5 * (double)L.k5[i] + DP_DC6 * (double)L.k6[i] + DP_DC7 * (double)L.k7[i]);
                sc[q] = fabs(ml_xval(L, i)) + dt * fabs((double)L.k1[i]);
                if (q > 0) pmax = sc[q] > pmax ? sc[q] : pmax;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double sv = q > 0 ? pmax : sc[q];
                sv = sv > floor[v] ? sv : floor[v];
                double e = fabs(xe[q]) / (ea + er * sv);
                e = (e == e) ? e : __builtin_inf();
                err = err < e ? e : err;
            }
        }
        return err;
    }
};

__device__ inline void piece20_rhs(const MLane<JSX_PIECE_NS> &L, const DModel &m, const DImage &im,
                                   const MConsts &c, float *r) {
    JSX_MIXED_FMA
    double p[3];
    ml_point_pc(L, im, c, p);
    MCoef k;
    m_coefficients<1>(m, c, p, k);
    const float *x = L.xt;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const float *y = x + 4 * v;
        float *o = r + 4 * v;
        const float e = v == 0 ? 1.0f : 0.0f;
        o[0] = e * k.eI - k.kI * y[0] - k.kQ * y[1] - k.kU * y[2] - k.kV * y[3];
        o[1] = e * k.eQ - k.kI * y[1] - k.kQ * y[0] - k.kF * y[2] - k.hQ * y[3];
        o[2] = e * k.eU - k.kI * y[2] - k.kU * y[0] + k.kF * y[1] - k.kC * y[3];
        o[3] = e * k.eV - k.kI * y[3] - k.kV * y[0] + k.hQ * y[1] + k.kC * y[2];
    }
}

/* the piece lane's transitions: one Stokes pass over [ta, tb], no segments,
 * no tau pass (the general machine's mutual recursion would keep the 20-
 * component lane in memory) */
__device__ __forceinline__ void p20_loop(MLane<JSX_PIECE_NS> &L) {
    if (!d_less_with_sign(L.t, L.L, L.dt)) { L.phase = PH_DONE; return; }
    if (d_less_with_sign(L.L, L.t + L.dt, L.dt)) L.dt = L.L - L.t;
    L.tries = 0;
    if (L.first) {
#pragma unroll
        for (int i = 0; i < JSX_PIECE_NS; ++i) L.xt[i] = L.x[i];
        L.tq = L.t;
        L.sub = SUB_INIT;
        L.first = 0;
        return;
    }
    L.n_steps++;
    ml_prepare_stage(L, 2);
}

__device__ __forceinline__ void p20_consume(MLane<JSX_PIECE_NS> &L, const DSolver &S, const MConsts &c,
                                            const float *r, const PieceErr20 &red) {
    L.n_rhs++;
#pragma unroll
    for (int i = 0; i < JSX_PIECE_NS; ++i)
        if (r[i] != r[i]) L.flags |= JSX_PIX_NAN;
    switch (L.sub) {
    case SUB_INIT:
#pragma unroll
        for (int i = 0; i < JSX_PIECE_NS; ++i) L.k1[i] = r[i];
        L.n_steps++;
        ml_prepare_stage(L, 2);
        return;
    case 2:
#pragma unroll
        for (int i = 0; i < JSX_PIECE_NS; ++i) L.k2[i] = r[i];
        ml_prepare_stage(L, 3); return;
    case 3:
#pragma unroll
        for (int i = 0; i < JSX_PIECE_NS; ++i) L.k3[i] = r[i];
        ml_prepare_stage(L, 4); return;
    case 4:
#pragma unroll
        for (int i = 0; i < JSX_PIECE_NS; ++i) L.k4[i] = r[i];
        ml_prepare_stage(L, 5); return;
    case 5:
#pragma unroll
        for (int i = 0; i < JSX_PIECE_NS; ++i) L.k5[i] = r[i];
        ml_prepare_stage(L, 6); return;
    case 6:
#pragma unroll
        for (int i = 0; i < JSX_PIECE_NS; ++i) L.k6[i] = r[i];
        ml_prepare_stage(L, 7); return;
    default: break;
    }
#pragma unroll
    for (int i = 0; i < JSX_PIECE_NS; ++i) L.k7[i] = r[i];
    const double h = L.dt;
    const int ok = ml_control(L, S, 0.0, red(L, c));
    if (L.tries++ >= S.max_fail) {
        L.flags |= JSX_PIX_FAIL_MAXTRIALS;
        L.phase = PH_DONE;
        return;
    }
    if (!ok) {
        L.n_steps++;
        ml_prepare_stage(L, 2);
        return;
    }
#pragma unroll
    for (int i = 0; i < JSX_PIECE_NS; ++i) {
        m_kahan_add(L.x[i], L.xlo[i], ml_increment(L, i, h));
        L.k1[i] = L.k7[i];
    }
    p20_loop(L);
}

__global__ __launch_bounds__(256) void jsx_k_split_piece20(const DParamsM *__restrict__ reqs, SplitWS ws) {
    const int lane = __lane_id();
    const unsigned npieces = *ws.dp_end;
    unsigned nrays = *ws.n_long;
    if (nrays > (unsigned)ws.cap) nrays = ws.cap;
    MLane<JSX_PIECE_NS> L;
    PieceErr20 red;
    int have = 0, exhausted = 0, slot = 0, pid = 0;
    unsigned long long rhs_acc = 0;
    for (;;) {
        int want = !have && !exhausted;
        unsigned long long mask = __ballot(want);
        if (mask) {
            int cnt = __popcll(mask);
            int leader = __ffsll((long long)mask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(ws.piece_head, (unsigned)cnt);
            base = __shfl(base, leader);
            unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
            unsigned id = base + (unsigned)__popcll(below);
            if (want) {
                if (id >= npieces) {
                    exhausted = 1;
                } else if ((unsigned)ws.pieces[id].ray < nrays) {
                    pid = (int)id;
                    const Piece pc = ws.pieces[pid];
                    const LongRay &R = ws.rays[pc.ray];
                    slot = R.slot;
                    const DParamsM *q = reqs + slot;
                    L.pix = pid;
                    L.phase = PH_STOKES;
                    L.nseg = 1;
                    L.seg = 0;
                    for (int i = 0; i < 3; ++i) {
                        double o = pc.interval ? R.s0out[i] : R.cout[i];
                        L.cout[i] = o;
                        L.cin[i] = o;
                    }
                    L.flags = 0; L.n_rhs = 0; L.n_steps = 0;
                    L.tau_bg = R.tau_bg; L.thick = 0.0;
#pragma unroll
                    for (int i = 0; i < JSX_PIECE_NS; ++i) {
                        L.x[i] = (i >= 4 && (i & 3) == (i >> 2) - 1) ? 1.0f : 0.0f;   /* columns e_0..e_3 */
                        L.xlo[i] = 0.f;
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) red.floor[v] = split_scale_floor(R, v, ws.floor_max);
                    L.t = pc.ta;
                    L.L = pc.tb;
                    L.dt = pc.dt0;
                    L.first = 1;
                    (void)q;
                    p20_loop(L);
                    have = L.phase == PH_DONE ? 2 : 1;
                }
            }
        }
        if (!__any(have) && __all(exhausted)) break;
        if (have == 1) {
            float r[JSX_PIECE_NS];
            const int s0 = __builtin_amdgcn_readfirstlane(slot);
            if (__all(slot == s0)) {
                const JSX_CONST DParamsM *q = (const JSX_CONST DParamsM *)(reqs + s0);
                JSX_OPAQUE_S(q);
                piece20_rhs(L, q->m, q->im, q->c, r);
                q = (const JSX_CONST DParamsM *)(reqs + s0);
                JSX_OPAQUE_S(q);
                p20_consume(L, q->S, q->c, r, red);
            } else {
                const DParamsM *q = reqs + slot;
                JSX_OPAQUE_V(q);
                piece20_rhs(L, q->m, q->im, q->c, r);
                q = reqs + slot;
                JSX_OPAQUE_V(q);
                p20_consume(L, q->S, q->c, r, red);
            }
            if (L.phase == PH_DONE) have = 2;
        }
        if (have == 2) {
            PieceOut &o = ws.pout[pid];
#pragma unroll
            for (int i = 0; i < 4; ++i) o.P[i] = ml_xval(L, i);
#pragma unroll
            for (int i = 0; i < 16; ++i) o.Phi[i] = ml_xval(L, 4 + i);
            o.flags = L.flags;
            o.n_rhs = L.n_rhs;
            o.n_steps = L.n_steps;
            rhs_acc += L.n_rhs;
            have = 0;
        }
    }
    unsigned long long tot = d_wave_sum_u64(rhs_acc);
    if (lane == 0 && tot) atomicAdd(ws.rhs_pieces, tot);
}

/* ---- 5c. pieces by exponential propagators (one lane per piece) ---- */
/* A piece's propagator and particular solution (20 doubles), with its pass
 * origin, end and error-scale floor (5 more), live outside the registers: in
 * the kernel, in LDS, component-major across the 256 lanes of the workgroup
 * (U[i] of lane l at [i][l]: consecutive lanes, consecutive addresses; 50 KB
 * per workgroup).  Held in registers they kept the step's coefficient and
 * exponential code at one wavefront per SIMD (256 VGPRs + ~70 AGPRs); the
 * step reads them once per step. */
#define JSX_MG_USTRIDE 256
/* per-lane doubles in LDS: Phi (16), P (4), the pass origin (3), the
 * piece's end and the particular solution's error-scale floor */
#define JSX_MG_O 20
#define JSX_MG_TB 23
#define JSX_MG_PF 24
#define JSX_MG_NU 25
struct MgU {
    double *p;                    /* component 0 of this lane's 20 */
    __device__ double &operator[](int i) const { return p[i * JSX_MG_USTRIDE]; }
};
struct MgLane {
    MgU U;                        /* Phi column-major (U[4 j + i]), then P (U[16 + i]),
                                     the pass origin [cm] (U[JSX_MG_O + i]), the
                                     piece's end (U[JSX_MG_TB]), p_floor (U[JSX_MG_PF]) */
    double t, h, rel_min;
    uint32_t flags, n_rhs, n_steps, tries;
    int done;
    int est;                      /* error estimate of mg_step (0 first order, 1 second exponential) */
};

__device__ inline void mg_lane_begin(MgLane &L, const LongRay &R, const Piece &pc, double rel_min, int est = 0) {
    L.rel_min = rel_min;
    L.est = est;
#pragma unroll
    for (int i = 0; i < 20; ++i) L.U[i] = (i < 16 && (i >> 2) == (i & 3)) ? 1.0 : 0.0;
    for (int i = 0; i < 3; ++i) L.U[JSX_MG_O + i] = pc.interval ? R.s0out[i] : R.cout[i];
    L.t = pc.ta;
    L.U[JSX_MG_TB] = pc.tb;
    const double len = pc.tb - pc.ta;
    L.h = pc.dt0 > 0.0 ? pc.dt0 : 0.0625 * len;
    L.U[JSX_MG_PF] = fabs(R.x[0]);
    L.flags = 0; L.n_rhs = 0; L.n_steps = 0; L.tries = 0;
    L.done = !(pc.tb > pc.ta);
}

/* one attempted step: accept (propagator updated, t advanced) or shrink h.
 * (The tolerance floor and the estimate's kind are kept per lane: passed as
 * launch-uniform values they let the compiler duplicate the step code, which
 * tripled the kernel's register spills.) */
__device__ inline void mg_lane_step(MgLane &L, const DModel &m, const DImage &im, const MConsts &c, const DSolver &S) {
    JSX_MIXED_FMA
    double h = L.h;
    int last = 0;
    if (!(L.t + h < L.U[JSX_MG_TB])) { h = L.U[JSX_MG_TB] - L.t; last = 1; }
    /* the two Gauss points (kept as two named values: an indexed pair would
       live in scratch memory) */
    MCoef A0, A1;
    {
        const double t0 = L.t + 0.21132486540518713 * h, t1 = L.t + 0.78867513459481287 * h;
        double p0[3], p1[3];
        for (int i = 0; i < 3; ++i) {
            p0[i] = L.U[JSX_MG_O + i] * c.inv_pc - t0 * im.dir[i];
            p1[i] = L.U[JSX_MG_O + i] * c.inv_pc - t1 * im.dir[i];
        }
        m_coefficients<1>(m, c, p0, A0);
        m_coefficients<1>(m, c, p1, A1);
    }
    L.n_rhs += 2;
    const double rel = c.st_rel > L.rel_min ? c.st_rel : L.rel_min;
    double E[16], w[4], err;
    const double P[4] = {L.U[16], L.U[17], L.U[18], L.U[19]};
    const int ok = mg_step(A0, A1, h, rel, L.U[JSX_MG_PF], P, E, w, &err, L.est);
    if (ok && err <= 1.0) {
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const double u[4] = {L.U[4 * j], L.U[4 * j + 1], L.U[4 * j + 2], L.U[4 * j + 3]};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double acc = E[i] * u[0] + E[4 + i] * u[1] + E[8 + i] * u[2] + E[12 + i] * u[3];
                L.U[4 * j + i] = j == 4 ? acc + w[i] : acc;
            }
        }
        L.t = last ? L.U[JSX_MG_TB] : L.t + h;
        L.n_steps++;
        L.tries = 0;
        float fac = 2.0f;
        if (err > 0.0) {
            fac = 0.9f * __builtin_amdgcn_exp2f(-(1.0f / 3.0f) * __builtin_amdgcn_logf((float)err));
            fac = fac < 2.0f ? fac : 2.0f;
        }
        L.h = h * (double)fac;
        if (last) L.done = 1;
    } else {
        float fac = 0.2f;
        if (ok && err == err) {
            fac = 0.9f * __builtin_amdgcn_exp2f(-(1.0f / 3.0f) * __builtin_amdgcn_logf((float)err));
            fac = fac > 0.2f ? fac : 0.2f;
        }
        L.h = h * (double)fac;
        if (++L.tries > (uint32_t)S.max_fail) {
            L.flags |= JSX_PIX_FAIL_MAXTRIALS;
            L.done = 1;
        }
    }
}

/* A piece that has used `budget` right-hand sides (0: no budget) and would
 * need more than `budget` again at its current step size stops where it is
 * and hands the rest of its interval, cut into up to JSX_MG_SPAWN equal parts
 * of about `budget` right-hand sides each, to the next round as new pieces
 * chained after it; the last round runs without a budget.  Cost per unit length varies by orders of magnitude along
 * a Faraday-thick ray (most steps sit in the first per-mille of the path past
 * the apex), so the planner's sampled cuts leave some pieces far longer than
 * the rest; the rounds bound the longest piece a lane holds. */
#define JSX_MG_SPAWN 8

/* a piece another wavefront of the running launch wrote: its five 8-byte
 * words read at the device's coherence point (no cache invalidation) */
__device__ inline Piece mg_piece_load(const Piece *p) {
    const unsigned long long *w = (const unsigned long long *)p;
    unsigned long long v[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) v[i] = __hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    Piece c;
    __builtin_memcpy(&c, v, sizeof c);
    return c;
}

/* and the writer's side: five 8-byte stores at the coherence point */
__device__ inline void mg_piece_store(Piece *p, const Piece &c) {
    unsigned long long v[5];
    __builtin_memcpy(v, &c, sizeof c);
    unsigned long long *w = (unsigned long long *)p;
#pragma unroll
    for (int i = 0; i < 5; ++i) __hip_atomic_store(w + i, v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* write the continuation pieces at slots base.. and chain them after pid */
__device__ inline void mg_spawn_at(const SplitWS &ws, int pid, const MgLane &L, unsigned base, int m,
                                    bool coherent = false) {
    /* in the single launch the piece may have been written by another
       wavefront in this launch: read at the coherence point, as taken */
    const Piece self = coherent ? mg_piece_load(ws.pieces + pid) : ws.pieces[pid];
    const double len = (L.U[JSX_MG_TB] - L.t) / (double)m;
    for (int i = 0; i < m; ++i) {
        Piece c;
        c.ray = self.ray;
        c.interval = self.interval;
        c.ta = L.t + (double)i * len;
        c.tb = i + 1 == m ? L.U[JSX_MG_TB] : L.t + (double)(i + 1) * len;
        c.dt0 = L.h;
        c.next = i + 1 == m ? self.next : (int)base + i + 1;
        c.pad = 0;
        if (coherent) mg_piece_store(ws.pieces + base + i, c);
        else ws.pieces[base + i] = c;
        ws.pout[base + i].flags = 0u;
    }
    ws.pieces[pid].next = (int)base;
}

/* slots reserved past the table's end by a failed reservation: the part
 * inside the table is marked empty (no ray), so a later round skips it */
__device__ inline void mg_spawn_void(const SplitWS &ws, unsigned base, int m, bool coherent = false) {
    for (unsigned i = base; i < base + (unsigned)m && i < (unsigned)ws.piece_cap; ++i) {
        if (coherent) __hip_atomic_store(&ws.pieces[i].ray, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else ws.pieces[i].ray = -1;
    }
}

/* continuation parts a piece asks for: the right-hand sides it still needs at
 * its current step size, in parts of about `budget` (0 or 1: go on alone) */
__device__ inline int mg_spawn_parts(const MgLane &L, unsigned budget) {
    const double est = 2.0 * (L.U[JSX_MG_TB] - L.t) / L.h;
    if (!(est > 2.0 * (double)budget)) return 0;
    const double m = ceil(est / (double)budget);
    return m < (double)JSX_MG_SPAWN ? (int)m : JSX_MG_SPAWN;
}

/* Two wavefronts per SIMD (at most 256 registers per lane).
 *
 * Rounds mode (part == 0): the launch integrates the pieces [mg_head,
 * round_end); with a budget, a piece past it hands the rest it estimates at
 * more than two budgets to the next launch (mg_spawn_parts).
 *
 * Single-launch mode (part > 0): the same rule with `part` right-hand sides
 * in place of the budget, but the parts run in the same launch, so the
 * launch has one tail instead of one per round.  An idle lane takes a
 * ticket, the next slot number of the piece table.  A slot below round_end
 * holds a planned piece; a slot past it becomes a piece when a running lane
 * reserves it (n_pieces), stores it at the coherence point (mg_piece_store)
 * and, after vmcnt(0), marks it with this launch's epoch; the ticket's
 * holder loads the mark, then the piece (mg_piece_load).  No cache is
 * flushed or invalidated on either side.  mg_pending counts slots reserved
 * and not finished: a wavefront adds the slots it reserves before marking
 * them, and takes off the pieces it finished in the same atomic or later, so
 * the count never drops below the slots still to finish; mg_pending = 0 means
 * no piece runs and none can appear, and a lane whose ticket is unfilled then,
 * or lies past the table, is done.  A wait is only ever on a lane that is
 * running (a spawner writes and marks its slots in the step that reserves
 * them), so the launch drains whatever its residency.  Wavefronts with no
 * piece running poll the shared counters every fourth trip and sleep in
 * between: polling them in every trip made them the launch's bottleneck. */
__global__ __launch_bounds__(256, 2) void jsx_k_split_magnus(const DParamsM *__restrict__ reqs, SplitWS ws,
                                                             unsigned budget, unsigned part) {
    const int lane = __lane_id();
    const unsigned npieces = *ws.round_end;
    unsigned nrays = *ws.n_long;
    unsigned *head = ws.mg_head;
    const unsigned cap = (unsigned)ws.piece_cap;
    const unsigned epoch = *ws.mg_epoch;
    if (nrays > (unsigned)ws.cap) nrays = ws.cap;
    MgLane L;
    __shared__ double sU[JSX_MG_NU * JSX_MG_USTRIDE];
    L.U.p = &sU[threadIdx.x];
    int have = 0, exhausted = 0, slot = 0, spawn_ok = 0;
    int holding = 0;              /* a ticket not yet filled */
    unsigned ticket = 0;          /* the slot taken: the piece's, once filled */
    unsigned long long rhs_acc = 0;
    /* single-launch mode: pieces this wavefront finished and has not yet
       taken off mg_pending (taking them off late only delays the end) */
    int dec = 0;
    unsigned it = 0;
    for (;;) {
        ++it;
        int want = !have && !exhausted && !holding;
        unsigned long long mask = __ballot(want);
        if (mask) {
            int cnt = __popcll(mask);
            int leader = __ffsll((long long)mask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(head, (unsigned)cnt);
            base = __shfl(base, leader);
            unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
            if (want) {
                ticket = base + (unsigned)__popcll(below);
                holding = 1;
            }
        }
        if (__any(holding)) {
            int fill = 0, fresh = 0;
            /* a wavefront with no piece running polls every fourth time */
            const int idle = !__any(have == 1);
            const int look = !idle || (it & 3u) == 0u;
            if (holding) {
                if (part == 0u) {
                    if (ticket < npieces) fill = 1;
                    else exhausted = 1;
                } else if (ticket >= cap) {
                    exhausted = 1;
                } else if (ticket < npieces) {
                    fill = 1;
                } else if (look) {
                    fill = fresh = __hip_atomic_load(ws.mg_ready + ticket, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT) == epoch;
                }
                if (exhausted) holding = 0;
            }
            int empty = 0;
            if (fill) {
                holding = 0;
                const Piece pc = fresh ? mg_piece_load(ws.pieces + ticket) : ws.pieces[ticket];
                if ((unsigned)pc.ray < nrays) {
                    const LongRay &R = ws.rays[pc.ray];
                    slot = R.slot;
                    mg_lane_begin(L, R, pc, (double)ws.mg_rel, ws.mg_err);
                    have = L.done ? 2 : 1;
                    spawn_ok = budget != 0u || part != 0u;
                } else {
                    empty = part != 0u;
                }
            }
            dec += __popcll(__ballot(empty));
            /* single-launch mode, a wavefront with no piece running: its own
               count goes in, then a ticket whose slot is not written is done
               if nothing runs anywhere */
            if (part && !__any(have == 1)) {
                if (dec && lane == 0) atomicSub(ws.mg_pending, (unsigned)dec);
                dec = 0;
                if (look && __any(holding)) {
                    const unsigned pend = __hip_atomic_load(ws.mg_pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (holding && __builtin_amdgcn_readfirstlane(pend) == 0u) {
                        holding = 0;
                        exhausted = 1;
                    }
                }
            }
        }
        if (!__any(have) && __all(exhausted)) {
            if (dec && lane == 0) atomicSub(ws.mg_pending, (unsigned)dec);
            break;
        }
        if (have == 1) {
            const int s0 = __builtin_amdgcn_readfirstlane(slot);
            if (__all(slot == s0)) {
                const JSX_CONST DParamsM *q = (const JSX_CONST DParamsM *)(reqs + s0);
                JSX_OPAQUE_S(q);
                mg_lane_step(L, q->m, q->im, q->c, q->S);
            } else {
                const DParamsM *q = reqs + slot;
                JSX_OPAQUE_V(q);
                mg_lane_step(L, q->m, q->im, q->c, q->S);
            }
            if (L.done) have = 2;
        } else if (part && !__any(have)) {
            __builtin_amdgcn_s_sleep(16);    /* every lane waits on a ticket */
        }
        /* continuation pieces: one reservation per wavefront */
        int spawned = 0, reserved = 0;
        unsigned sbase = 0;
        {
            /* a piece past its budget (rounds) or its part (single launch)
               hands on the rest it estimates at more than two of them */
            const unsigned lim = part ? part : budget;
            int m = (have == 1 && spawn_ok && L.n_rhs >= lim) ? mg_spawn_parts(L, lim) : 0;
            const unsigned long long smask = __ballot(m > 0);
            if (smask) {
                /* inclusive scan of the part counts over the wavefront */
                int incl = m;
                for (int off = 1; off < 64; off <<= 1) {
                    const int v = __shfl_up(incl, off);
                    if (lane >= off) incl += v;
                }
                const int total = __shfl(incl, 63);
                unsigned base = 0;
                if (lane == 0) base = atomicAdd(ws.n_pieces, (unsigned)total);
                base = __shfl(base, 0);
                if (base < cap) reserved = (int)(base + (unsigned)total < cap ? (unsigned)total : cap - base);
                const unsigned mine = base + (unsigned)(incl - m);
                if (m > 0) {
                    if (mine + (unsigned)m <= cap) {
                        mg_spawn_at(ws, (int)ticket, L, mine, m, part != 0u);
                        L.U[JSX_MG_TB] = L.t;
                        L.done = 1;
                        have = 2;
                    } else {
                        mg_spawn_void(ws, mine, m, part != 0u);
                        spawn_ok = 0;   /* table full: this piece runs to its end */
                    }
                }
                spawned = part ? (int)m : 0;
                sbase = mine;
            }
        }
        dec += __popcll(__ballot(have == 2));
        if (part) {
            if (reserved || (it & 15u) == 0u) {
                /* one atomic: the reserved slots on, the finished pieces off */
                const int d = reserved - dec;
                if (d != 0 && lane == 0) atomicAdd(ws.mg_pending, (unsigned)d);
                dec = 0;
            }
            /* the new slots' marks once their pieces (stored at the coherence
               point) and the pending count are complete: vmcnt(0) */
            if (__any(spawned > 0)) {
                __builtin_amdgcn_s_waitcnt(0x0F70);
                for (unsigned i = sbase; i < sbase + (unsigned)spawned && i < cap; ++i)
                    __hip_atomic_store(ws.mg_ready + i, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (have == 2) {
            PieceOut &o = ws.pout[ticket];
#pragma unroll
            for (int i = 0; i < 4; ++i) o.P[i] = L.U[16 + i];
#pragma unroll
            for (int i = 0; i < 16; ++i) o.Phi[i] = L.U[i];
            o.flags = L.flags;
            o.n_rhs = L.n_rhs;
            o.n_steps = L.n_steps;
            rhs_acc += L.n_rhs;
            have = 0;
        }
    }
    unsigned long long tot = d_wave_sum_u64(rhs_acc);
    if (lane == 0) {
        if (tot) atomicAdd(ws.rhs_mg, tot);
        atomicAdd(ws.mg_stat, (unsigned long long)it);
    }
}

/* ---- 6a. chain jumping: after round r every piece holds the composition of
 * the 2^r pieces of its chain that start at it, and links to the piece after
 * them (synchronous rounds, ping-pong buffers).  Four rounds shorten every
 * chain 16-fold before the per-ray walk; composing in a tree instead of
 * left to right changes only the FP64 rounding.  Slots that are not in any
 * chain (marked empty) may hold stale links: links outside the table end the
 * chain. */
__device__ inline void split_compose(const PieceOut &a, const PieceOut &b, PieceOut &o) {
    /* o = b after a: Phi = Phi_b Phi_a, P = Phi_b P_a + P_b */
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            o.Phi[4 * j + i] = b.Phi[i] * a.Phi[4 * j] + b.Phi[4 + i] * a.Phi[4 * j + 1] +
                               b.Phi[8 + i] * a.Phi[4 * j + 2] + b.Phi[12 + i] * a.Phi[4 * j + 3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        o.P[i] = b.Phi[i] * a.P[0] + b.Phi[4 + i] * a.P[1] + b.Phi[8 + i] * a.P[2] + b.Phi[12 + i] * a.P[3] + b.P[i];
    o.flags = a.flags | b.flags;
    o.n_rhs = a.n_rhs + b.n_rhs;
    o.n_steps = a.n_steps + b.n_steps;
    o.pad = 0;
}

__global__ __launch_bounds__(256) void jsx_k_split_jump(SplitWS ws, int round) {
    unsigned n = *ws.n_pieces;
    if (n > (unsigned)ws.piece_cap) n = (unsigned)ws.piece_cap;
    const PieceOut *src = (round & 1) ? ws.pout2 : ws.pout;
    PieceOut *dst = (round & 1) ? ws.pout : ws.pout2;
    const int *snext = round == 0 ? nullptr : ((round & 1) ? ws.next_a : ws.next_b);
    int *dnext = (round & 1) ? ws.next_b : ws.next_a;
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int nk = snext ? snext[k] : ws.pieces[k].next;
        if (nk < 0 || nk >= (int)n) {
            dst[k] = src[k];
            dnext[k] = -1;
        } else {
            PieceOut o;
            split_compose(src[k], src[nk], o);
            dst[k] = o;
            dnext[k] = snext ? snext[nk] : ws.pieces[nk].next;
        }
    }
}

/* ---- 6. compose pieces and write the pixel ---- */
__global__ __launch_bounds__(256) void jsx_k_split_combine(const DOut *__restrict__ outs, SplitWS ws, int jumped) {
    const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned n = *ws.n_long;
    if (n > (unsigned)ws.cap) n = ws.cap;
    if (r >= n) return;
    const LongRay &R = ws.rays[r];
    double S[4] = {R.x[0], R.x[1], R.x[2], R.x[3]};
    uint32_t fl = R.flags, nr = R.n_rhs, nst = R.n_steps;
    if (R.piece0 + R.npieces > ws.piece_cap) {
        fl |= JSX_PIX_FAIL_MAXTRIALS;   /* planner overflow: cannot happen with the sizing in use */
    } else {
        /* Walk the chain.  Without jumping, chains are runs of consecutive
           slots (a ray's planned pieces, each round's continuation parts), so
           the next four slots' links and outputs are loaded together and
           applied while the links stay consecutive: one load latency per run
           of up to four.  After jumping the links are in next_b. */
        int k = R.piece0;
        int guard = 0;
        if (jumped) {
            for (; k >= 0 && k < ws.piece_cap && guard < ws.piece_cap; ++guard) {
                const PieceOut o = ws.pout[k];
                double T[4];
                for (int c = 0; c < 4; ++c)
                    T[c] = o.Phi[c] * S[0] + o.Phi[4 + c] * S[1] + o.Phi[8 + c] * S[2] + o.Phi[12 + c] * S[3] + o.P[c];
                for (int c = 0; c < 4; ++c) S[c] = T[c];
                fl |= o.flags;
                nr += o.n_rhs;
                nst += o.n_steps;
                k = ws.next_b[k];
            }
            k = -1;
        }
        while (k >= 0 && guard < ws.piece_cap) {
            int nx[4];
            PieceOut o[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int kk = k + i < ws.piece_cap ? k + i : k;
                nx[i] = ws.pieces[kk].next;
                o[i] = ws.pout[kk];
            }
            int nk = -1;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double T[4];
                for (int c = 0; c < 4; ++c)
                    T[c] = o[i].Phi[c] * S[0] + o[i].Phi[4 + c] * S[1] + o[i].Phi[8 + c] * S[2] +
                           o[i].Phi[12 + c] * S[3] + o[i].P[c];
                for (int c = 0; c < 4; ++c) S[c] = T[c];
                fl |= o[i].flags;
                nr += o[i].n_rhs;
                nst += o[i].n_steps;
                ++guard;
                nk = nx[i];
                if (i == 3 || nx[i] != k + i + 1 || k + i + 1 >= ws.piece_cap) break;
            }
            k = nk;
        }
    }
    const DOut &O = outs[R.slot];
    const int64_t q = R.pix;
    if ((fl & JSX_PIX_HIT) && S[0] != S[0]) fl |= JSX_PIX_NAN;
    if (O.tau) O.tau[q] = R.tau_bg;
    if (O.I) O.I[q] = S[0];
    if (O.Q) O.Q[q] = S[1];
    if (O.U) O.U[q] = S[2];
    if (O.V) O.V[q] = S[3];
    if (O.l) O.l[q] = R.thick;
    if (O.flags) O.flags[q] = fl & ~JSX_PIX_HZ_SEG1;
    if (O.n_rhs) O.n_rhs[q] = nr;
    if (O.n_steps) O.n_steps[q] = nst;
}

/* zero the split counters (a kernel: graph-safe) */
__global__ void jsx_k_split_reset(SplitWS ws) {
    if (threadIdx.x == 0) {
        *ws.n_long = 0u;
        *ws.n_pieces = 0u;
        *ws.piece_head = 0u;
        *ws.dp_end = 0u;
        *ws.mg_head = 0u;
        *ws.round_end = 0u;
        *ws.mg_pending = 0u;
        *ws.mg_epoch += 1u;
        *ws.n_mg = 0u;
        *ws.rhs_pieces = 0ull;
        *ws.rhs_mg = 0ull;
        *ws.mg_stat = 0ull;
    }
}

/* next round of exponential pieces: the pieces spawned by the last one */
__global__ void jsx_k_split_round(SplitWS ws) {
    if (threadIdx.x == 0) {
        const unsigned e = *ws.round_end;
        const unsigned n = *ws.n_pieces;
        *ws.mg_head = e;
        *ws.round_end = n < (unsigned)ws.piece_cap ? n : (unsigned)ws.piece_cap;
    }
}

#endif /* JSX_SPLIT_H */
