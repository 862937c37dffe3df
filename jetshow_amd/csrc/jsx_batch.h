/*
 * jsx_batch.h -- one persistent launch over the rays of many images.
 *
 * Ray ids 0..total-1 enumerate every pixel of every request (request r owns
 * [ray_off[r], ray_off[r+1])).  A triage pass first tests every ray once:
 * rays that need no integration (misses of the jet, rays that end at their
 * first point) are written out there, the others are appended to a live list,
 * one contiguous chunk per wavefront so neighbouring pixels stay together.
 * The live list is dealt round-robin in blocks of `block` entries to 8
 * queues, one per XCD; a wavefront drains the queue of the XCD it runs on and,
 * when that queue is empty, steals from the others in turn.  Each lane that
 * finishes takes a new ray, so wavefronts stay full while rays of very
 * different step counts drain; one atomic per wavefront per refill round
 * hands out entries.  Without the triage list (live == nullptr) the queues
 * deal ray ids directly and a lane whose ray misses refills on the next trip.
 * A lane's request index selects its launch constants; when every active lane
 * of the wavefront serves the same request the constants are read with scalar
 * loads, otherwise with vector loads.
 *
 * Once every queue is dry the launch's tail begins.  The wavefronts keep a
 * count of lanes that still hold a ray (`active`); long full-Stokes rays are
 * parked for the time-parallel path only when that count has fallen to
 * `park_lanes`, i.e. when the machine is running out of work, so a launch
 * whose remaining rays still fill the chip finishes them sequentially; or
 * when at most `park_wave` lanes of their own wavefront still hold a ray
 * (the wavefront then runs mostly empty).
 *
 * Within a request, ids map to pixels either row-major or apex columns first
 * (longest rays first).
 */
#ifndef JSX_BATCH_H
#define JSX_BATCH_H

#include "jsx_split.h"

#define JSX_NQ 8

struct BatchQueues {
    unsigned long long head[JSX_NQ * 8];   /* one counter per queue, 64 B apart */
    unsigned long long rhs;                /* right-hand sides evaluated by the launch */
    unsigned long long trips;              /* wavefront loop trips that evaluated RHS */
    unsigned long long trips_uniform;      /* ... of them with all lanes on one request */
    unsigned long long trips_tail;         /* ... of them after the queues ran dry */
    unsigned long long n_live;             /* rays the triage pass left for integration */
    unsigned bucket_n[32];                 /* live rays per cost bucket (JSX_LPT_BUCKETS) */
    unsigned bucket_fill[32];              /* scatter cursors */
    unsigned long long active;             /* lanes holding a ray (sum of per-wave deltas) */
    unsigned drained;                      /* set once every queue has run dry */
};

/* policy: FP64 machine */
template <int NS>
struct PolFP64 {
    using LaneT = Lane<NS>;
    using Params = DParams;
    template <class Q>
    __device__ __forceinline__ static void begin(LaneT &L, const Q &q, int64_t pix) { lane_begin(L, q.m, q.im, q.S, pix); }
    template <class Q>
    __device__ __forceinline__ static void rhs_consume(LaneT &L, const Q *q0) {
        double r[NS];
        const Q *q = q0;
        JSX_OPAQUE_V(q);
        lane_rhs(L, q->m, q->im, q->S, r);
        q = q0;
        JSX_OPAQUE_V(q);
        lane_consume(L, q->S, q->im, r);
    }
    template <class Q>
    __device__ __forceinline__ static void rhs_consume_uniform(LaneT &L, const Q *q0) {
        double r[NS];
        const JSX_CONST Q *q = (const JSX_CONST Q *)q0;
        JSX_OPAQUE_S(q);
        lane_rhs(L, q->m, q->im, q->S, r);
        q = (const JSX_CONST Q *)q0;
        JSX_OPAQUE_S(q);
        lane_consume(L, q->S, q->im, r);
    }
    __device__ __forceinline__ static void write(const LaneT &L, const DOut &o, int64_t pix) { lane_write(L, o, pix); }
    __device__ static bool done(const LaneT &L) { return L.phase == PH_DONE; }
    __device__ static int64_t pix(const LaneT &L) { return L.pix; }
    __device__ static unsigned rhs(const LaneT &L) { return L.n_rhs; }
    /* stiffness k_I + |k_F| + |k_C| + |h_Q| + |k_pol| at a point [cm], per cm */
    template <class Q>
    __device__ static float stiffness(const Q &q, const double *p_cm) {
        const double n_los[3] = {-1. * q.im.dir[0], -1. * q.im.dir[1], -1. * q.im.dir[2]};
        DCoef c;
        d_coefficients<NS == 4 ? 1 : 0>(q.m, p_cm, n_los, q.S.nu, c);
        const double pol = sqrt(c.kQ * c.kQ + c.kU * c.kU + c.kV * c.kV);
        return (float)(fabs(c.kI) + fabs(c.kF) + fabs(c.kC) + fabs(c.hQ) + pol);
    }
    /* FP64 rays are never split: results stay identical to the sequential oracle */
    __device__ static bool park(const LaneT &, const Params &, int, const SplitWS &, unsigned) { return false; }
};

/* policy: mixed-precision machine */
template <int NS>
struct PolMixed {
    using LaneT = MLane<NS>;
    using Params = DParamsM;
    template <class Q>
    __device__ static void begin(LaneT &L, const Q &q, int64_t pix) { ml_begin(L, q.m, q.im, q.S, q.c, pix); }
    template <class Q>
    __device__ static void rhs_consume(LaneT &L, const Q *q0) {
        float r[ml_nr<NS>()];
        const Q *q = q0;
        JSX_OPAQUE_V(q);
        ml_rhs(L, q->m, q->im, q->c, r);
        q = q0;
        JSX_OPAQUE_V(q);
        ml_consume(L, q->S, q->im, q->c, r);
    }
    template <class Q>
    __device__ static void rhs_consume_uniform(LaneT &L, const Q *q0) {
        float r[ml_nr<NS>()];
        const JSX_CONST Q *q = (const JSX_CONST Q *)q0;
        JSX_OPAQUE_S(q);
        ml_rhs(L, q->m, q->im, q->c, r);
        q = (const JSX_CONST Q *)q0;
        JSX_OPAQUE_S(q);
        ml_consume(L, q->S, q->im, q->c, r);
    }
    __device__ static void write(const LaneT &L, const DOut &o, int64_t pix) { ml_write(L, o, pix); }
    __device__ static bool done(const LaneT &L) { return L.phase == PH_DONE; }
    __device__ static int64_t pix(const LaneT &L) { return L.pix; }
    __device__ static unsigned rhs(const LaneT &L) { return L.n_rhs; }
    template <class Q>
    __device__ static float stiffness(const Q &q, const double *p_cm) {
        const double p[3] = {p_cm[0] * q.c.inv_pc, p_cm[1] * q.c.inv_pc, p_cm[2] * q.c.inv_pc};
        MCoef k;
        m_coefficients<NS == 4 ? 1 : 0>(q.m, q.c, p, k);
        return split_stiffness(k) * (float)q.c.inv_pc;   /* per pc -> per cm */
    }
    /* full-Stokes rays past the budget go to the time-parallel path */
    __device__ static bool park(const LaneT &L, const Params &q, int slot, const SplitWS &ws, unsigned defer_rhs) {
        if constexpr (NS == 4) {
            if (ws.rays && q.c.split_ok && split_should_park(L, defer_rhs)) return split_park(L, slot, ws);
        }
        return false;
    }
};

/* workgroups are dealt to the 8 XCDs round-robin by workgroup id */
__device__ inline int d_xcd_of_block() { return (int)(blockIdx.x % JSX_NQ); }

/* id -> (request, pixel) */
__device__ inline int d_find_request(const int64_t *__restrict__ off, int n_req, int64_t id) {
    int lo = 0, hi = n_req - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= id) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* queue q, local counter h -> global id with blocks of `block` ids dealt round-robin */
__device__ inline int64_t d_queue_id(int q, unsigned long long h, int64_t block) {
    unsigned long long b = h / (unsigned long long)block, r = h % (unsigned long long)block;
    return (int64_t)(b * JSX_NQ * (unsigned long long)block + (unsigned long long)q * block + r);
}

/* Estimated cost of a live ray: its stiffness integrated over its segments
 * with 4 midpoint samples each (~ the Dormand-Prince step count; 8 coefficient
 * evaluations against ~10^3 for an average ray), as a bucket 0..31 of its
 * log2 (31 = costliest).  The live list is ordered by bucket, costliest
 * first, so the long rays start early and the launch's tail is short ones
 * (longest-processing-time-first scheduling). */
#define JSX_LPT_BUCKETS 32

template <class P, class L_, class Q>
__device__ inline int d_cost_bucket(const L_ &L, const Q &q) {
    float w = 0.f;
    for (int s = 0; s < L.nseg; ++s) {
        const double *a = s ? L.s1in : L.s0in, *b = s ? L.s1out : L.s0out;
        const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const float len = (float)d_norm3(d);
        for (int i = 0; i < 4; ++i) {
            const double f = 0.125 + 0.25 * i;
            const double p[3] = {a[0] + f * d[0], a[1] + f * d[1], a[2] + f * d[2]};
            const float k = P::stiffness(q, p);
            w += (k == k && k < 1e30f) ? k * 0.25f * len : 1e30f;
        }
    }
    const int e = (int)__builtin_amdgcn_logf(1.0f + w);   /* log2 */
    return e < 0 ? 0 : (e > JSX_LPT_BUCKETS - 1 ? JSX_LPT_BUCKETS - 1 : e);
}

/* Triage: one pass over every ray id.  Rays finished by their first point are
   written; the rest go to `live`, each wavefront's survivors as one chunk. */
template <class P>
__global__ __launch_bounds__(256) void jsx_k_triage(const typename P::Params *__restrict__ reqs,
                                                    const int64_t *__restrict__ ray_off,
                                                    const DOut *__restrict__ outs, int n_req, int64_t total,
                                                    int order, BatchQueues *qs, int64_t *__restrict__ live,
                                                    unsigned char *__restrict__ bucket, int pixel_walk) {
    using LaneT = typename P::LaneT;
    __shared__ unsigned s_hist[JSX_LPT_BUCKETS];
    const int lane = __lane_id();
    if (threadIdx.x < JSX_LPT_BUCKETS) s_hist[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base_id = (int64_t)blockIdx.x * blockDim.x; base_id < total; base_id += stride) {
        /* pixel_walk: threads walk the pixels in memory order, so that the
           planes a wavefront writes for the rays that need no integration
           are contiguous; else they walk the ray ids (the launch's order).
           The live list holds ray ids either way. */
        const int64_t t = base_id + threadIdx.x;
        int64_t id = t;
        int alive = 0, bk = 0;
        if (t < total) {
            LaneT L;
            const int req = d_find_request(ray_off, n_req, t);
            const typename P::Params *qp = reqs + req;
            int64_t pix;
            if (pixel_walk) {
                pix = t - ray_off[req];
                if (order == 1) {
                    const int64_t row = pix / qp->im.nl, col = pix - row * qp->im.nl;
                    id = ray_off[req] + col * qp->im.na + row;
                }
            } else {
                pix = d_ray_to_pixel(qp->im, t - ray_off[req], order);
            }
            P::begin(L, *qp, pix);
            if (P::done(L)) {
                P::write(L, outs[req], pix);
            } else {
                alive = 1;
                if (bucket) {
                    bk = d_cost_bucket<P>(L, *qp);
                    atomicAdd(&s_hist[bk], 1u);
                }
            }
        }
        const unsigned long long mask = __ballot(alive);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            unsigned long long at = 0;
            if (lane == leader) at = atomicAdd(&qs->n_live, (unsigned long long)__popcll(mask));
            const unsigned lo = __shfl((unsigned)(at & 0xffffffffull), leader);
            const unsigned hi = __shfl((unsigned)(at >> 32), leader);
            at = ((unsigned long long)hi << 32) | lo;
            const unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
            if (alive) {
                live[at + (unsigned long long)__popcll(below)] = id;
                if (bucket) bucket[at + (unsigned long long)__popcll(below)] = (unsigned char)bk;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < JSX_LPT_BUCKETS && s_hist[threadIdx.x]) atomicAdd(&qs->bucket_n[threadIdx.x], s_hist[threadIdx.x]);
}

/* Order the live list by cost bucket, costliest first: each entry goes to
 * its bucket's range (bucket ranges in descending bucket order, offsets from
 * the bucket counts of the triage pass); order inside a bucket is arbitrary. */
__global__ __launch_bounds__(256) void jsx_k_lpt_scatter(BatchQueues *qs, const int64_t *__restrict__ live,
                                                         const unsigned char *__restrict__ bucket,
                                                         int64_t *__restrict__ sorted,
                                                         unsigned char *__restrict__ sorted_bucket) {
    __shared__ unsigned s_base[JSX_LPT_BUCKETS], s_cnt[JSX_LPT_BUCKETS], s_res[JSX_LPT_BUCKETS];
    if (threadIdx.x == 0) {
        unsigned acc = 0;
        for (int b = JSX_LPT_BUCKETS - 1; b >= 0; --b) {
            s_base[b] = acc;
            acc += qs->bucket_n[b];
        }
    }
    const int64_t n = (int64_t)qs->n_live;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    /* every thread of the block runs the same number of rounds */
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        if (threadIdx.x < JSX_LPT_BUCKETS) s_cnt[threadIdx.x] = 0u;
        __syncthreads();
        int b = -1;
        unsigned rank = 0;
        if (i < n) {
            b = bucket[i];
            rank = atomicAdd(&s_cnt[b], 1u);
        }
        __syncthreads();
        if (threadIdx.x < JSX_LPT_BUCKETS)
            s_res[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(&qs->bucket_fill[threadIdx.x], s_cnt[threadIdx.x]) : 0u;
        __syncthreads();
        if (b >= 0) {
            const unsigned at = s_base[b] + s_res[b] + rank;
            sorted[at] = live[i];
            sorted_bucket[at] = (unsigned char)b;
        }
        __syncthreads();
    }
}

template <class P>
__global__ __launch_bounds__(256) void jsx_k_batch(const typename P::Params *__restrict__ reqs,
                                                   const int64_t *__restrict__ ray_off, const DOut *__restrict__ outs,
                                                   int n_req, int64_t total, BatchQueues *qs, int64_t block,
                                                   int order, int steal, SplitWS ws, unsigned defer_rhs,
                                                   const int64_t *__restrict__ live, unsigned park_lanes,
                                                   int park_wave, const unsigned char *__restrict__ lbucket,
                                                   int park_bucket) {
    using LaneT = typename P::LaneT;
    using Params = typename P::Params;
    LaneT L;
    int have = 0, exhausted = 0, req = 0;
    int early = 0;   /* the ray's cost bucket marks it for splitting from its first Stokes step */
    /* entries handed out by the queues: the triage list when there is one */
    int64_t n_ids = total;
    if (live) {
        const unsigned long long nl = __hip_atomic_load(&qs->n_live, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(nl & 0xffffffffull));
        const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(nl >> 32));
        n_ids = (int64_t)(((unsigned long long)hi << 32) | lo);
    }
    /* Long rays are parked for the time-parallel path only in the launch's
       tail: every queue dry (`drained`) and at most `park_lanes` lanes of the
       whole launch still holding a ray.  Both wave-uniform. */
    int drained = 0, may_park = 0;
    unsigned trip = 0;
    int held = 0;   /* lanes of this wavefront counted in qs->active */
    unsigned long long rhs_acc = 0, n_trips = 0, n_uniform = 0, n_tail = 0;
    const int lane = __lane_id();
    int myq = steal ? d_xcd_of_block() : 0;
    int qtry = 0;   /* queues tried since the last successful grab */
    const int nq = steal ? JSX_NQ : 1;
    for (;;) {
        int want = !have && !exhausted;
        unsigned long long mask = __ballot(want);
        if (mask) {
            int cnt = __popcll(mask);
            int leader = __ffsll((long long)mask) - 1;
            unsigned long long base = 0;
            int q = (myq + qtry) % nq;
            if (lane == leader) base = atomicAdd(&qs->head[q * 8], (unsigned long long)cnt);
            unsigned lo = __shfl((unsigned)(base & 0xffffffffull), leader);
            unsigned hi = __shfl((unsigned)(base >> 32), leader);
            base = ((unsigned long long)hi << 32) | lo;
            unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
            int64_t idx = steal ? d_queue_id(q, base + (unsigned long long)__popcll(below), block)
                                : (int64_t)(base + (unsigned long long)__popcll(below));
            /* the smallest entry handed out this round decides whether the queue ran dry */
            int64_t first = steal ? d_queue_id(q, base, block) : (int64_t)base;
            if (first >= n_ids) {
                qtry++;
                if (qtry >= nq) exhausted = want ? 1 : exhausted;
                if (qtry >= nq) qtry = nq;
            }
            if (want && !exhausted) {
                if (idx < n_ids) {
                    const int64_t id = live ? live[idx] : idx;
                    early = (lbucket && park_bucket > 0) ? (int)(lbucket[idx] >= park_bucket) : 0;
                    req = d_find_request(ray_off, n_req, id);
                    const Params *qp = reqs + req;
                    int64_t local = id - ray_off[req];
                    int64_t pix = d_ray_to_pixel(qp->im, local, order);
                    P::begin(L, *qp, pix);
                    if (P::done(L)) {
                        P::write(L, outs[req], pix);
                        rhs_acc += P::rhs(L);
                    } else {
                        have = 1;
                    }
                }
            }
        }
        {
            const int nh = __popcll(__ballot(have));
            if (nh != held) {
                if (lane == 0) atomicAdd(&qs->active, (unsigned long long)(long long)(nh - held));
                held = nh;
            }
        }
        if (!__any(have) && __all(exhausted || have)) {
            if (__all(exhausted)) break;
        }
        if (!drained) {
            if (__any(exhausted)) {
                drained = 1;
                if (lane == 0) __hip_atomic_store(&qs->drained, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if ((++trip & 15u) == 0u) {
                unsigned v = __hip_atomic_load(&qs->drained, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                drained = __builtin_amdgcn_readfirstlane((int)v);
            }
        } else if (!may_park && (++trip & 15u) == 0u) {
            const long long act = (long long)__hip_atomic_load(&qs->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            may_park = __builtin_amdgcn_readfirstlane((int)(act <= (long long)park_lanes));
        }
        if (__any(have)) {
            n_trips++;
            if (drained) n_tail++;
        }
        int uni = 0;
        if (have) {
            int r0 = __builtin_amdgcn_readfirstlane(req);
            uni = __all(req == r0);
            if (uni) {
                P::rhs_consume_uniform(L, reqs + r0);
            } else {
                P::rhs_consume(L, reqs + req);
            }
            if (P::done(L)) {
                P::write(L, outs[req], P::pix(L));
                rhs_acc += P::rhs(L);
                have = 0;
            } else if ((early || may_park || (drained && held <= park_wave)) &&
                       P::park(L, reqs[req], req, ws, early ? 0u : defer_rhs)) {
                rhs_acc += P::rhs(L);
                have = 0;
            }
        }
        if (__any(uni)) n_uniform++;   /* counted by every lane: lane 0 reports */
    }
    unsigned long long tot = d_wave_sum_u64(rhs_acc);
    if (lane == 0) {
        if (tot) atomicAdd(&qs->rhs, tot);
        if (n_trips) atomicAdd(&qs->trips, n_trips);
        if (n_uniform) atomicAdd(&qs->trips_uniform, n_uniform);
        if (n_tail) atomicAdd(&qs->trips_tail, n_tail);
    }
}

/* zero the queue heads (a kernel, so the sequence can be captured in a graph) */
__global__ void jsx_k_reset_queues(BatchQueues *qs) {
    int i = threadIdx.x;
    if (i < JSX_NQ * 8) qs->head[i] = 0ull;
    if (i == 0) {
        qs->rhs = 0ull;
        qs->trips = 0ull;
        qs->trips_uniform = 0ull;
        qs->trips_tail = 0ull;
        qs->n_live = 0ull;
        for (int b = 0; b < 32; ++b) { qs->bucket_n[b] = 0u; qs->bucket_fill[b] = 0u; }
        qs->active = 0ull;
        qs->drained = 0u;
    }
}

#endif /* JSX_BATCH_H */
