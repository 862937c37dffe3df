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
 * whose remaining rays still fill the chip finishes them sequentially.
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
    unsigned long long n_live;             /* rays the triage pass left for integration */
    unsigned long long active;             /* lanes holding a ray (sum of per-wave deltas) */
    unsigned drained;                      /* set once every queue has run dry */
};

/* policy: FP64 machine */
template <int NS>
struct PolFP64 {
    using LaneT = Lane<NS>;
    using Params = DParams;
    template <class Q>
    __device__ static void begin(LaneT &L, const Q &q, int64_t pix) { lane_begin(L, q.m, q.im, q.S, pix); }
    template <class Q>
    __device__ static void rhs_consume(LaneT &L, const Q *q0) {
        double r[NS];
        const Q *q = q0;
        JSX_OPAQUE_V(q);
        lane_rhs(L, q->m, q->im, q->S, r);
        q = q0;
        JSX_OPAQUE_V(q);
        lane_consume(L, q->S, q->im, r);
    }
    template <class Q>
    __device__ static void rhs_consume_uniform(LaneT &L, const Q *q0) {
        double r[NS];
        const JSX_CONST Q *q = (const JSX_CONST Q *)q0;
        JSX_OPAQUE_S(q);
        lane_rhs(L, q->m, q->im, q->S, r);
        q = (const JSX_CONST Q *)q0;
        JSX_OPAQUE_S(q);
        lane_consume(L, q->S, q->im, r);
    }
    __device__ static void write(const LaneT &L, const DOut &o, int64_t pix) { lane_write(L, o, pix); }
    __device__ static bool done(const LaneT &L) { return L.phase == PH_DONE; }
    __device__ static int64_t pix(const LaneT &L) { return L.pix; }
    __device__ static unsigned rhs(const LaneT &L) { return L.n_rhs; }
    /* FP64 rays are never split: results stay identical to the sequential oracle */
    __device__ static bool park(const LaneT &, int, const SplitWS &, unsigned) { return false; }
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
        float r[NS];
        const Q *q = q0;
        JSX_OPAQUE_V(q);
        ml_rhs(L, q->m, q->im, q->c, r);
        q = q0;
        JSX_OPAQUE_V(q);
        ml_consume(L, q->S, q->im, q->c, r);
    }
    template <class Q>
    __device__ static void rhs_consume_uniform(LaneT &L, const Q *q0) {
        float r[NS];
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
    /* full-Stokes rays past the budget go to the time-parallel path */
    __device__ static bool park(const LaneT &L, int slot, const SplitWS &ws, unsigned defer_rhs) {
        if constexpr (NS == 4) {
            if (ws.rays && split_should_park(L, defer_rhs)) return split_park(L, slot, ws);
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

/* Triage: one pass over every ray id.  Rays finished by their first point are
   written; the rest go to `live`, each wavefront's survivors as one chunk. */
template <class P>
__global__ __launch_bounds__(256) void jsx_k_triage(const typename P::Params *__restrict__ reqs,
                                                    const int64_t *__restrict__ ray_off,
                                                    const DOut *__restrict__ outs, int n_req, int64_t total,
                                                    int order, BatchQueues *qs, int64_t *__restrict__ live) {
    using LaneT = typename P::LaneT;
    const int lane = __lane_id();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base_id = (int64_t)blockIdx.x * blockDim.x; base_id < total; base_id += stride) {
        const int64_t id = base_id + threadIdx.x;
        int alive = 0;
        if (id < total) {
            LaneT L;
            const int req = d_find_request(ray_off, n_req, id);
            const typename P::Params *qp = reqs + req;
            const int64_t pix = d_ray_to_pixel(qp->im, id - ray_off[req], order);
            P::begin(L, *qp, pix);
            if (P::done(L)) P::write(L, outs[req], pix);
            else alive = 1;
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
            if (alive) live[at + (unsigned long long)__popcll(below)] = id;
        }
    }
}

template <class P>
__global__ __launch_bounds__(256) void jsx_k_batch(const typename P::Params *__restrict__ reqs,
                                                   const int64_t *__restrict__ ray_off, const DOut *__restrict__ outs,
                                                   int n_req, int64_t total, BatchQueues *qs, int64_t block,
                                                   int order, int steal, SplitWS ws, unsigned defer_rhs,
                                                   const int64_t *__restrict__ live, unsigned park_lanes) {
    using LaneT = typename P::LaneT;
    using Params = typename P::Params;
    LaneT L;
    int have = 0, exhausted = 0, req = 0;
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
    unsigned long long rhs_acc = 0, n_trips = 0, n_uniform = 0;
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
        if (__any(have)) n_trips++;
        if (have) {
            int r0 = __builtin_amdgcn_readfirstlane(req);
            if (__all(req == r0)) {
                n_uniform++;
                P::rhs_consume_uniform(L, reqs + r0);
            } else {
                P::rhs_consume(L, reqs + req);
            }
            if (P::done(L)) {
                P::write(L, outs[req], P::pix(L));
                rhs_acc += P::rhs(L);
                have = 0;
            } else if (may_park && P::park(L, req, ws, defer_rhs)) {
                rhs_acc += P::rhs(L);
                have = 0;
            }
        }
    }
    unsigned long long tot = d_wave_sum_u64(rhs_acc);
    if (lane == 0) {
        if (tot) atomicAdd(&qs->rhs, tot);
        if (n_trips) atomicAdd(&qs->trips, n_trips);
        if (n_uniform) atomicAdd(&qs->trips_uniform, n_uniform);
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
        qs->n_live = 0ull;
        qs->active = 0ull;
        qs->drained = 0u;
    }
}

#endif /* JSX_BATCH_H */
