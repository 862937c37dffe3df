/*
 * jsx_serve.h -- continuous-batching request server (included by
 * jsx_host.hip, which provides the context and descriptor lowering).
 *
 * Requests (jet model + image + solver parameters, optionally with their own
 * field grids) are submitted from any thread.  A scheduler thread:
 *   1. admits pending requests in arrival order while their HBM footprint
 *      (output planes + grids) fits the server budget and a request slot is
 *      free: grids are uploaded, descriptors lowered into the device request
 *      table, and the request's rays are appended to the ray id space;
 *   2. runs one epoch: a persistent kernel over the shared ray queue with an
 *      iteration budget.  Lanes first resume ray states saved by the previous
 *      epoch, then take new rays of any admitted request; when the budget is
 *      spent, a lane writes its complete ray state to the device-resident
 *      resume pool and exits, so every epoch is short and newly admitted
 *      requests start within one epoch while older requests are still in
 *      flight.  The two epoch variants (resume pool A->B and B->A) are
 *      captured once as HIP graphs and replayed;
 *   3. reads the per-request remaining-ray counters, copies finished requests'
 *      planes to their host buffers, frees their memory and stamps latency.
 * A ray's arithmetic does not depend on when it runs or what runs beside it,
 * so results are identical alone and in the mix.
 */
#ifndef JSX_SERVE_H
#define JSX_SERVE_H

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

/* device-side epoch description (pointers fixed for the server's lifetime) */
struct ServeEpoch {
    int64_t total;          /* ray ids handed to the queue so far         */
    int n_table;            /* entries in the ray table                   */
    int budget;             /* loop trips per epoch                       */
    int order;
};

template <class P>
struct ServeSaved {
    typename P::LaneT L;
    int slot;
};

template <class P>
__global__ __launch_bounds__(256) void jsx_k_serve(const typename P::Params *__restrict__ reqs,
                                                   const DOut *__restrict__ outs, const int64_t *__restrict__ tab_off,
                                                   const int *__restrict__ tab_slot, const ServeEpoch *__restrict__ ep,
                                                   unsigned long long *head, ServeSaved<P> *res_in,
                                                   const unsigned *res_in_count, unsigned *res_take,
                                                   ServeSaved<P> *res_out, unsigned *res_out_count,
                                                   unsigned *remaining) {
    using LaneT = typename P::LaneT;
    using Params = typename P::Params;
    LaneT L;
    int have = 0, exhausted = 0, slot = 0, resumed_all = 0;
    const int lane = __lane_id();
    const int64_t total = ep->total;
    const int n_table = ep->n_table, budget = ep->budget, order = ep->order;
    const unsigned n_in = *res_in_count;
    for (int it = 0;; ++it) {
        int want = !have && !exhausted;
        unsigned long long mask = __ballot(want);
        if (mask) {
            int cnt = __popcll(mask);
            int leader = __ffsll((long long)mask) - 1;
            unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
            int rank = __popcll(below);
            if (!resumed_all) {
                /* saved ray states of the previous epoch first */
                unsigned b0 = 0;
                if (lane == leader) b0 = atomicAdd(res_take, (unsigned)cnt);
                b0 = __shfl(b0, leader);
                if (b0 >= n_in) resumed_all = 1;
                if (want && b0 + rank < n_in) {
                    const ServeSaved<P> &sv = res_in[b0 + rank];
                    L = sv.L;
                    slot = sv.slot;
                    have = 1;
                }
            } else {
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(head, (unsigned long long)cnt);
                unsigned lo = __shfl((unsigned)(base & 0xffffffffull), leader);
                unsigned hi = __shfl((unsigned)(base >> 32), leader);
                base = ((unsigned long long)hi << 32) | lo;
                if ((int64_t)base >= total) exhausted = 1;
                int64_t id = (int64_t)(base + (unsigned long long)rank);
                if (want && id < total && n_table > 0) {
                    int lo_i = 0, hi_i = n_table - 1;
                    while (lo_i < hi_i) {
                        int mid = (lo_i + hi_i + 1) >> 1;
                        if (tab_off[mid] <= id) lo_i = mid;
                        else hi_i = mid - 1;
                    }
                    slot = tab_slot[lo_i];
                    const Params *qp = reqs + slot;
                    int64_t pix = d_ray_to_pixel(qp->im, id - tab_off[lo_i], order);
                    P::begin(L, *qp, pix);
                    if (P::done(L)) {
                        P::write(L, outs[slot], pix);
                        atomicSub(&remaining[slot], 1u);
                    } else {
                        have = 1;
                    }
                }
            }
        }
        if (it >= budget) {
            /* out of budget: park the ray state for the next epoch */
            if (have) {
                unsigned k = atomicAdd(res_out_count, 1u);
                res_out[k].L = L;
                res_out[k].slot = slot;
                have = 0;
            }
            break;
        }
        if (!__any(have) && __all(exhausted)) break;
        if (have) {
            int s0 = __builtin_amdgcn_readfirstlane(slot);
            if (__all(slot == s0)) P::rhs_consume_uniform(L, reqs + s0);
            else P::rhs_consume(L, reqs + slot);
            if (P::done(L)) {
                P::write(L, outs[slot], P::pix(L));
                atomicSub(&remaining[slot], 1u);
                have = 0;
            }
        }
    }
}

/* zero the resume bookkeeping of the next epoch */
__global__ void jsx_k_serve_prep(unsigned *res_take, unsigned *res_out_count) {
    if (threadIdx.x == 0) { *res_take = 0u; *res_out_count = 0u; }
}

struct jsx_server;

struct ServeReq {
    uint64_t id;
    jsx_request r;
    std::vector<jsx_grid_payload> grids;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int slot = -1;
    int status = 0;
    bool done = false;
    size_t footprint = 0;
    jsx_outputs dev;          /* device planes */
    std::vector<int32_t> grid_handles;
    int64_t npix = 0;
};

struct jsx_server {
    jsx_ctx *ctx = nullptr;
    int precision = JSX_PREC_MIXED, stokes = JSX_STOKES_FULL;
    int max_active = 64;
    size_t budget = 0, used = 0;
    int epoch_budget = 4000;
    hipStream_t st = nullptr;
    /* device tables */
    void *d_reqs = nullptr;
    DOut *d_outs = nullptr;
    int64_t *d_tab_off = nullptr;
    int *d_tab_slot = nullptr;
    ServeEpoch *d_ep = nullptr, *h_ep = nullptr;
    unsigned long long *d_head = nullptr;
    void *d_res[2] = {nullptr, nullptr};
    unsigned *d_res_count = nullptr;   /* [0],[1]: counts of pools A,B; [2]: take counter */
    unsigned *d_remaining = nullptr, *h_remaining = nullptr;
    int64_t *h_tab_off = nullptr;
    int *h_tab_slot = nullptr;
    size_t res_cap = 0, lane_bytes = 0;
    int grid_blocks = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int cur = 0;  /* resume pool read by the next epoch */
    /* host bookkeeping */
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<ServeReq *> pending;
    std::vector<ServeReq *> slots;
    std::unordered_map<uint64_t, ServeReq *> all;
    std::vector<std::pair<int64_t, int>> table;    /* (first id, slot) of requests with rays left to hand out */
    std::vector<int64_t> slot_end;                  /* one past the last ray id of the slot's request */
    int64_t total = 0;
    int64_t head_seen = 0;  /* queue head after the last epoch */
    uint64_t next_id = 1;
    bool stop = false;
    std::thread th;
    std::string err;
    int64_t epochs = 0;
};

static double serve_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class P>
static size_t serve_lane_bytes() { return sizeof(ServeSaved<P>); }

template <class P>
static void serve_launch(jsx_server *S, int from) {
    using Params = typename P::Params;
    hipLaunchKernelGGL(jsx_k_serve_prep, dim3(1), dim3(64), 0, S->st, S->d_res_count + 2, S->d_res_count + (1 - from));
    hipLaunchKernelGGL(jsx_k_serve<P>, dim3(S->grid_blocks), dim3(256), 0, S->st, (const Params *)S->d_reqs, S->d_outs,
                       S->d_tab_off, S->d_tab_slot, S->d_ep, S->d_head, (ServeSaved<P> *)S->d_res[from],
                       S->d_res_count + from, S->d_res_count + 2, (ServeSaved<P> *)S->d_res[1 - from],
                       S->d_res_count + (1 - from), S->d_remaining);
}

static void serve_dispatch_launch(jsx_server *S, int from) {
    if (S->precision == JSX_PREC_MIXED) {
        if (S->stokes == JSX_STOKES_FULL) serve_launch<PolMixed<4>>(S, from);
        else serve_launch<PolMixed<1>>(S, from);
    } else {
        if (S->stokes == JSX_STOKES_FULL) serve_launch<PolFP64<4>>(S, from);
        else serve_launch<PolFP64<1>>(S, from);
    }
}

static size_t serve_lane_size(int precision, int stokes) {
    if (precision == JSX_PREC_MIXED)
        return stokes == JSX_STOKES_FULL ? serve_lane_bytes<PolMixed<4>>() : serve_lane_bytes<PolMixed<1>>();
    return stokes == JSX_STOKES_FULL ? serve_lane_bytes<PolFP64<4>>() : serve_lane_bytes<PolFP64<1>>();
}

static size_t req_footprint(const ServeReq *q) {
    size_t n = (size_t)(q->r.image.n_across * q->r.image.n_along);
    size_t f = n * (6 * sizeof(double) + 3 * sizeof(uint32_t));
    for (auto &g : q->grids) f += sizeof(double) * (size_t)g.desc.nz * (size_t)g.desc.nr;
    return f;
}

/* admit one request into slot s (scheduler thread, S->mu held) */
static int serve_admit(jsx_server *S, ServeReq *q, int s) {
    jsx_ctx *ctx = S->ctx;
    jsx_request &r = q->r;
    q->npix = r.image.n_across * r.image.n_along;
    for (auto &g : q->grids) {
        int32_t h = 0;
        int rc = jsx_grid_upload(ctx, &g.desc, g.values, &h);
        if (rc) return rc;
        q->grid_handles.push_back(h);
        if (g.field == JSX_FIELD_B) r.jet.grid_b = h;
        else if (g.field == JSX_FIELD_B2) r.jet.grid_b2 = h;
        else if (g.field == JSX_FIELD_N) r.jet.grid_n = h;
        else if (g.field == JSX_FIELD_V) r.jet.grid_v = h;
    }
    DModel m;
    DSolver Sv;
    DImage D;
    int rc;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if ((rc = lower_model(ctx, r.jet, r.params.flags, m))) return rc;
        if ((rc = lower_solver(ctx, r.params, Sv))) return rc;
        if ((rc = lower_image(ctx, r.image, D, nullptr))) return rc;
    }
    if ((rc = jsx_alloc_outputs(ctx, q->npix, JSX_STOKES_FULL, &q->dev))) return rc;
    DOut O = lower_out(q->dev);
    if (S->precision == JSX_PREC_MIXED) {
        DParamsM pm;
        pm.m = m; pm.im = D; pm.S = Sv;
        lower_mixed(r.jet, r.params, m, D, pm.c);
        HIPCHECK(ctx, hipMemcpyAsync((char *)S->d_reqs + s * sizeof(DParamsM), &pm, sizeof pm, hipMemcpyHostToDevice, S->st));
    } else {
        DParams pd;
        pd.m = m; pd.im = D; pd.S = Sv;
        HIPCHECK(ctx, hipMemcpyAsync((char *)S->d_reqs + s * sizeof(DParams), &pd, sizeof pd, hipMemcpyHostToDevice, S->st));
    }
    HIPCHECK(ctx, hipMemcpyAsync(S->d_outs + s, &O, sizeof O, hipMemcpyHostToDevice, S->st));
    S->h_remaining[s] = (unsigned)q->npix;
    HIPCHECK(ctx, hipMemcpyAsync(S->d_remaining + s, &S->h_remaining[s], sizeof(unsigned), hipMemcpyHostToDevice, S->st));
    HIPCHECK(ctx, hipStreamSynchronize(S->st));
    q->slot = s;
    q->t_admit = serve_now();
    /* ids below the queue head are spent (lanes overshoot it when the queue
       runs dry), so a new request's rays start at or above it */
    if (S->total < S->head_seen) S->total = S->head_seen;
    S->table.push_back({S->total, s});
    S->total += q->npix;
    S->slot_end[s] = S->total;
    return JSX_OK;
}

static void serve_finish(jsx_server *S, ServeReq *q) {
    jsx_ctx *ctx = S->ctx;
    size_t nd = (size_t)q->npix * sizeof(double), nu = (size_t)q->npix * sizeof(uint32_t);
    const jsx_outputs &h = q->r.out_host;
    struct { void *h; void *d; size_t b; } cp[] = {
        {h.tau, q->dev.tau, nd}, {h.I, q->dev.I, nd}, {h.Q, q->dev.Q, nd}, {h.U, q->dev.U, nd}, {h.V, q->dev.V, nd},
        {h.l, q->dev.l, nd}, {h.flags, q->dev.flags, nu}, {h.n_rhs, q->dev.n_rhs, nu}, {h.n_steps, q->dev.n_steps, nu}};
    for (auto &c : cp)
        if (c.h && hipMemcpy(c.h, c.d, c.b, hipMemcpyDeviceToHost) != hipSuccess) q->status = JSX_EHIP;
    jsx_free_outputs(ctx, &q->dev);
    for (int32_t g : q->grid_handles) jsx_grid_release(ctx, g);
    q->t_done = serve_now();
    q->done = true;
    S->used -= q->footprint;
    S->slots[q->slot] = nullptr;
}

static int serve_build_graphs(jsx_server *S) {
    for (int k = 0; k < 2; ++k) {
        hipGraph_t g;
        HIPCHECK(S->ctx, hipStreamBeginCapture(S->st, hipStreamCaptureModeThreadLocal));
        serve_dispatch_launch(S, k);
        HIPCHECK(S->ctx, hipStreamEndCapture(S->st, &g));
        HIPCHECK(S->ctx, hipGraphInstantiate(&S->graph[k], g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
    }
    return JSX_OK;
}

static void serve_loop(jsx_server *S) {
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        /* 1. admission in arrival order */
        while (!S->pending.empty()) {
            ServeReq *q = S->pending.front();
            int s = -1;
            for (int i = 0; i < S->max_active; ++i)
                if (!S->slots[i]) { s = i; break; }
            if (s < 0) break;
            if (q->footprint > S->budget) {
                q->status = JSX_EBUDGET; q->done = true; q->t_done = serve_now();
                S->pending.pop_front();
                S->cv_done.notify_all();
                continue;
            }
            if (S->used + q->footprint > S->budget) break;
            S->pending.pop_front();
            S->slots[s] = q;
            S->used += q->footprint;
            int rc = serve_admit(S, q, s);
            if (rc) {
                q->status = rc; q->done = true; q->t_done = serve_now();
                S->slots[s] = nullptr; S->used -= q->footprint;
                S->cv_done.notify_all();
            }
        }
        bool any = false;
        for (auto *q : S->slots) any = any || q;
        if (!any) {
            if (S->stop && S->pending.empty()) return;
            S->cv.wait(lk);
            continue;
        }
        /* 2. one epoch over the current ray table */
        S->h_ep->total = S->total;
        S->h_ep->n_table = (int)S->table.size();
        S->h_ep->budget = S->epoch_budget;
        S->h_ep->order = 1;
        for (size_t i = 0; i < S->table.size(); ++i) { S->h_tab_off[i] = S->table[i].first; S->h_tab_slot[i] = S->table[i].second; }
        lk.unlock();
        hipError_t e = hipSuccess;
        if (!S->table.empty()) {
            e = hipMemcpyAsync(S->d_tab_off, S->h_tab_off, sizeof(int64_t) * S->table.size(), hipMemcpyHostToDevice, S->st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(S->d_tab_slot, S->h_tab_slot, sizeof(int) * S->table.size(), hipMemcpyHostToDevice, S->st);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(S->d_ep, S->h_ep, sizeof(ServeEpoch), hipMemcpyHostToDevice, S->st);
        if (e == hipSuccess) e = hipGraphLaunch(S->graph[S->cur], S->st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(S->h_remaining, S->d_remaining, sizeof(unsigned) * S->max_active, hipMemcpyDeviceToHost, S->st);
        if (e == hipSuccess) e = hipStreamSynchronize(S->st);
        lk.lock();
        S->epochs++;
        if (e != hipSuccess) {
            S->err = hipGetErrorString(e);
            for (auto *q : S->slots)
                if (q) { q->status = JSX_EHIP; q->done = true; }
            S->cv_done.notify_all();
            return;
        }
        S->cur = 1 - S->cur;
        /* 3. completions; prune requests whose rays are all handed out */
        for (int i = 0; i < S->max_active; ++i) {
            ServeReq *q = S->slots[i];
            if (q && S->h_remaining[i] == 0) serve_finish(S, q);
        }
        S->cv_done.notify_all();
        unsigned long long head = 0;
        if (hipMemcpy(&head, S->d_head, sizeof head, hipMemcpyDeviceToHost) == hipSuccess) S->head_seen = (int64_t)head;
        std::vector<std::pair<int64_t, int>> keep;
        for (auto &t : S->table)
            if (S->slot_end[t.second] > S->head_seen && S->slots[t.second]) keep.push_back(t);
        S->table.swap(keep);
    }
}

extern "C" int jsx_server_create(jsx_ctx *ctx, int32_t precision, int32_t stokes, int32_t max_active,
                                 size_t hbm_budget, int32_t epoch_budget, jsx_server **out) {
    if (!ctx || !out || max_active <= 0) return JSX_EINVAL;
    if ((precision != JSX_PREC_FP64 && precision != JSX_PREC_MIXED) || (stokes != JSX_STOKES_I && stokes != JSX_STOKES_FULL))
        return fail(ctx, JSX_EINVAL, "bad precision/stokes");
    jsx_server *S = new jsx_server();
    S->ctx = ctx; S->precision = precision; S->stokes = stokes; S->max_active = max_active;
    S->budget = hbm_budget ? hbm_budget : ctx->budget / 2;
    S->epoch_budget = epoch_budget > 0 ? epoch_budget : 4000;
    S->slots.assign(max_active, nullptr);
    S->slot_end.assign(max_active, 0);
    (void)hipSetDevice(ctx->device);
    size_t esz = precision == JSX_PREC_MIXED ? sizeof(DParamsM) : sizeof(DParams);
    S->grid_blocks = ctx->n_cu * ctx->persist_blocks_per_cu;
    S->lane_bytes = serve_lane_size(precision, stokes);
    S->res_cap = (size_t)S->grid_blocks * 256;
    bool ok = hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&S->d_reqs, esz * max_active) == hipSuccess &&
              hipMalloc(&S->d_outs, sizeof(DOut) * max_active) == hipSuccess &&
              hipMalloc(&S->d_tab_off, sizeof(int64_t) * max_active) == hipSuccess &&
              hipMalloc(&S->d_tab_slot, sizeof(int) * max_active) == hipSuccess &&
              hipMalloc(&S->d_ep, sizeof(ServeEpoch)) == hipSuccess &&
              hipHostMalloc(&S->h_ep, sizeof(ServeEpoch), hipHostMallocDefault) == hipSuccess &&
              hipMalloc(&S->d_head, sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&S->d_res[0], S->lane_bytes * S->res_cap) == hipSuccess &&
              hipMalloc(&S->d_res[1], S->lane_bytes * S->res_cap) == hipSuccess &&
              hipMalloc(&S->d_res_count, 4 * sizeof(unsigned)) == hipSuccess &&
              hipMalloc(&S->d_remaining, sizeof(unsigned) * max_active) == hipSuccess &&
              hipHostMalloc(&S->h_remaining, sizeof(unsigned) * max_active, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&S->h_tab_off, sizeof(int64_t) * max_active, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&S->h_tab_slot, sizeof(int) * max_active, hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        jsx_server_destroy(S);
        return fail(ctx, JSX_ENOMEM, "server allocation failed");
    }
    /* zero the counters with a kernel-free copy before any capture */
    std::vector<unsigned char> z(std::max(sizeof(unsigned long long), 4 * sizeof(unsigned)), 0);
    (void)hipMemcpy(S->d_head, z.data(), sizeof(unsigned long long), hipMemcpyHostToDevice);
    (void)hipMemcpy(S->d_res_count, z.data(), 4 * sizeof(unsigned), hipMemcpyHostToDevice);
    std::vector<unsigned> zr(max_active, 0);
    (void)hipMemcpy(S->d_remaining, zr.data(), sizeof(unsigned) * max_active, hipMemcpyHostToDevice);
    int rc = serve_build_graphs(S);
    if (rc) { jsx_server_destroy(S); return rc; }
    S->th = std::thread(serve_loop, S);
    *out = S;
    return JSX_OK;
}

extern "C" int jsx_server_submit(jsx_server *S, const jsx_request *req, int32_t n_grids, const jsx_grid_payload *grids,
                                 uint64_t *id_out) {
    if (!S || !req || !id_out || (n_grids && !grids)) return JSX_EINVAL;
    if (req->image.n_across <= 0 || req->image.n_along <= 0 || (req->image.n_across % 2))
        return fail(S->ctx, JSX_EINVAL, "image needs n_along > 0 and even n_across > 0");
    if (req->params.precision != S->precision || req->params.stokes != S->stokes)
        return fail(S->ctx, JSX_EINVAL, "request precision/stokes differ from the server's kernel variant");
    if (!req->out_host.I) return fail(S->ctx, JSX_EINVAL, "out_host.I must be set");
    ServeReq *q = new ServeReq();
    q->r = *req;
    q->grids.assign(grids, grids + n_grids);
    q->t_submit = serve_now();
    q->footprint = req_footprint(q);
    std::lock_guard<std::mutex> lk(S->mu);
    q->id = S->next_id++;
    S->all[q->id] = q;
    S->pending.push_back(q);
    *id_out = q->id;
    S->cv.notify_all();
    return JSX_OK;
}

static void fill_timing(const ServeReq *q, jsx_request_timing *t) {
    if (!t) return;
    t->t_submit = q->t_submit;
    t->t_admit = q->t_admit;
    t->t_done = q->t_done;
    t->n_pixels = q->r.image.n_across * q->r.image.n_along;
    t->status = q->status;
}

extern "C" int jsx_server_poll(jsx_server *S, uint64_t id, int32_t *done, jsx_request_timing *t) {
    if (!S || !done) return JSX_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return JSX_ENOTFOUND;
    *done = it->second->done ? 1 : 0;
    if (*done) fill_timing(it->second, t);
    return JSX_OK;
}

extern "C" int jsx_server_wait(jsx_server *S, uint64_t id, jsx_request_timing *t) {
    if (!S) return JSX_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return JSX_ENOTFOUND;
    ServeReq *q = it->second;
    S->cv_done.wait(lk, [&] { return q->done; });
    fill_timing(q, t);
    int st = q->status;
    S->all.erase(it);
    delete q;
    return st;
}

extern "C" int jsx_server_stats(jsx_server *S, int64_t *epochs, size_t *hbm_used) {
    if (!S) return JSX_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    if (epochs) *epochs = S->epochs;
    if (hbm_used) *hbm_used = S->used;
    return JSX_OK;
}

extern "C" int jsx_server_destroy(jsx_server *S) {
    if (!S) return JSX_EINVAL;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv.notify_all();
    }
    if (S->th.joinable()) S->th.join();
    (void)hipSetDevice(S->ctx->device);
    if (S->st) (void)hipStreamSynchronize(S->st);
    for (int k = 0; k < 2; ++k)
        if (S->graph[k]) (void)hipGraphExecDestroy(S->graph[k]);
    void *dv[] = {S->d_reqs, S->d_outs, S->d_tab_off, S->d_tab_slot, S->d_ep, S->d_head, S->d_res[0], S->d_res[1],
                  S->d_res_count, S->d_remaining};
    for (void *p : dv)
        if (p) (void)hipFree(p);
    void *hv[] = {S->h_ep, S->h_remaining, S->h_tab_off, S->h_tab_slot};
    for (void *p : hv)
        if (p) (void)hipHostFree(p);
    if (S->st) (void)hipStreamDestroy(S->st);
    for (auto &kv : S->all) delete kv.second;
    delete S;
    return JSX_OK;
}

#endif /* JSX_SERVE_H */
