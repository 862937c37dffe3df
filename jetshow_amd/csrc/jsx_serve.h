/*
 * jsx_serve.h -- continuous-batching request server (included by
 * jsx_host.hip, which provides the context, the descriptor lowering and the
 * batched launch).
 *
 * Requests (jet model + image + solver parameters, optionally with their own
 * field grids) are submitted from any thread.  A scheduler thread:
 *   1. admits pending requests in arrival order while fewer than max_active
 *      are admitted and what they add to HBM fits the server budget: a
 *      request's grids stay resident after it (keyed by the caller's array,
 *      so the next request naming the same grids finds them there), and
 *      output planes are taken from and returned to a pool (hipFree waits
 *      for the device, which would stall the launches in flight); unused
 *      grids and pooled planes are freed only when the budget needs room;
 *   2. as soon as a launch lane is free, runs every admitted request that is
 *      not yet on the device as one batched launch on that lane -- the whole
 *      persistent pipeline of jsx_observe_batch_device: triage of rays that
 *      need no integration, longest-first work-stealing queues, and the
 *      time-parallel split of long full-Stokes rays.  Up to `launches` (2)
 *      launches are in flight at once, each on its own stream with its own
 *      queues and split workspace, so requests that arrive while others are
 *      on the GPU start at once and their work overlaps the running launch;
 *      under load, requests that arrive while every lane is busy form the
 *      next, larger launch;
 *   3. polls the launches' completion events, copies finished requests'
 *      planes to their host buffers, frees their memory and stamps latency.
 * A ray's arithmetic does not depend on the launch it runs in, except that a
 * long full-Stokes ray may be split into time-parallel pieces at places that
 * depend on the other long rays of its launch; results alone and in a mix
 * then agree to the solver's tolerance, not bit for bit.
 */
#ifndef JSX_SERVE_H
#define JSX_SERVE_H

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

struct jsx_server;

struct ServeReq {
    uint64_t id;
    jsx_request r;
    std::vector<jsx_grid_payload> grids;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = 0;
    bool done = false;
    size_t footprint = 0;
    jsx_outputs dev{};        /* device planes */
    int64_t cap = 0;          /* pixels the planes hold */
    std::vector<int32_t> grid_handles;
    int64_t npix = 0;
    BatchItem item{};         /* lowered descriptors */
};

/* one launch lane: a context of its own (queues, request tables, triage
 * lists and split workspace of the batched launch), its stream, and the
 * event after its launch */
struct ServeLane {
    jsx_ctx *sub = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    std::vector<ServeReq *> reqs;
    bool busy = false;
};

/* a payload grid kept in HBM while requests that name it are admitted, and
 * after, until its memory is needed (keyed by the caller's host array and its
 * description: the array must not change while a request naming it is
 * queued or admitted) */
struct ServeGrid {
    int32_t handle = 0;
    int refs = 0;
    size_t bytes = 0;
    uint64_t last_use = 0;
};

struct jsx_server {
    jsx_ctx *ctx = nullptr;
    int precision = JSX_PREC_MIXED, stokes = JSX_STOKES_FULL;
    int max_active = 64;
    /* HBM held: admitted requests' planes, free planes kept for reuse, and
       resident grids; bounded by budget */
    size_t budget = 0, used = 0;
    /* device planes are reused, not freed (hipFree waits for the device,
       which would stall launches in flight): free planes by capacity */
    std::multimap<int64_t, jsx_outputs> plane_pool;
    std::map<std::vector<double>, ServeGrid> grid_cache;
    uint64_t use_clock = 0;
    int n_active = 0;                    /* admitted, not yet finished */
    std::vector<ServeLane> lanes;
    /* host bookkeeping */
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<ServeReq *> pending;      /* submitted, not admitted */
    std::vector<ServeReq *> ready;       /* admitted, not launched  */
    std::unordered_map<uint64_t, ServeReq *> all;
    uint64_t next_id = 1;
    bool stop = false;
    bool dead = false;                   /* a HIP error ended the scheduler */
    std::thread th;
    std::string err;
    int64_t epochs = 0;                  /* launches run */
};

static double serve_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static size_t plane_bytes(int64_t n) { return (size_t)n * (6 * sizeof(double) + 3 * sizeof(uint32_t)); }

static size_t req_footprint(const ServeReq *q) {
    size_t f = plane_bytes(q->r.image.n_across * q->r.image.n_along);
    for (auto &g : q->grids) f += sizeof(double) * (size_t)g.desc.nz * (size_t)g.desc.nr;
    return f;
}

static std::vector<double> grid_key(const jsx_grid_payload &g) {
    return {(double)(uintptr_t)g.values, (double)g.field, (double)g.desc.nz, (double)g.desc.nr, g.desc.z0, g.desc.dz,
            g.desc.r0, g.desc.dr, g.desc.nan_value};
}

/* bytes a request adds to S->used when admitted: planes unless a pooled set
 * fits, grids unless resident (S->mu held) */
static size_t serve_need(jsx_server *S, const ServeReq *q) {
    const int64_t n = q->r.image.n_across * q->r.image.n_along;
    auto it = S->plane_pool.lower_bound(n);
    size_t need = (it != S->plane_pool.end() && it->first <= 2 * n) ? 0 : plane_bytes(n);
    for (auto &g : q->grids)
        if (!S->grid_cache.count(grid_key(g))) need += sizeof(double) * (size_t)g.desc.nz * (size_t)g.desc.nr;
    return need;
}

/* free unreferenced grids (least recently used first) and pooled planes
 * until `need` more bytes fit the budget; returns whether they do (S->mu
 * held; the frees wait for the device) */
static bool serve_make_room(jsx_server *S, size_t need) {
    while (S->used + need > S->budget) {
        auto victim = S->grid_cache.end();
        for (auto it = S->grid_cache.begin(); it != S->grid_cache.end(); ++it)
            if (it->second.refs == 0 && (victim == S->grid_cache.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim != S->grid_cache.end()) {
            jsx_grid_release(S->ctx, victim->second.handle);
            S->used -= victim->second.bytes;
            S->grid_cache.erase(victim);
            continue;
        }
        if (!S->plane_pool.empty()) {
            auto it = std::prev(S->plane_pool.end());   /* largest first */
            jsx_free_outputs(S->ctx, &it->second);
            S->used -= plane_bytes(it->first);
            S->plane_pool.erase(it);
            continue;
        }
        return false;
    }
    return true;
}

/* the launch options a lane's context takes from the server's context */
static const char *const kServeOptions[] = {"blocks_per_cu", "queue_block", "steal", "defer_rhs", "split_w_min",
                                            "split_floor_max", "piece_layout", "piece_target", "mg_rel", "mg_err",
                                            "mg_profile", "mg_cost", "piece_budget", "mg_part", "triage", "lpt", "park_bucket",
                                            "park_wave", "park_lanes", "triage_walk"};

static int serve_copy_options(jsx_ctx *dst, jsx_ctx *src) {
    for (const char *k : kServeOptions) {
        double v = 0.0;
        int rc = jsx_ctx_get_option(src, k, &v);
        if (rc == JSX_OK) rc = jsx_ctx_set_option(dst, k, v);
        if (rc) return rc;
    }
    return JSX_OK;
}

/* admit one request: its grids (resident or uploaded), its descriptors
 * lowered, its device planes (pooled or allocated).  Scheduler thread, S->mu
 * held (the caches are the scheduler's); room was made for what it adds. */
static int serve_admit(jsx_server *S, ServeReq *q) {
    jsx_ctx *ctx = S->ctx;
    jsx_request &r = q->r;
    q->npix = r.image.n_across * r.image.n_along;
    for (auto &g : q->grids) {
        auto key = grid_key(g);
        auto it = S->grid_cache.find(key);
        int32_t h = 0;
        if (it == S->grid_cache.end()) {
            int rc = jsx_grid_upload(ctx, &g.desc, g.values, &h);
            if (rc) return rc;
            ServeGrid sg;
            sg.handle = h;
            sg.bytes = sizeof(double) * (size_t)g.desc.nz * (size_t)g.desc.nr;
            S->used += sg.bytes;
            it = S->grid_cache.emplace(key, sg).first;
        }
        it->second.refs++;
        it->second.last_use = ++S->use_clock;
        h = it->second.handle;
        q->grid_handles.push_back(h);
        if (g.field == JSX_FIELD_B) r.jet.grid_b = h;
        else if (g.field == JSX_FIELD_B2) r.jet.grid_b2 = h;
        else if (g.field == JSX_FIELD_N) r.jet.grid_n = h;
        else if (g.field == JSX_FIELD_V) r.jet.grid_v = h;
    }
    BatchItem &it = q->item;
    it.npix = q->npix;
    int rc;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if ((rc = lower_model(ctx, r.jet, r.params.flags, it.m))) return rc;
        if ((rc = lower_solver(ctx, r.params, it.S))) return rc;
        if ((rc = lower_image(ctx, r.image, it.im, nullptr))) return rc;
        if (S->precision == JSX_PREC_MIXED) lower_mixed(r.jet, r.params, it.m, it.im, it.c);
        else std::memset(&it.c, 0, sizeof it.c);
    }
    auto pit = S->plane_pool.lower_bound(q->npix);
    if (pit != S->plane_pool.end() && pit->first <= 2 * q->npix) {
        q->dev = pit->second;
        q->cap = pit->first;
        S->plane_pool.erase(pit);
    } else {
        if ((rc = jsx_alloc_outputs(ctx, q->npix, JSX_STOKES_FULL, &q->dev))) return rc;
        q->cap = q->npix;
        S->used += plane_bytes(q->npix);
    }
    it.out = lower_out(q->dev);
    q->t_admit = serve_now();
    return JSX_OK;
}

/* copy a finished request's planes to the caller (S->mu not held) */
static void serve_copy_out(ServeReq *q, int status) {
    size_t nd = (size_t)q->npix * sizeof(double), nu = (size_t)q->npix * sizeof(uint32_t);
    const jsx_outputs &h = q->r.out_host;
    q->status = status;
    if (status == JSX_OK && q->dev.I) {
        struct { void *h; void *d; size_t b; } cp[] = {
            {h.tau, q->dev.tau, nd}, {h.I, q->dev.I, nd}, {h.Q, q->dev.Q, nd}, {h.U, q->dev.U, nd},
            {h.V, q->dev.V, nd}, {h.l, q->dev.l, nd}, {h.flags, q->dev.flags, nu}, {h.n_rhs, q->dev.n_rhs, nu},
            {h.n_steps, q->dev.n_steps, nu}};
        for (auto &c : cp)
            if (c.h && hipMemcpy(c.h, c.d, c.b, hipMemcpyDeviceToHost) != hipSuccess) q->status = JSX_EHIP;
    }
}

/* a request leaves the device: planes back to the pool, grid references
 * dropped; both stay held until their memory is needed (S->mu held) */
static void serve_release(jsx_server *S, ServeReq *q) {
    if (q->dev.I) {
        S->plane_pool.emplace(q->cap, q->dev);
        q->dev = jsx_outputs{};
    }
    for (int32_t h : q->grid_handles)
        for (auto &kv : S->grid_cache)
            if (kv.second.handle == h) { kv.second.refs--; break; }
    q->grid_handles.clear();
}

/* fail every request not yet done (S->mu held) and stop admitting: after a
 * HIP error the device state of in-flight launches is unknown */
static void serve_fail_all(jsx_server *S, const std::string &why) {
    S->dead = true;
    S->err = why;
    auto kill = [&](ServeReq *q) {
        if (q->done) return;
        q->status = JSX_EHIP;
        q->done = true;
        q->t_done = serve_now();
    };
    for (auto *q : S->pending) kill(q);
    S->pending.clear();
    for (auto *q : S->ready) { serve_release(S, q); kill(q); }
    S->ready.clear();
    for (auto &L : S->lanes) {
        for (auto *q : L.reqs) { serve_release(S, q); kill(q); }
        L.reqs.clear();
        L.busy = false;
    }
    S->cv_done.notify_all();
}

static void serve_loop(jsx_server *S) {
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        /* 1. completions */
        bool finished = false;
        for (auto &L : S->lanes) {
            if (!L.busy) continue;
            hipError_t e = hipEventQuery(L.done);
            if (e == hipErrorNotReady) continue;
            std::vector<ServeReq *> reqs;
            reqs.swap(L.reqs);
            L.busy = false;
            lk.unlock();
            for (auto *q : reqs) serve_copy_out(q, e == hipSuccess ? JSX_OK : JSX_EHIP);
            lk.lock();
            for (auto *q : reqs) {
                serve_release(S, q);
                q->t_done = serve_now();
                q->done = true;
                S->n_active--;
            }
            finished = true;
            if (e != hipSuccess) {
                serve_fail_all(S, hipGetErrorString(e));
                return;
            }
        }
        if (finished) S->cv_done.notify_all();
        /* 2. admission in arrival order */
        while (!S->pending.empty() && S->n_active < S->max_active) {
            ServeReq *q = S->pending.front();
            if (q->footprint > S->budget) {
                q->status = JSX_EBUDGET; q->done = true; q->t_done = serve_now();
                S->pending.pop_front();
                S->cv_done.notify_all();
                continue;
            }
            const size_t need = serve_need(S, q);
            if (S->used + need > S->budget && !serve_make_room(S, need)) break;
            S->pending.pop_front();
            S->n_active++;
            int rc = serve_admit(S, q);
            if (rc) {
                serve_release(S, q);
                q->status = rc; q->done = true; q->t_done = serve_now();
                S->n_active--;
                S->cv_done.notify_all();
            } else {
                S->ready.push_back(q);
            }
        }
        /* 3. launch the admitted requests on a free lane */
        ServeLane *free_lane = nullptr;
        for (auto &L : S->lanes)
            if (!L.busy) { free_lane = &L; break; }
        if (!S->ready.empty() && free_lane) {
            ServeLane &L = *free_lane;
            L.reqs.swap(S->ready);
            L.busy = true;
            S->epochs++;
            lk.unlock();
            std::vector<BatchItem> items;
            items.reserve(L.reqs.size());
            for (auto *q : L.reqs) items.push_back(q->item);
            int rc = serve_copy_options(L.sub, S->ctx);
            if (rc == JSX_OK) {
                std::lock_guard<std::mutex> lk2(L.sub->mu);
                rc = launch_batch(L.sub, L.st, S->precision, S->stokes, items, 1);
            }
            hipError_t e = rc == JSX_OK ? hipEventRecord(L.done, L.st) : hipErrorUnknown;
            lk.lock();
            if (rc != JSX_OK || e != hipSuccess) {
                serve_fail_all(S, rc != JSX_OK ? std::string(jsx_last_error(L.sub)) : std::string(hipGetErrorString(e)));
                return;
            }
            continue;
        }
        bool busy = false;
        for (auto &L : S->lanes) busy = busy || L.busy;
        if (!busy && S->ready.empty() && S->pending.empty()) {
            if (S->stop) return;
            S->cv.wait(lk);
        } else if (busy) {
            /* poll the launches' events; a submission wakes the loop early */
            S->cv.wait_for(lk, std::chrono::microseconds(200));
        } else {
            /* pending requests that do not fit yet: wait for room */
            S->cv.wait_for(lk, std::chrono::milliseconds(1));
        }
    }
}

extern "C" int jsx_server_create(jsx_ctx *ctx, int32_t precision, int32_t stokes, int32_t max_active,
                                 size_t hbm_budget, int32_t epoch_budget, jsx_server **out) {
    if (!ctx || !out || max_active <= 0) return JSX_EINVAL;
    if ((precision != JSX_PREC_FP64 && precision != JSX_PREC_MIXED) || (stokes != JSX_STOKES_I && stokes != JSX_STOKES_FULL))
        return fail(ctx, JSX_EINVAL, "bad precision/stokes");
    (void)epoch_budget;   /* launches are not cut into epochs */
    jsx_server *S = new jsx_server();
    S->ctx = ctx; S->precision = precision; S->stokes = stokes; S->max_active = max_active;
    S->budget = hbm_budget ? hbm_budget : ctx->budget / 2;
    (void)hipSetDevice(ctx->device);
    const int n_lanes = 2;
    S->lanes.resize(n_lanes);
    for (auto &L : S->lanes) {
        int rc = jsx_ctx_create(ctx->device, ctx->budget, &L.sub);
        if (rc) { jsx_server_destroy(S); return fail(ctx, rc, "server launch lane: context creation failed"); }
        if (hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) {
            jsx_server_destroy(S);
            return fail(ctx, JSX_EHIP, "server launch lane: stream creation failed");
        }
    }
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
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->dead) return fail(S->ctx, JSX_EHIP, "server stopped after a HIP error: %s", S->err.c_str());
    ServeReq *q = new ServeReq();
    q->r = *req;
    q->grids.assign(grids, grids + n_grids);
    q->t_submit = serve_now();
    q->footprint = req_footprint(q);
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
    for (auto &L : S->lanes) {
        if (L.st) (void)hipStreamSynchronize(L.st);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.st) (void)hipStreamDestroy(L.st);
        if (L.sub) jsx_ctx_destroy(L.sub);
    }
    for (auto &kv : S->all)
        if (kv.second->dev.I) jsx_free_outputs(S->ctx, &kv.second->dev);
    for (auto &kv : S->plane_pool) jsx_free_outputs(S->ctx, &kv.second);
    for (auto &kv : S->grid_cache) jsx_grid_release(S->ctx, kv.second.handle);
    for (auto &kv : S->all) delete kv.second;
    delete S;
    return JSX_OK;
}

#endif /* JSX_SERVE_H */
