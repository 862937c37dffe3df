/*
 * jsx_host.hip -- the C ABI (include/jetshow_amd.h): contexts, descriptor
 * lowering, grids in HBM, image tables, kernel launches.
 *
 * Lowering evaluates every parameter-only constant with the C library in the
 * same expression form the reference evaluates it on each call
 * (src/utils.cpp:40-144, src/Cone.cpp:22,69-70, src/VField.cpp:28), so the
 * kernels see exactly the values the reference would compute.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "jsx_kernels.h"
#include "jsx_mixed.h"
#include "jsx_batch.h"
#include "jsx_fourier.h"
#include "jsx_delaunay.h"
#include "jsx_walk.h"


#define JSX_OPT_PERSISTENT 0x100u   /* internal: persistent queue kernel        */

struct GridRec {
    jsx_grid_desc desc;
    double *dev;
    size_t bytes;
    int hull = 0;        /* face planes of a convex hull (jsx_hull_upload) */
    /* a triangulated point set (jsx_grid_upload_points): dev holds the node
       values, z, r [pc], then the int32 triangles, bucket offsets and lists */
    int tri = 0;
    int64_t n_nodes = 0, n_tri = 0, n_list = 0;
    int nbz = 0, nbr = 0;
    double bz0 = 0, bdz = 1, br0 = 0, bdr = 1, bzl = 0, brl = 0;
};

struct ImageRec {
    double *along = nullptr, *across_half = nullptr;
    float *along_f = nullptr, *across_half_f = nullptr;
    int64_t na = 0, nl = 0;
    size_t bytes = 0;
};

/* what the batched launch's kernel sequence depends on (compared bytewise) */
struct BatchGraphKey {
    int precision, stokes, order, nreq;
    int64_t total;
    int blocks, tri_blocks;
    unsigned park_lanes;
    const void *reqs, *off, *outs, *queues, *live, *sorted, *bucket, *sbucket;
    int64_t queue_block;
    int steal;
    unsigned defer_rhs;
    int park_wave, park_bucket, piece_layout;
    unsigned piece_budget, mg_part;
    int n_cu, split, triage_walk;
    SplitWS ws;
};

struct jsx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0, used = 0;
    std::string err;
    std::mutex mu;
    std::map<int32_t, GridRec> grids;
    int32_t next_grid = 1;
    std::map<std::vector<double>, ImageRec> images;
    unsigned long long *d_head = nullptr;
    /* ring of launch-constant slots: pinned host mirror -> device, one event per slot */
    static const int kSlots = 64;
    DParamsM *h_params = nullptr, *d_params = nullptr;   /* slots sized for the larger struct */
    hipEvent_t slot_ev[kSlots];
    int slot_next = 0;
    int n_cu = 256;
    /* persistent batch launch shape (autotuned) */
    int persist_blocks_per_cu = 4;
    int64_t queue_block = 64;
    int steal = 1;
    /* batch buffers (grow-only) */
    BatchQueues *d_queues = nullptr;
    void *d_reqs = nullptr;
    size_t reqs_cap = 0;
    int64_t *d_off = nullptr;
    DOut *d_outs = nullptr;
    size_t n_cap = 0;
    /* request tables of batched launches: two pinned host buffers used in
       turn (each with the event after its copy to the device), and the event
       after the last launch (a launch on another stream waits for it before
       reusing the device buffers) */
    unsigned char *h_stage[2] = {nullptr, nullptr};
    size_t h_stage_cap[2] = {0, 0};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int stage_next = 0;
    hipEvent_t batch_ev = nullptr;
    hipStream_t batch_stream = nullptr;
    /* triage pass: live-ray list, its cost buckets and the list ordered by
       them (grow-only) */
    int triage = 2;                     /* 0 off, 1 on, 2 mixed-precision launches only */
    int lpt = 1;
    int64_t *d_live = nullptr, *d_sorted = nullptr;
    unsigned char *d_bucket = nullptr, *d_sbucket = nullptr;
    /* rays in this cost bucket or above are split from their first Stokes
       step (0: off) */
    int park_bucket = 0;
    size_t live_cap = 0;
    /* lanes still holding a ray below which long rays may be parked (0: auto),
       and the same per wavefront (0: off) */
    unsigned park_lanes = 0;
    int park_wave = 0;
    /* pieces the split planner aims at (0: auto, about two per resident piece lane) */
    int piece_target = 0;
    /* exponential pieces: right-hand sides a piece may use in each of the
       first rounds before handing the rest of its interval on (0: one round) */
    unsigned piece_budget = 64;
    /* exponential pieces in one launch: a piece that has used this many
       right-hand sides hands the rest it estimates at more than two such parts
       on to the same launch (0: the budgeted rounds above) */
    unsigned mg_part = 48;
    /* time-parallel path for long full-Stokes rays (mixed precision) */
    unsigned defer_rhs = 512;           /* 0 disables parking                  */
    float split_w_min = 256.0f;
    float split_floor_max = 1.0f;
    int piece_layout = 3;               /* 0: 5-lane groups; 1: one lane, 20 components;
                                           2: one lane, exponential (Magnus) propagator;
                                           3: 1 or 2 per ray by its rotation */
    float mg_cost = 0.5f;               /* layout 3: cost ratio below which a ray's pieces are exponential */
    int mg_profile = 0;                 /* cuts of exponential rays: 0 stiffness, 1 step estimate */
    float mg_rel = 5e-6f;               /* smallest relative tolerance of exponential steps */
    int mg_err = 0;                     /* their error estimate: 0 first order, 1 second exponential */
    SplitWS split{};
    void *split_mem = nullptr;
    /* visibilities: chunk partial sums (grow-only) and the event after their
       last reader, which a launch on another stream waits for */
    double2 *d_vis_part = nullptr;
    size_t vis_part_cap = 0;
    hipEvent_t vis_ev = nullptr;
    double2 *vis_scratch = nullptr;   /* this device's address of the file-scope `scratch` */
    int vis_trig = 1;                   /* mixed visibilities: 0 polynomial sine/cosine, 1 hardware */
    /* optional event timing of the batched launches (option "timing") */
    int timing = 0;
    /* the batched launch as a HIP graph (option "graph", default on) */
    int graph = 1;
    /* triage threads walk pixels in memory order (1) or ray ids (0) */
    int triage_walk = 1;
    /* mixed requests on regular-grid fields integrated cell by cell
       (option "grid_walk", default on) and the walk kernel's ray counter */
    int grid_walk = 1;
    unsigned long long *d_walk_ctr = nullptr;
    /* the walk launch's request table, offsets and outputs (device, and a
       pinned staging copy); walk_ev marks the last launch that read them */
    DParamsM *d_walk_reqs = nullptr;
    int64_t *d_walk_off = nullptr;
    DOut *d_walk_outs = nullptr;
    char *h_walk = nullptr;
    size_t walk_cap = 0;
    hipEvent_t walk_ev = nullptr;
    hipGraphExec_t bg_exec = nullptr;
    BatchGraphKey bg_key;
    bool bg_seen = false;
    int64_t graph_captures = 0, graph_launches = 0;
    struct EvSet { hipEvent_t e[7]; };
    std::vector<EvSet> evs;     /* created on first use, reused */
    int ev_n = 0;               /* launches recorded since timing was (re)enabled */
    static const int kEvMax = 64;
};

/* Partial sums of jsx_visibilities_device_noalloc: a file-scope device array
 * (each GPU the library runs on holds its own copy until the process ends).
 * In host code `scratch` names a host-side stand-in, not device memory: its
 * device address is taken once per device with hipGetSymbolAddress (at
 * context creation) and only that address is handed to kernels. */
static __device__ __align__(256) double2 scratch[JSX_VIS_SCRATCH_MAX];
static std::mutex g_scratch_mu;
static double2 *g_scratch_ptr[64];      /* device address of `scratch`, per device id */

/* with the device current: resolve (once) and return its address of `scratch` */
static double2 *scratch_on_device(int device) {
    if (device < 0 || device >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = (double2 *)ptr;
    }
    return g_scratch_ptr[device];
}

static int fail(jsx_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

#define HIPCHECK(ctx, call)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail((ctx), JSX_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
    } while (0)

extern "C" int jsx_abi_version(void) { return JSX_ABI_VERSION; }

extern "C" const char *jsx_last_error(jsx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int jsx_ctx_create(int device, size_t hbm_budget, jsx_ctx **out) {
    if (!out) return JSX_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return JSX_ENODEV;
    if (device < 0 || device >= ndev) return JSX_EINVAL;
    jsx_ctx *ctx = new jsx_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return JSX_ENODEV; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    ctx->budget = hbm_budget ? hbm_budget : (size_t)(0.9 * (double)free_b);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return JSX_EHIP; }
    if (hipMalloc(&ctx->d_head, 64 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&ctx->d_params, sizeof(DParamsM) * jsx_ctx::kSlots) != hipSuccess ||
        hipHostMalloc(&ctx->h_params, sizeof(DParamsM) * jsx_ctx::kSlots, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return JSX_ENOMEM;
    }
    for (int i = 0; i < jsx_ctx::kSlots; ++i) (void)hipEventCreateWithFlags(&ctx->slot_ev[i], hipEventDisableTiming);
    ctx->vis_scratch = scratch_on_device(device);
    if (!ctx->vis_scratch) {
        jsx_ctx_destroy(ctx);
        return JSX_EHIP;
    }
    *out = ctx;
    return JSX_OK;
}

extern "C" int jsx_ctx_destroy(jsx_ctx *ctx) {
    if (!ctx) return JSX_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->grids) (void)hipFree(kv.second.dev);
    for (auto &kv : ctx->images) {
        (void)hipFree(kv.second.along); (void)hipFree(kv.second.across_half);
        (void)hipFree(kv.second.along_f); (void)hipFree(kv.second.across_half_f);
    }
    (void)hipFree(ctx->d_head);
    (void)hipFree(ctx->d_params);
    (void)hipHostFree(ctx->h_params);
    for (int i = 0; i < jsx_ctx::kSlots; ++i) (void)hipEventDestroy(ctx->slot_ev[i]);
    (void)hipFree(ctx->d_queues); (void)hipFree(ctx->d_reqs); (void)hipFree(ctx->d_off); (void)hipFree(ctx->d_outs);
    for (int k = 0; k < 2; ++k) {
        (void)hipHostFree(ctx->h_stage[k]);
        if (ctx->stage_ev[k]) (void)hipEventDestroy(ctx->stage_ev[k]);
    }
    if (ctx->bg_exec) {
        if (ctx->batch_ev) (void)hipEventSynchronize(ctx->batch_ev);
        (void)hipGraphExecDestroy(ctx->bg_exec);
    }
    if (ctx->batch_ev) (void)hipEventDestroy(ctx->batch_ev);
    if (ctx->walk_ev) {
        (void)hipEventSynchronize(ctx->walk_ev);
        (void)hipEventDestroy(ctx->walk_ev);
    }
    (void)hipFree(ctx->d_walk_ctr);
    (void)hipFree(ctx->d_walk_reqs); (void)hipFree(ctx->d_walk_off); (void)hipFree(ctx->d_walk_outs);
    if (ctx->h_walk) (void)hipHostFree(ctx->h_walk);
    (void)hipFree(ctx->d_live); (void)hipFree(ctx->d_sorted); (void)hipFree(ctx->d_bucket);
    (void)hipFree(ctx->d_sbucket);
    (void)hipFree(ctx->split_mem);
    (void)hipFree(ctx->d_vis_part);
    if (ctx->vis_ev) (void)hipEventDestroy(ctx->vis_ev);
    for (auto &es : ctx->evs)
        for (int k = 0; k < 7; ++k) (void)hipEventDestroy(es.e[k]);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return JSX_OK;
}

static void split_target(jsx_ctx *ctx);

/* ---------------- options ---------------- */
extern "C" int jsx_ctx_set_option(jsx_ctx *ctx, const char *name, double v) {
    if (!ctx || !name) return JSX_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::string k(name);
    if (k == "blocks_per_cu") {
        if (!(v >= 1 && v <= 16)) return fail(ctx, JSX_EINVAL, "blocks_per_cu must be in [1, 16]");
        ctx->persist_blocks_per_cu = (int)v;
        if (ctx->split_mem) split_target(ctx);
    } else if (k == "vis_trig") {
        if (!(v == 0.0 || v == 1.0)) return fail(ctx, JSX_EINVAL, "vis_trig must be 0 or 1");
        ctx->vis_trig = (int)v;
    } else if (k == "queue_block") {
        if (!(v >= 1 && v <= (double)(1 << 20))) return fail(ctx, JSX_EINVAL, "queue_block must be in [1, 2^20]");
        ctx->queue_block = (int64_t)v;
    } else if (k == "steal") {
        ctx->steal = v != 0.0;
    } else if (k == "defer_rhs") {
        if (!(v >= 0 && v < 4e9)) return fail(ctx, JSX_EINVAL, "defer_rhs out of range");
        ctx->defer_rhs = (unsigned)v;
    } else if (k == "split_w_min") {
        if (!(v > 0)) return fail(ctx, JSX_EINVAL, "split_w_min must be positive");
        ctx->split_w_min = (float)v;
        if (ctx->split_mem) ctx->split.w_min = (float)v;
    } else if (k == "split_floor_max") {
        if (!(v >= 1.0)) return fail(ctx, JSX_EINVAL, "split_floor_max must be >= 1");
        ctx->split_floor_max = (float)v;
        if (ctx->split_mem) ctx->split.floor_max = (float)v;
    } else if (k == "piece_layout") {
        if (!(v == 0.0 || v == 1.0 || v == 2.0 || v == 3.0))
            return fail(ctx, JSX_EINVAL, "piece_layout must be 0, 1, 2 or 3");
        ctx->piece_layout = (int)v;
        if (ctx->split_mem) split_target(ctx);
    } else if (k == "piece_target") {
        if (!(v >= 0 && v <= (double)(1 << 20))) return fail(ctx, JSX_EINVAL, "piece_target must be in [0, 2^20]");
        ctx->piece_target = (int)v;
        if (ctx->split_mem) split_target(ctx);
    } else if (k == "mg_rel") {
        if (!(v >= 1e-9 && v <= 1e-3)) return fail(ctx, JSX_EINVAL, "mg_rel must be in [1e-9, 1e-3]");
        ctx->mg_rel = (float)v;
        if (ctx->split_mem) ctx->split.mg_rel = (float)v;
    } else if (k == "mg_err") {
        if (!(v == 0.0 || v == 1.0)) return fail(ctx, JSX_EINVAL, "mg_err must be 0 or 1");
        ctx->mg_err = (int)v;
        if (ctx->split_mem) ctx->split.mg_err = ctx->mg_err;
    } else if (k == "mg_profile") {
        ctx->mg_profile = v != 0.0;
        if (ctx->split_mem) ctx->split.mg_profile = ctx->mg_profile;
    } else if (k == "mg_cost") {
        if (!(v > 0)) return fail(ctx, JSX_EINVAL, "mg_cost must be positive");
        ctx->mg_cost = (float)v;
        if (ctx->split_mem) ctx->split.mg_cost = (float)v;
    } else if (k == "piece_budget") {
        if (!(v >= 0 && v < 1e9)) return fail(ctx, JSX_EINVAL, "piece_budget out of range");
        ctx->piece_budget = (unsigned)v;
    } else if (k == "mg_part") {
        if (!(v >= 0 && v < 1e6)) return fail(ctx, JSX_EINVAL, "mg_part out of range");
        ctx->mg_part = (unsigned)v;
    } else if (k == "triage") {
        if (!(v == 0.0 || v == 1.0 || v == 2.0)) return fail(ctx, JSX_EINVAL, "triage must be 0, 1 or 2");
        ctx->triage = (int)v;
    } else if (k == "lpt") {
        ctx->lpt = v != 0.0;
    } else if (k == "park_bucket") {
        if (!(v >= 0 && v < 32)) return fail(ctx, JSX_EINVAL, "park_bucket must be in [0, 31]");
        ctx->park_bucket = (int)v;
    } else if (k == "park_wave") {
        if (!(v >= 0 && v <= 64)) return fail(ctx, JSX_EINVAL, "park_wave must be in [0, 64]");
        ctx->park_wave = (int)v;
    } else if (k == "park_lanes") {
        if (!(v >= 0 && v < 4e9)) return fail(ctx, JSX_EINVAL, "park_lanes out of range");
        ctx->park_lanes = (unsigned)v;
    } else if (k == "timing") {
        ctx->timing = v != 0.0;
        ctx->ev_n = 0;
    } else if (k == "graph") {
        ctx->graph = v != 0.0;
    } else if (k == "triage_walk") {
        ctx->triage_walk = v != 0.0;
    } else if (k == "grid_walk") {
        ctx->grid_walk = v != 0.0;
    } else {
        return fail(ctx, JSX_EINVAL, "unknown option '%s'", name);
    }
    return JSX_OK;
}

extern "C" int jsx_ctx_get_option(jsx_ctx *ctx, const char *name, double *v) {
    if (!ctx || !name || !v) return JSX_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::string k(name);
    if (k == "blocks_per_cu") *v = ctx->persist_blocks_per_cu;
    else if (k == "queue_block") *v = (double)ctx->queue_block;
    else if (k == "vis_trig") *v = ctx->vis_trig;
    else if (k == "steal") *v = ctx->steal;
    else if (k == "defer_rhs") *v = ctx->defer_rhs;
    else if (k == "split_w_min") *v = ctx->split_w_min;
    else if (k == "split_floor_max") *v = ctx->split_floor_max;
    else if (k == "piece_layout") *v = ctx->piece_layout;
    else if (k == "triage") *v = ctx->triage;
    else if (k == "lpt") *v = ctx->lpt;
    else if (k == "park_bucket") *v = ctx->park_bucket;
    else if (k == "piece_target") *v = ctx->piece_target;
    else if (k == "piece_budget") *v = ctx->piece_budget;
    else if (k == "mg_part") *v = ctx->mg_part;
    else if (k == "mg_cost") *v = ctx->mg_cost;
    else if (k == "mg_profile") *v = ctx->mg_profile;
    else if (k == "mg_rel") *v = ctx->mg_rel;
    else if (k == "mg_err") *v = ctx->mg_err;
    else if (k == "park_lanes") *v = ctx->park_lanes;
    else if (k == "graph") *v = ctx->graph;
    else if (k == "triage_walk") *v = ctx->triage_walk;
    else if (k == "grid_walk") *v = ctx->grid_walk;
    else if (k == "stat_graph_captures") *v = (double)ctx->graph_captures;
    else if (k == "stat_graph_launches") *v = (double)ctx->graph_launches;
    else if (k == "park_wave") *v = ctx->park_wave;
    else if (k == "stat_parked" || k == "stat_pieces") {
        unsigned c[2] = {0, 0};
        if (ctx->split_mem) {
            HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
            HIPCHECK(ctx, hipMemcpy(&c[0], ctx->split.n_long, sizeof(unsigned), hipMemcpyDeviceToHost));
            HIPCHECK(ctx, hipMemcpy(&c[1], ctx->split.n_pieces, sizeof(unsigned), hipMemcpyDeviceToHost));
        }
        unsigned parked = c[0] < (unsigned)ctx->split.cap ? c[0] : (unsigned)ctx->split.cap;
        unsigned pieces = c[1] < (unsigned)ctx->split.piece_cap ? c[1] : (unsigned)ctx->split.piece_cap;
        *v = k == "stat_parked" ? (double)parked : (double)pieces;
    } else if (k == "stat_mg_trips") {
        /* wavefront loop trips of the last launch's exponential pieces */
        unsigned long long c = 0;
        if (ctx->split_mem) {
            HIPCHECK(ctx, hipDeviceSynchronize());
            HIPCHECK(ctx, hipMemcpy(&c, ctx->split.mg_stat, sizeof c, hipMemcpyDeviceToHost));
        }
        *v = (double)c;
    } else if (k == "stat_parked_exp") {
        unsigned c = 0;
        if (ctx->split_mem) {
            HIPCHECK(ctx, hipDeviceSynchronize());
            HIPCHECK(ctx, hipMemcpy(&c, ctx->split.n_mg, sizeof c, hipMemcpyDeviceToHost));
        }
        *v = c;
    } else if (k == "stat_piece_steps_max" || k == "stat_piece_steps_mean") {
        /* step attempts (right-hand sides) of the last launch's pieces: the
           longest piece bounds the piece kernel */
        unsigned np = 0;
        double mx = 0.0, sum = 0.0;
        if (ctx->split_mem) {
            HIPCHECK(ctx, hipDeviceSynchronize());
            HIPCHECK(ctx, hipMemcpy(&np, ctx->split.n_pieces, sizeof np, hipMemcpyDeviceToHost));
            if (np > (unsigned)ctx->split.piece_cap) np = (unsigned)ctx->split.piece_cap;
            std::vector<PieceOut> po(np);
            if (np) HIPCHECK(ctx, hipMemcpy(po.data(), ctx->split.pout, sizeof(PieceOut) * np, hipMemcpyDeviceToHost));
            for (unsigned i = 0; i < np; ++i) {
                mx = po[i].n_rhs > mx ? po[i].n_rhs : mx;
                sum += po[i].n_rhs;
            }
        }
        *v = k == "stat_piece_steps_max" ? mx : (np ? sum / np : 0.0);
    } else if (k == "stat_rhs_batch" || k == "stat_rhs_pieces" || k == "stat_trips" || k == "stat_trips_uniform" ||
               k == "stat_rhs_exp" || k == "stat_trips_tail") {
        unsigned long long r = 0;
        HIPCHECK(ctx, hipDeviceSynchronize());
        if (k == "stat_rhs_batch" && ctx->d_queues)
            HIPCHECK(ctx, hipMemcpy(&r, &ctx->d_queues->rhs, sizeof r, hipMemcpyDeviceToHost));
        if (k == "stat_trips" && ctx->d_queues)
            HIPCHECK(ctx, hipMemcpy(&r, &ctx->d_queues->trips, sizeof r, hipMemcpyDeviceToHost));
        if (k == "stat_trips_uniform" && ctx->d_queues)
            HIPCHECK(ctx, hipMemcpy(&r, &ctx->d_queues->trips_uniform, sizeof r, hipMemcpyDeviceToHost));
        if (k == "stat_trips_tail" && ctx->d_queues)
            HIPCHECK(ctx, hipMemcpy(&r, &ctx->d_queues->trips_tail, sizeof r, hipMemcpyDeviceToHost));
        if (k == "stat_rhs_pieces" && ctx->split_mem)
            HIPCHECK(ctx, hipMemcpy(&r, ctx->split.rhs_pieces, sizeof r, hipMemcpyDeviceToHost));
        if (k == "stat_rhs_exp" && ctx->split_mem)
            HIPCHECK(ctx, hipMemcpy(&r, ctx->split.rhs_mg, sizeof r, hipMemcpyDeviceToHost));
        *v = (double)r;
    } else if (k == "stat_launches") {
        *v = ctx->ev_n;
    } else if (k == "stat_batch_ms" || k == "stat_piece_ms" || k == "stat_split_ms" || k == "stat_triage_ms" ||
               k == "stat_piece_dp_ms" || k == "stat_piece_exp_ms") {
        /* mean over the launches recorded since timing was enabled; events:
           5 triage 0 batch 1 planner 2 Dormand-Prince pieces 6 exponential
           pieces 3 combine 4 */
        int a = 1, b = 4;
        if (k == "stat_batch_ms") { a = 0; b = 1; }
        else if (k == "stat_piece_ms") { a = 2; b = 3; }
        else if (k == "stat_triage_ms") { a = 5; b = 0; }
        else if (k == "stat_piece_dp_ms") { a = 2; b = 6; }
        else if (k == "stat_piece_exp_ms") { a = 6; b = 3; }
        double sum = 0.0;
        int n = ctx->ev_n < jsx_ctx::kEvMax ? ctx->ev_n : jsx_ctx::kEvMax;
        for (int i = 0; i < n; ++i) {
            float ms = 0.f;
            HIPCHECK(ctx, hipEventSynchronize(ctx->evs[i].e[b]));
            HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->evs[i].e[a], ctx->evs[i].e[b]));
            sum += ms;
        }
        *v = n ? sum / n : 0.0;
    } else {
        return fail(ctx, JSX_EINVAL, "unknown option '%s'", name);
    }
    return JSX_OK;
}

/* ---------------- grids ---------------- */
extern "C" int jsx_grid_upload(jsx_ctx *ctx, const jsx_grid_desc *desc, const double *values, int32_t *handle_out) {
    if (!ctx || !desc || !values || !handle_out) return fail(ctx, JSX_EINVAL, "null argument");
    if (desc->nz < 2 || desc->nr < 2 || !(desc->dz > 0) || !(desc->dr > 0))
        return fail(ctx, JSX_EINVAL, "grid needs nz, nr >= 2 and positive spacing");
    std::lock_guard<std::mutex> lk(ctx->mu);
    size_t bytes = sizeof(double) * (size_t)desc->nz * (size_t)desc->nr;
    if (ctx->used + bytes > ctx->budget) return fail(ctx, JSX_EBUDGET, "grid of %zu bytes exceeds HBM budget", bytes);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    GridRec r;
    r.desc = *desc;
    r.bytes = bytes;
    HIPCHECK(ctx, hipMalloc(&r.dev, bytes));
    HIPCHECK(ctx, hipMemcpy(r.dev, values, bytes, hipMemcpyHostToDevice));
    ctx->used += bytes;
    int32_t h = ctx->next_grid++;
    ctx->grids[h] = r;
    *handle_out = h;
    return JSX_OK;
}

extern "C" int jsx_triangulate(const double *zr, int64_t n_points, int32_t *tri, int64_t cap, int64_t *n_tri) {
    if (!zr || !n_tri || n_points < 0 || n_points > 0x7fffffff || (cap > 0 && !tri)) return JSX_EINVAL;
    std::vector<double> z((size_t)n_points), r((size_t)n_points);
    for (int64_t i = 0; i < n_points; ++i) {
        z[i] = zr[2 * i];
        r[i] = zr[2 * i + 1];
        if (!std::isfinite(z[i]) || !std::isfinite(r[i])) return JSX_EINVAL;
    }
    std::vector<int32_t> T;
    if (!jsx_del::triangulate(z, r, T)) {
        *n_tri = 0;
        return JSX_EUNSUPPORTED;
    }
    *n_tri = (int64_t)T.size() / 3;
    if (cap == 0) return JSX_OK;
    if (cap < *n_tri) return JSX_EINVAL;
    std::memcpy(tri, T.data(), sizeof(int32_t) * T.size());
    return JSX_OK;
}

extern "C" int jsx_grid_upload_points(jsx_ctx *ctx, const double *zrv, int64_t n_points, double nan_value,
                                      int32_t *handle_out) {
    if (!ctx || !zrv || !handle_out) return fail(ctx, JSX_EINVAL, "null argument");
    if (n_points < 3 || n_points > 0x7fffffff) return fail(ctx, JSX_EINVAL, "need 3 to 2^31 - 1 points");
    /* nodes that form a complete regular grid keep the regular layout */
    jsx_grid_desc d;
    int rc = jsx_grid_from_points(zrv, n_points, nan_value, &d, nullptr, 0);
    if (rc == JSX_OK) {
        std::vector<double> vals((size_t)d.nz * (size_t)d.nr);
        if ((rc = jsx_grid_from_points(zrv, n_points, nan_value, &d, vals.data(), (int64_t)vals.size())))
            return fail(ctx, rc, "regular grid points");
        return jsx_grid_upload(ctx, &d, vals.data(), handle_out);
    }
    if (rc != JSX_EUNSUPPORTED) return fail(ctx, rc, "bad field points");
    /* create_triangulation (src/SimulationInterpolater.cpp:105-117):
       coordinates divided by pc */
    const double PC = 3.0856775814671913e+18;
    const size_t n = (size_t)n_points;
    std::vector<double> z(n), r(n), v(n);
    for (size_t i = 0; i < n; ++i) {
        z[i] = zrv[3 * i] / PC;
        r[i] = zrv[3 * i + 1] / PC;
        v[i] = zrv[3 * i + 2];
        if (!std::isfinite(z[i]) || !std::isfinite(r[i])) return fail(ctx, JSX_EINVAL, "non-finite point");
    }
    std::vector<int32_t> T;
    if (!jsx_del::triangulate(z, r, T))
        return fail(ctx, JSX_EUNSUPPORTED, "the points span no triangle (fewer than three, or all on one line)");
    jsx_del::Buckets B;
    jsx_del::build_buckets(z, r, T, B);
    GridRec g;
    std::memset(&g.desc, 0, sizeof g.desc);
    g.desc.nan_value = nan_value;
    g.tri = 1;
    g.n_nodes = (int64_t)n;
    g.n_tri = (int64_t)T.size() / 3;
    g.n_list = (int64_t)B.list.size();
    g.nbz = B.nbz; g.nbr = B.nbr;
    g.bz0 = B.z0; g.bdz = B.dz; g.br0 = B.r0; g.bdr = B.dr; g.bzl = B.zlast; g.brl = B.rlast;
    const size_t nd = 3 * n, ni = T.size() + B.off.size() + B.list.size();
    g.bytes = sizeof(double) * nd + sizeof(int32_t) * ni;
    std::vector<char> blob(g.bytes);
    double *dp = (double *)blob.data();
    std::memcpy(dp, v.data(), sizeof(double) * n);
    std::memcpy(dp + n, z.data(), sizeof(double) * n);
    std::memcpy(dp + 2 * n, r.data(), sizeof(double) * n);
    int32_t *ip = (int32_t *)(dp + nd);
    std::memcpy(ip, T.data(), sizeof(int32_t) * T.size());
    std::memcpy(ip + T.size(), B.off.data(), sizeof(int32_t) * B.off.size());
    std::memcpy(ip + T.size() + B.off.size(), B.list.data(), sizeof(int32_t) * B.list.size());
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->used + g.bytes > ctx->budget)
        return fail(ctx, JSX_EBUDGET, "triangulated grid of %zu bytes exceeds HBM budget", g.bytes);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    HIPCHECK(ctx, hipMalloc(&g.dev, g.bytes));
    HIPCHECK(ctx, hipMemcpy(g.dev, blob.data(), g.bytes, hipMemcpyHostToDevice));
    ctx->used += g.bytes;
    const int32_t h = ctx->next_grid++;
    ctx->grids[h] = g;
    *handle_out = h;
    return JSX_OK;
}

extern "C" int jsx_hull_upload(jsx_ctx *ctx, const double *planes, int32_t n_planes, int32_t *handle_out) {
    if (!ctx || !planes || !handle_out) return fail(ctx, JSX_EINVAL, "null argument");
    if (n_planes < 4) return fail(ctx, JSX_EINVAL, "a bounded convex hull needs at least 4 planes");
    for (int64_t i = 0; i < 4 * (int64_t)n_planes; ++i)
        if (!std::isfinite(planes[i])) return fail(ctx, JSX_EINVAL, "non-finite plane coefficient");
    jsx_grid_desc d;
    std::memset(&d, 0, sizeof d);
    d.nz = n_planes;
    d.nr = 4;
    d.dz = d.dr = 1.0;
    int rc = jsx_grid_upload(ctx, &d, planes, handle_out);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->grids[*handle_out].hull = 1;
    return JSX_OK;
}

extern "C" int jsx_grid_release(jsx_ctx *ctx, int32_t handle) {
    if (!ctx) return JSX_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->grids.find(handle);
    if (it == ctx->grids.end()) return fail(ctx, JSX_ENOTFOUND, "unknown grid handle %d", handle);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(it->second.dev);
    ctx->used -= it->second.bytes;
    ctx->grids.erase(it);
    return JSX_OK;
}

/* ---------------- lowering ---------------- */
static void h_normalize3(double *a) {
    double z = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
    if (z > 0.0) { double r = sqrt(z); a[0] /= r; a[1] /= r; a[2] /= r; }
}

static int lower_grid(jsx_ctx *ctx, int32_t h, DGrid &g, int &has) {
    has = 0;
    if (h == 0) return JSX_OK;
    auto it = ctx->grids.find(h);
    if (it == ctx->grids.end()) return fail(ctx, JSX_ENOTFOUND, "unknown grid handle %d", h);
    const jsx_grid_desc &d = it->second.desc;
    g.pz = g.pr = nullptr;
    g.tri = g.boff = g.blist = nullptr;
    g.nbz = g.nbr = 0;
    if (it->second.tri) {
        const GridRec &R = it->second;
        const size_t n = (size_t)R.n_nodes;
        g.v = R.dev;
        g.pz = R.dev + n;
        g.pr = R.dev + 2 * n;
        g.tri = (const int32_t *)(R.dev + 3 * n);
        g.boff = g.tri + 3 * R.n_tri;
        g.blist = g.boff + ((size_t)R.nbz * (size_t)R.nbr + 1);
        g.nz = 0; g.nr = 0;
        g.nbz = R.nbz; g.nbr = R.nbr;
        g.z0 = R.bz0; g.dz = R.bdz; g.r0 = R.br0; g.dr = R.bdr;
        g.zlast = R.bzl; g.rlast = R.brl;
        g.nan_value = d.nan_value;
        has = 1;
        return JSX_OK;
    }
    g.v = it->second.dev;
    g.nz = d.nz; g.nr = d.nr;
    g.z0 = d.z0; g.dz = d.dz; g.r0 = d.r0; g.dr = d.dr; g.nan_value = d.nan_value;
    g.zlast = d.z0 + (double)(d.nz - 1) * d.dz;
    g.rlast = d.r0 + (double)(d.nr - 1) * d.dr;
    has = 1;
    return JSX_OK;
}

/* parameter-only constants of the further reference classes, evaluated with
 * the C library as their constructors and bf()/vf() evaluate them */
static void lower_model_ext(const jsx_jet_desc &j, DModel &m) {
    const double PC = 3.0856775814671913e+18;
    m.ext = j.geometry == JSX_GEOM_PARABOLOID_CONE || j.bfield >= JSX_B_CONST_CYLINDER ||
            j.nfield >= JSX_N_COMPOSITE_BK || j.vfield >= JSX_V_SHEARED_FLAT;
    if (j.geometry == JSX_GEOM_PARABOLOID_CONE) {
        /* ParabaloidCone.cpp:12-16: the cone below the paraboloid's tangent at z0 */
        const double r0 = j.geo_param, z0 = j.geo_z0;
        const double shift = -(r0 * sqrt(z0 / PC) / (0.5 * r0 * pow(z0 * PC, -0.5)) - z0);
        const double angle = atan(0.5 * r0 * pow(z0 * PC, -0.5));
        m.origin[0] = m.origin[0] + 0.0;
        m.origin[1] = m.origin[1] + 0.0;
        m.origin[2] = m.origin[2] + shift;
        m.cos_angle = cos(angle);
        const double cc = cos(angle) * cos(angle);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) m.M[r][c] = m.axis[r] * m.axis[c] - cc * (r == c ? 1.0 : 0.0);
        m.geo_z0 = z0;
    }
    m.b_negexp2 = -j.b_exp2; m.b_z0 = j.b_z0;
    m.b_pitch = j.b_pitch; m.tan_pitch = tan(j.b_pitch);
    if (j.bfield == JSX_B_FORCE_FREE_CYLINDRICAL) jsx_bessel_i01(j.b_mu, &m.b_i0, &m.b_i1);
    m.n_negexp2 = -j.n_exp2; m.n_z0 = j.n_z0;
    m.v_r = j.v_r; m.v_theta = j.v_theta; m.v_R0 = j.v_R0; m.v_Rz0 = j.v_Rz0; m.v_z0 = j.v_z0;
    if (j.vfield == JSX_V_ACC_PARABOLIC_CONST_CONE || j.vfield == JSX_V_SHEARED_ACC_PARABOLIC_CONST_CONE) {
        /* VField.cpp:222-257: the paraboloid part has R0 = z0; the cone part is
           central from (0, 0, -oz) with the paraboloid's tangent angle */
        const double Rz0 = j.v_Rz0, z0 = j.v_z0;
        m.v_R0 = z0;
        m.v_origin[0] = 0.0; m.v_origin[1] = 0.0;
        m.v_origin[2] = Rz0 * sqrt(z0 / PC) / (0.5 * Rz0 * pow(z0 * PC, -0.5)) - z0;
        if (j.vfield == JSX_V_SHEARED_ACC_PARABOLIC_CONST_CONE) m.v_theta = atan(sqrt(Rz0 * Rz0 * z0 / PC) / z0);
    }
}

static int lower_model(jsx_ctx *ctx, const jsx_jet_desc &j, unsigned opt, DModel &m) {
    std::memset(&m, 0, sizeof m);
    if (j.geometry < JSX_GEOM_CONE || j.geometry > JSX_GEOM_PARABOLOID_CONE)
        return fail(ctx, JSX_EUNSUPPORTED, "geometry kind %d", j.geometry);
    if (j.bfield < JSX_B_RANDOM_SCALAR || j.bfield > JSX_B_COMPOSITE_RANDOM_SCALAR_Z)
        return fail(ctx, JSX_EUNSUPPORTED, "bfield kind %d", j.bfield);
    if (j.nfield < JSX_N_CONST || j.nfield > JSX_N_COMPOSITE_BK)
        return fail(ctx, JSX_EUNSUPPORTED, "nfield kind %d", j.nfield);
    if (j.vfield < JSX_V_CONST_FLAT || j.vfield > JSX_V_SHEARED_ACC_PARABOLIC_CONST_CONE)
        return fail(ctx, JSX_EUNSUPPORTED, "vfield kind %d", j.vfield);
    m.geometry = j.geometry; m.bfield = j.bfield; m.nfield = j.nfield; m.vfield = j.vfield;
    m.b_plasma = j.b_in_plasma_frame; m.n_plasma = j.n_in_plasma_frame;
    m.opt = (int)opt;
    for (int i = 0; i < 3; ++i) { m.origin[i] = j.geo_origin[i]; m.axis[i] = j.geo_axis[i]; m.v_origin[i] = j.v_origin[i]; }
    h_normalize3(m.axis);
    m.cos_angle = cos(j.geo_param);
    double cc = cos(j.geo_param) * cos(j.geo_param);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m.M[r][c] = m.axis[r] * m.axis[c] - cc * (r == c ? 1.0 : 0.0);
    m.geo_param = j.geo_param;
    m.big_scale = j.geometry == JSX_GEOM_CYLINDER ? 100. * j.geo_param : j.geo_big_scale;
    m.clip = j.z_clip_max > j.z_clip_min;
    m.zmin = j.z_clip_min; m.zmax = j.z_clip_max;
    m.b0 = j.b0; m.b_negexp = -j.b_exp;
    m.sin_pitch = sin(j.b_pitch); m.cos_pitch = cos(j.b_pitch);
    m.n0 = j.n0; m.n_negexp = -j.n_exp;
    m.thermal_fraction = j.thermal_fraction;
    m.gamma0 = j.gamma0; m.gamma1 = j.gamma1; m.gz0 = j.gamma_z0; m.gz1 = j.gamma_z1;
    m.vz_flat = 29979245800.0 * sqrt(1. - 1. / (j.gamma0 * j.gamma0));
    const double s = j.s, PI = 3.141592653589793238462643383279502884;
    m.f_ki_v = (pow(3., (s + 1.) / 2.) / 4.) * tgamma(s / 4. + 11. / 6.) * tgamma(s / 4. + 1. / 6.);
    m.f_ki_s = m.f_ki_v * (sqrt(PI / 4.) * tgamma((6. + s) / 4.) / tgamma((8. + s) / 4.));
    m.f_ei_v = pow(3., s / 2.) / (2. * (s + 1)) * tgamma(s / 4. + 19. / 12.) * tgamma(s / 4. - 1. / 12.);
    m.f_ei_s = m.f_ei_v * (sqrt(PI / 4.) * tgamma((5. + s) / 4.) / tgamma((7. + s) / 4.));
    m.r_kq = (s + 2.) / (s + 10. / 3);
    m.r_eq = (s + 1.0) / (s + 7. / 3.);
    m.f_kv = pow(3., s / 2.) * (s + 3.) * (s + 2.) / (4. * (s + 1.)) * tgamma(s / 4. + 11. / 12.) * tgamma(s / 4. + 7. / 12.);
    m.f_ev = pow(3., (s - 1.) / 2.) * (s + 2.) / (2. * s) * tgamma(s / 4. + 2. / 3.) * tgamma(s / 4. + 1. / 3.);
    const double gm = j.gamma_min;
    m.f_kFr = (s + 2.) * log(gm) / ((s + 1.) * pow(gm, s + 1.));
    m.kCr_pre = 2. / (s - 2.);
    m.kCr_gm = pow(gm, 2. - s);
    m.e_ki = s / 2.;
    m.e_ei = (s - 1.) / 2.;
    m.e_kv = (s + 1.) / 2.;
    m.e_kCr = (s - 2.) / 2.;
    lower_model_ext(j, m);
    int rc;
    if (j.geometry == JSX_GEOM_CONVEX_HULL) {
        auto it = ctx->grids.find(j.geo_table);
        if (it == ctx->grids.end() || !it->second.hull)
            return fail(ctx, JSX_ENOTFOUND, "geo_table %d is not a hull handle", j.geo_table);
        if ((rc = lower_grid(ctx, j.geo_table, m.hull, m.has_hull))) return rc;
    }
    if (j.bfield == JSX_B_GRID_SCALAR || j.bfield == JSX_B_GRID_VECTOR) {
        if ((rc = lower_grid(ctx, j.grid_b, m.gb, m.has_gb))) return rc;
        if (j.bfield == JSX_B_GRID_VECTOR && (rc = lower_grid(ctx, j.grid_b2, m.gb2, m.has_gb2))) return rc;
    }
    if (j.nfield == JSX_N_GRID && (rc = lower_grid(ctx, j.grid_n, m.gn, m.has_gn))) return rc;
    if (j.vfield == JSX_V_GRID && (rc = lower_grid(ctx, j.grid_v, m.gv, m.has_gv))) return rc;
    /* triangulated point sets are looked up by the kernels built for the
       further kinds */
    if ((m.has_gb && m.gb.nz == 0) || (m.has_gb2 && m.gb2.nz == 0) || (m.has_gn && m.gn.nz == 0) ||
        (m.has_gv && m.gv.nz == 0))
        m.ext = 1;
    return JSX_OK;
}

static int lower_solver(jsx_ctx *ctx, const jsx_solver_params &p, DSolver &S) {
    if (!(p.tau_max > 0)) return fail(ctx, JSX_EINVAL, "tau_max must be > 0");
    if (p.n < 1) return fail(ctx, JSX_EINVAL, "n must be >= 1");
    if (!(p.nu > 0)) return fail(ctx, JSX_EINVAL, "nu must be > 0");
    if (p.stokes != JSX_STOKES_I && p.stokes != JSX_STOKES_FULL) return fail(ctx, JSX_EINVAL, "stokes must be 1 or 4");
    if (p.precision != JSX_PREC_FP64 && p.precision != JSX_PREC_MIXED)
        return fail(ctx, JSX_EINVAL, "precision must be 0 or 1");
    if (!(p.tau_n_min > 0) || !(p.tau_n_max > p.tau_n_min)) return fail(ctx, JSX_EINVAL, "bad steps_schedule thresholds");
    std::memset(&S, 0, sizeof S);
    S.nu = p.nu; S.tau_max = p.tau_max; S.tau_min = p.tau_min; S.dt_max = p.dt_max;
    S.tau_abs = p.tau_abs_tol; S.tau_rel = p.tau_rel_tol; S.st_abs = p.stokes_abs_tol; S.st_rel = p.stokes_rel_tol;
    S.tau_prec = p.tau_precision;
    S.tau_n_min = p.tau_n_min; S.tau_n_max = p.tau_n_max;
    S.n = p.n;
    S.n_max = p.n_max_factor * p.n;
    S.max_fail = p.max_failed_tries;
    S.stokes = p.stokes;
    S.sched_b = -log(p.tau_n_min);
    S.sched_k = S.n_max / pow((log(p.tau_n_max) + S.sched_b), 2.0);
    S.emin = pow(5.0, -5.0);
    return JSX_OK;
}

/* image plane tables: include/logspace.h:25-33, src/Image.cpp:56-126 */
static int image_tables(const jsx_image_desc &im, std::vector<double> &size_along, std::vector<double> &along,
                        std::vector<double> &across_half) {
    if (im.n_across <= 0 || im.n_along <= 0 || (im.n_across % 2) != 0) return JSX_EINVAL;
    int64_t na = im.n_across, nl = im.n_along, h = na / 2;
    size_along.resize(nl);
    along.resize(nl);
    across_half.resize((size_t)(h * nl));
    double cur = pow(10, im.lg_pixel_size_start_cm);
    double base = pow(10, (im.lg_pixel_size_stop_cm - im.lg_pixel_size_start_cm) / (int)(nl - 1));
    for (int64_t j = 0; j < nl; ++j) { size_along[j] = cur; cur *= base; }
    double acc = 0.0;
    for (int64_t j = 0; j < nl; ++j) {
        acc = j == 0 ? size_along[0] : acc + size_along[j];
        along[j] = acc - 0.5 * size_along[j];
    }
    for (int64_t col = 0; col < nl; ++col) {
        double sz = size_along[col] * 1.0;
        double c = 0.0;
        for (int64_t k = 0; k < h; ++k) {
            c = k == 0 ? sz : c + sz;
            across_half[k * nl + col] = c - 0.5 * sz;
        }
    }
    return JSX_OK;
}

extern "C" int jsx_image_pixel_sizes(const jsx_image_desc *image, double *sizes) {
    if (!image || !sizes) return JSX_EINVAL;
    std::vector<double> sa, al, ah;
    int rc = image_tables(*image, sa, al, ah);
    if (rc) return rc;
    for (int64_t i = 0; i < image->n_across; ++i)
        for (int64_t j = 0; j < image->n_along; ++j) sizes[i * image->n_along + j] = sa[j];
    return JSX_OK;
}

static int lower_image(jsx_ctx *ctx, const jsx_image_desc &im, DImage &D, ImageRec **rec_out) {
    std::vector<double> key = {(double)im.n_across, (double)im.n_along, im.lg_pixel_size_start_cm,
                               im.lg_pixel_size_stop_cm};
    auto it = ctx->images.find(key);
    ImageRec *rec;
    if (it == ctx->images.end()) {
        std::vector<double> sa, al, ah;
        if (image_tables(im, sa, al, ah)) return fail(ctx, JSX_EINVAL, "image needs n_along > 0 and even n_across > 0");
        ImageRec r;
        r.na = im.n_across; r.nl = im.n_along;
        r.bytes = sizeof(double) * (al.size() + ah.size());
        HIPCHECK(ctx, hipMalloc(&r.along, sizeof(double) * al.size()));
        HIPCHECK(ctx, hipMalloc(&r.across_half, sizeof(double) * (ah.size() ? ah.size() : 1)));
        HIPCHECK(ctx, hipMemcpy(r.along, al.data(), sizeof(double) * al.size(), hipMemcpyHostToDevice));
        if (ah.size()) HIPCHECK(ctx, hipMemcpy(r.across_half, ah.data(), sizeof(double) * ah.size(), hipMemcpyHostToDevice));
        std::vector<float> alf(al.begin(), al.end()), ahf(ah.begin(), ah.end());
        HIPCHECK(ctx, hipMalloc(&r.along_f, sizeof(float) * alf.size()));
        HIPCHECK(ctx, hipMalloc(&r.across_half_f, sizeof(float) * (ahf.size() ? ahf.size() : 1)));
        HIPCHECK(ctx, hipMemcpy(r.along_f, alf.data(), sizeof(float) * alf.size(), hipMemcpyHostToDevice));
        if (ahf.size()) HIPCHECK(ctx, hipMemcpy(r.across_half_f, ahf.data(), sizeof(float) * ahf.size(), hipMemcpyHostToDevice));
        rec = &(ctx->images[key] = r);
    } else {
        rec = &it->second;
    }
    D.along = rec->along;
    D.across_half = rec->across_half;
    D.na = im.n_across; D.nl = im.n_along;
    double raw[3] = {-sin(im.los_angle), 0, -cos(im.los_angle)};
    h_normalize3(raw);
    for (int i = 0; i < 3; ++i) { D.dir[i] = raw[i]; D.d2[i] = raw[i]; }
    h_normalize3(D.d2);
    D.zscale = 1. / sin(im.los_angle);
    if (rec_out) *rec_out = rec;
    return JSX_OK;
}

static DOut lower_out(const jsx_outputs &o) {
    DOut d;
    d.tau = o.tau; d.I = o.I; d.Q = o.Q; d.U = o.U; d.V = o.V; d.l = o.l;
    d.flags = o.flags; d.n_rhs = o.n_rhs; d.n_steps = o.n_steps;
    return d;
}

/* stage launch constants into a device slot on stream st */
/* Launch constants go through a ring of device slots.  A slot is reused only
 * after the kernel that read it has finished: the caller records slot_ev[k]
 * on its stream after that launch (stage_done), so a later launch on any
 * stream waits for the earlier reader, not only for the earlier copy. */
static int stage_bytes(jsx_ctx *ctx, const void *src, size_t n, hipStream_t st, void **dev, int *slot) {
    int k = ctx->slot_next;
    ctx->slot_next = (k + 1) % jsx_ctx::kSlots;
    HIPCHECK(ctx, hipEventSynchronize(ctx->slot_ev[k]));
    DParamsM *h = ctx->h_params + k;
    std::memcpy((void *)h, src, n);
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d_params + k, h, n, hipMemcpyHostToDevice, st));
    *dev = (void *)(ctx->d_params + k);
    *slot = k;
    return JSX_OK;
}

static int stage_done(jsx_ctx *ctx, int slot, hipStream_t st) {
    HIPCHECK(ctx, hipEventRecord(ctx->slot_ev[slot], st));
    return JSX_OK;
}

static int stage_params(jsx_ctx *ctx, const DModel &m, const DImage &D, const DSolver &S, hipStream_t st,
                        DParams **dev, int *slot) {
    DParams h;
    h.m = m; h.im = D; h.S = S;
    return stage_bytes(ctx, &h, sizeof h, st, (void **)dev, slot);
}

#ifndef JSX_MIXED_TAU_FLOOR
#define JSX_MIXED_TAU_FLOOR 1e-7
#endif
#ifndef JSX_MIXED_STOKES_FLOOR_FULL
#define JSX_MIXED_STOKES_FLOOR_FULL 1e-7
#endif
/* the smallest relative tolerance the mixed path holds the Stokes state to
   (the state is FP32 + compensation, ~1e-14) */
#ifndef JSX_MIXED_STATE_FLOOR
#define JSX_MIXED_STATE_FLOOR 1e-10
#endif
#ifndef JSX_MIXED_STOKES_FLOOR_I
#define JSX_MIXED_STOKES_FLOOR_I 1e-7
#endif
/* decay exponent from the observer at which the mixed Stokes pass starts
   (JSX_PIX_HORIZON): Stokes I past tau = 50; full Stokes past
   int (k_I - |eta|) ds = 50, the bound on the transfer operator's norm decay
   (dichroism eta = (k_Q, k_U, k_V)).  Emission from deeper reaches the
   observer attenuated by at least e^-50 = 2e-22. */
/* longest mixed Stokes-pass step through grid fields, in cells of the
   finest grid spacing (requests tighter than the FP32 floor) */
#ifndef JSX_MIXED_GRID_STEP
#define JSX_MIXED_GRID_STEP 4.0
#endif
#ifndef JSX_MIXED_TAU_HORIZON
#define JSX_MIXED_TAU_HORIZON 50.0
#endif

/* FP32 constants of the mixed path (see jsx_mixed.h for the formulas) */
static void lower_mixed(const jsx_jet_desc &j, const jsx_solver_params &p, const DModel &m, const DImage &D,
                        MConsts &c) {
    std::memset(&c, 0, sizeof c);
    const double PI = 3.141592653589793238462643383279502884, PC = 3.0856775814671913e+18;
    const double ME = 9.10938356e-28, QE = 4.80320467299766e-10, CL = 29979245800.0;
    const double CB = QE / (2. * PI * ME * CL), CP = QE * QE / (PI * ME), nu = p.nu, s = j.s;
    const bool vec = !(j.bfield == JSX_B_RANDOM_SCALAR || j.bfield == JSX_B_RANDOM_SCALAR_Z ||
                       j.bfield == JSX_B_GRID_SCALAR || j.bfield == JSX_B_COMPOSITE_RANDOM_SCALAR_Z);
    const double fk = vec ? m.f_ki_v : m.f_ki_s, fe = vec ? m.f_ei_v : m.f_ei_s;
    const double g = CB / nu;
    c.A_k = (float)(PI * CP * CB * fk * pow(g, s / 2.) / (CL * nu * nu) * PC);
    c.A_e = (float)(PI * CP * CB * ME * fe * pow(g, (s - 1.) / 2.) / CL * PC);
    c.A_kv = (float)(PI * CP * CB * m.f_kv * pow(g, (s + 1.) / 2.) / (CL * nu * nu) * PC);
    c.A_ev = (float)(PI * CP * CB * ME * m.f_ev * pow(g, s / 2.) / CL * PC);
    c.A_F = (float)(2. * PI * CP * CB / (CL * nu * nu) * PC);
    c.A_C = (float)(PI * CP * CB * CB / (CL * nu * nu * nu) * PC);
    c.r_kq = (float)m.r_kq; c.r_eq = (float)m.r_eq; c.f_kFr = (float)m.f_kFr;
    {   /* k_C_r through expm1 (jsx_mixed.h): y = (s-2)(ln gmin + ln(g x)/2), x = 2^l2 */
        const double e = s - 2., lg = log(j.gamma_min) + 0.5 * log(g);
        if (e != 0.0) {
            c.kC_a = (float)(-(2. / e) * pow(j.gamma_min, -e));
            c.kC_lin = 0.f;
            c.kC_y0 = e * lg;
            c.kC_y1 = 0.5 * e * 0.69314718055994530942;
        } else {
            c.kC_a = -2.f;
            c.kC_lin = 1.f;
            c.kC_y0 = lg;
            c.kC_y1 = 0.5 * 0.69314718055994530942;
        }
    }
    c.e_ki = (float)m.e_ki; c.e_ei = (float)m.e_ei; c.e_kv = (float)m.e_kv;
    c.n0 = (float)j.n0; c.b0 = (float)j.b0; c.n_negexp = (float)m.n_negexp; c.b_negexp = (float)m.b_negexp;
    double gam = j.gamma0 > 1.0 ? j.gamma0 : 1.0;
    double beta = sqrt(1. - 1. / (gam * gam));
    c.gam_flat = (float)gam; c.beta_flat = (float)beta; c.omb_flat = (float)(1. / (gam * gam * (1. + beta)));
    double nz = -D.dir[2];
    /* 1 - cos(theta) without cancellation: |n - z|^2 / 2 */
    c.one_minus_mu = (float)(0.5 * (D.dir[0] * D.dir[0] + D.dir[1] * D.dir[1] + (nz - 1.0) * (nz - 1.0)));
    for (int i = 0; i < 3; ++i) { c.n_los[i] = (float)(-D.dir[i]); c.axis[i] = (float)m.axis[i]; }
    c.sin_pitch = (float)m.sin_pitch; c.cos_pitch = (float)m.cos_pitch;
    c.thermal_fraction = (float)j.thermal_fraction;
    c.inv_pc = 1.0 / PC;
    /* FP32 stage values put a floor under the usable relative tolerance: the
       embedded error estimate dt * sum(dc_j k_j) carries the stages' rounding,
       ~1e-7 |k| x |dc| ~ 8e-9 dt |k|, i.e. an error-norm noise of ~8e-9 / rel,
       harmless (< 0.1) down to rel = 1e-7.  (Q, U, V are controlled relative
       to the polarized intensity, see ml_step_error, which keeps their noise
       in the same range.) */
    const double ft = JSX_MIXED_TAU_FLOOR;
    const double fs = p.stokes == JSX_STOKES_FULL ? JSX_MIXED_STOKES_FLOOR_FULL : JSX_MIXED_STOKES_FLOOR_I;
    c.tau_abs = p.tau_abs_tol > ft ? p.tau_abs_tol : ft;
    c.tau_rel = p.tau_rel_tol > ft ? p.tau_rel_tol : ft;
    c.st_abs = p.stokes_abs_tol;
    c.st_rel = p.stokes_rel_tol > fs ? p.stokes_rel_tol : fs;
    c.st_rel_x = p.stokes_rel_tol > JSX_MIXED_STATE_FLOOR ? p.stokes_rel_tol : JSX_MIXED_STATE_FLOOR;
    /* the pieces' error control is floored at the FP32 tolerance floor: a
       request asking for a tighter Stokes tolerance (grid fields, whose
       kinks the error estimate undershoots, run at 1e-9) is integrated by
       the sequential lane, whose norm holds the state to the caller's
       tolerance */
    c.split_ok = p.stokes_rel_tol >= 0.999 * JSX_MIXED_STOKES_FLOOR_FULL ? 1 : 0;
    c.dt_max_pc = p.dt_max / PC;
    c.tau_hz = (p.flags & JSX_OPT_FULL_DEPTH) || !(JSX_MIXED_TAU_HORIZON < p.tau_max) ? 0.0 : JSX_MIXED_TAU_HORIZON;
    /* the terms' rounding floor matters where the state tolerance is tighter
       than the FP32 floor: there the rounding of emission and absorption,
       ~1e-8 dt k_I I per step, exceeds rel |I| in an optically thick region
       once dt k_I > 0.1 rel / 1e-8; at rel >= 1e-7 stability bounds dt k_I
       first (3.3), so those requests skip the 16 operations per RHS */
    c.mag_floor = c.st_rel_x < JSX_MIXED_STOKES_FLOOR_I ? 1 : 0;
    /* fields interpolated on a regular grid are linear inside each triangle
       and kinked on its edges, where the embedded error estimate does not
       see the error; the FP32 floor on the step's increment (1e-7, above a
       tight caller's tolerance) then let a step across several cells through
       with its whole kink error -- 3e-4 of a thin ray's intensity from one
       first step of 0.08 pc through cells of 0.004 pc.  A request tighter than
       the floor therefore steps at most JSX_MIXED_GRID_STEP cells of the
       finest grid spacing: 3330 pixels of a 256^2 image at rel 1e-9 against
       FP64 at 1e-11, 512 x 128 grid: max 4.7e-6 (no cap: 3.1e-4); 2048 x 512:
       1.5e-6 (no cap: 4.5e-5) for 16% more evaluations; 8 cells: 1.0e-5 */
    c.st_dt_max_pc = 0.0;
    if (c.mag_floor) {
        double cell = 0.0;
        const DGrid *gs[4] = {m.has_gb ? &m.gb : nullptr, m.has_gb2 ? &m.gb2 : nullptr, m.has_gn ? &m.gn : nullptr,
                              m.has_gv ? &m.gv : nullptr};
        for (const DGrid *g : gs) {
            if (!g || g->nz == 0) continue;
            const double d = g->dz < g->dr ? g->dz : g->dr;
            cell = (cell == 0.0 || d < cell) ? d : cell;
        }
        if (cell > 0.0) c.st_dt_max_pc = JSX_MIXED_GRID_STEP * cell;
    }
}

/* ---------------- persistent batch launch ---------------- */
struct BatchItem {
    DModel m;
    DImage im;
    DSolver S;
    MConsts c;
    DOut out;
    int64_t npix;
};

/* pieces the planner aims at.  Dormand-Prince pieces: about two per piece
   worker the piece kernel keeps resident (5-lane groups at two wavefronts per
   SIMD, or single lanes at one wavefront per SIMD).  Exponential pieces: one
   per resident lane in the first round ("piece_target" overrides), the
   budgeted rounds cut the costly ones further. */
static void split_target(jsx_ctx *ctx) {
    SplitWS &w = ctx->split;
    long t = ctx->piece_layout != 0 ? (long)ctx->n_cu * 4 * 64 * 2
                                    : (long)ctx->n_cu * ctx->persist_blocks_per_cu * 4 * (64 / JSX_SPLIT_GROUP) * 2;
    long tm = (long)ctx->n_cu * 256;
    if (ctx->piece_target > 0) tm = ctx->piece_target;
    w.target_pieces = (int)(t < w.piece_cap / 2 ? t : w.piece_cap / 2);
    w.target_mg = (int)(tm < w.piece_cap / 2 ? tm : w.piece_cap / 2);
    w.layout = ctx->piece_layout;
}

/* one allocation for the split path's tables (kept for the context's life) */
#define JSX_SPLIT_RAY_CAP 65536
#define JSX_SPLIT_PIECE_CAP (1 << 20)

/* bytes of the split workspace (allocated once per context, on the first
 * mixed full-Stokes batched launch) */
static size_t split_workspace_bytes() {
    const size_t cap = JSX_SPLIT_RAY_CAP, piece_cap = JSX_SPLIT_PIECE_CAP;
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return r(128 * sizeof(unsigned) + 3 * sizeof(unsigned long long)) + r(sizeof(LongRay) * cap) +
           r(sizeof(RaySum) * cap) + r(sizeof(float) * cap * JSX_SPLIT_NSAMP) + r(sizeof(Piece) * piece_cap) +
           2 * r(sizeof(PieceOut) * piece_cap) + 3 * r(sizeof(int) * piece_cap);
}

static int split_workspace(jsx_ctx *ctx, hipStream_t st) {
    if (ctx->split_mem) return JSX_OK;
    const int cap = JSX_SPLIT_RAY_CAP, piece_cap = JSX_SPLIT_PIECE_CAP;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    size_t o_cnt = take(128 * sizeof(unsigned) + 3 * sizeof(unsigned long long));
    size_t o_rays = take(sizeof(LongRay) * cap);
    size_t o_rsum = take(sizeof(RaySum) * cap);
    size_t o_prof = take(sizeof(float) * (size_t)cap * JSX_SPLIT_NSAMP);
    size_t o_pcs = take(sizeof(Piece) * piece_cap);
    size_t o_pout = take(sizeof(PieceOut) * piece_cap);
    size_t o_pout2 = take(sizeof(PieceOut) * piece_cap);
    size_t o_na = take(sizeof(int) * piece_cap);
    size_t o_nb = take(sizeof(int) * piece_cap);
    size_t o_ready = take(sizeof(unsigned) * piece_cap);
    HIPCHECK(ctx, hipMalloc(&ctx->split_mem, off));
    /* the launch epoch and the slot marks start at zero (the first launch's
       reset makes the epoch 1).  Cleared on the launch's own stream and waited
       for: a launch on a non-blocking stream is not ordered after the null
       stream, and a clear that landed while the split kernels ran wiped their
       counters.  (The first split launch of a context is never captured: a
       capture needs an earlier launch of the same shape.) */
    HIPCHECK(ctx, hipMemsetAsync(ctx->split_mem, 0, off, st));
    HIPCHECK(ctx, hipStreamSynchronize(st));
    char *b = (char *)ctx->split_mem;
    SplitWS &w = ctx->split;
    /* the three counters the exponential launch polls sit on lines of their own */
    w.n_long = (unsigned *)(b + o_cnt);
    w.piece_head = w.n_long + 2;
    w.round_end = w.n_long + 3;
    w.dp_end = w.n_long + 4;
    w.n_mg = w.n_long + 6;
    w.mg_epoch = w.n_long + 8;
    w.n_pieces = w.n_long + 32;
    w.mg_head = w.n_long + 64;
    w.mg_pending = w.n_long + 96;
    w.mg_ready = (unsigned *)(b + o_ready);
    w.rhs_pieces = (unsigned long long *)(b + o_cnt + 128 * sizeof(unsigned));
    w.rhs_mg = w.rhs_pieces + 1;
    w.mg_stat = w.rhs_pieces + 2;
    w.rays = (LongRay *)(b + o_rays);
    w.rsum = (RaySum *)(b + o_rsum);
    w.cap = cap;
    w.prof = (float *)(b + o_prof);
    w.pieces = (Piece *)(b + o_pcs);
    w.pout = (PieceOut *)(b + o_pout);
    w.pout2 = (PieceOut *)(b + o_pout2);
    w.next_a = (int *)(b + o_na);
    w.next_b = (int *)(b + o_nb);
    w.piece_cap = piece_cap;
    split_target(ctx);
    w.w_min = ctx->split_w_min;
    w.floor_max = ctx->split_floor_max;
    w.mg_cost = ctx->mg_cost;
    w.mg_profile = ctx->mg_profile;
    w.mg_rel = ctx->mg_rel;
    w.mg_err = ctx->mg_err;
    return JSX_OK;
}

/* Upload per-request constants and launch one persistent kernel over all
 * rays.  All items share precision and Stokes mode.  Synchronises the stream
 * before reusing the context's batch buffers. */
/* The plain form: one thread per pixel, one pass per image.  Models with a
 * field or geometry kind outside the workloads' set run in the kernels built
 * with that code (the common kernels leave it out: it costs them registers). */
template <int NS, bool EXT>
static void plain_kernel(hipStream_t st, int precision, const void *dp, const DOut &O, int64_t npix) {
    const int threads = 256;
    const dim3 grid((unsigned)((npix + threads - 1) / threads));
    if (precision == JSX_PREC_MIXED)
        hipLaunchKernelGGL((jsx_km_plain<NS, EXT>), grid, dim3(threads), 0, st, (const DParamsM *)dp, O, npix);
    else
        hipLaunchKernelGGL((jsx_k_plain<NS, EXT>), grid, dim3(threads), 0, st, (const DParams *)dp, O, npix);
}

/* a mixed request whose B (random-direction magnitude), N and Gamma come
   from regular grids of one geometry: integrated cell by cell (jsx_walk.h) */
static bool walk_eligible(const jsx_ctx *ctx, int precision, const DModel &m) {
    if (!ctx->grid_walk || precision != JSX_PREC_MIXED || m.ext) return false;
    if (m.bfield != JSX_B_GRID_SCALAR || m.nfield != JSX_N_GRID || m.vfield != JSX_V_GRID) return false;
    if (!m.has_gb || !m.has_gn || !m.has_gv) return false;
    const DGrid *g[3] = {&m.gb, &m.gn, &m.gv};
    for (const DGrid *x : g)
        if (x->nz < 2 || x->nr < 2 || x->nz != m.gb.nz || x->nr != m.gb.nr || x->z0 != m.gb.z0 || x->dz != m.gb.dz ||
            x->r0 != m.gb.r0 || x->dr != m.gb.dr)
            return false;
    return true;
}

/* every walk-eligible request of a launch in one walk launch (rays of all
   of them in one pool, so a request's rays fill the device together with
   the others'); the tables are reused once the last walk launch that read
   them has finished */
static int launch_walk(jsx_ctx *ctx, hipStream_t st, bool full, const std::vector<const BatchItem *> &items) {
    const size_t n = items.size();
    if (n == 0) return JSX_OK;
    if (!ctx->d_walk_ctr) HIPCHECK(ctx, hipMalloc(&ctx->d_walk_ctr, sizeof(unsigned long long)));
    if (!ctx->walk_ev) HIPCHECK(ctx, hipEventCreateWithFlags(&ctx->walk_ev, hipEventDisableTiming));
    HIPCHECK(ctx, hipEventSynchronize(ctx->walk_ev));
    if (n > ctx->walk_cap) {
        (void)hipFree(ctx->d_walk_reqs); (void)hipFree(ctx->d_walk_off); (void)hipFree(ctx->d_walk_outs);
        if (ctx->h_walk) (void)hipHostFree(ctx->h_walk);
        ctx->d_walk_reqs = nullptr; ctx->d_walk_off = nullptr; ctx->d_walk_outs = nullptr; ctx->h_walk = nullptr;
        ctx->walk_cap = 0;
        const size_t cap = n < 64 ? 64 : n;
        HIPCHECK(ctx, hipMalloc(&ctx->d_walk_reqs, sizeof(DParamsM) * cap));
        HIPCHECK(ctx, hipMalloc(&ctx->d_walk_off, sizeof(int64_t) * (cap + 1)));
        HIPCHECK(ctx, hipMalloc(&ctx->d_walk_outs, sizeof(DOut) * cap));
        HIPCHECK(ctx, hipHostMalloc(&ctx->h_walk, (sizeof(DParamsM) + sizeof(int64_t) + sizeof(DOut)) * cap + 64,
                                    hipHostMallocDefault));
        ctx->walk_cap = cap;
    }
    DParamsM *hr = (DParamsM *)ctx->h_walk;
    int64_t *ho = (int64_t *)(hr + ctx->walk_cap);
    DOut *hd = (DOut *)(ho + ctx->walk_cap + 1);
    int64_t total = 0;
    for (size_t k = 0; k < n; ++k) {
        const BatchItem &it = *items[k];
        hr[k].m = it.m; hr[k].im = it.im; hr[k].S = it.S; hr[k].c = it.c;
        ho[k] = total;
        hd[k] = it.out;
        total += it.npix;
    }
    ho[n] = total;
    if (total == 0) return JSX_OK;
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d_walk_reqs, hr, sizeof(DParamsM) * n, hipMemcpyHostToDevice, st));
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d_walk_off, ho, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d_walk_outs, hd, sizeof(DOut) * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(jsx_k_walk_reset, dim3(1), dim3(64), 0, st, ctx->d_walk_ctr);
    const dim3 grid(ctx->n_cu * 8), block(256);
    if (full)
        hipLaunchKernelGGL(jsx_k_walk<4>, grid, block, 0, st, ctx->d_walk_reqs, ctx->d_walk_off, ctx->d_walk_outs, (int)n,
                           total, ctx->d_walk_ctr);
    else
        hipLaunchKernelGGL(jsx_k_walk<1>, grid, block, 0, st, ctx->d_walk_reqs, ctx->d_walk_off, ctx->d_walk_outs, (int)n,
                           total, ctx->d_walk_ctr);
    HIPCHECK(ctx, hipGetLastError());
    HIPCHECK(ctx, hipEventRecord(ctx->walk_ev, st));
    return JSX_OK;
}

static int launch_plain(jsx_ctx *ctx, hipStream_t st, int precision, bool full, const DModel &m, const DImage &D,
                        const DSolver &S, const MConsts *c, const DOut &O, int64_t npix) {
    if (npix == 0) return JSX_OK;
    if (walk_eligible(ctx, precision, m)) {
        BatchItem it;
        it.m = m; it.im = D; it.S = S; it.c = *c; it.out = O; it.npix = npix;
        std::vector<const BatchItem *> one{&it};
        return launch_walk(ctx, st, full, one);
    }
    void *dp = nullptr;
    int slot = 0, rc;
    if (precision == JSX_PREC_MIXED) {
        DParamsM hm;
        hm.m = m; hm.im = D; hm.S = S; hm.c = *c;
        if ((rc = stage_bytes(ctx, &hm, sizeof hm, st, &dp, &slot))) return rc;
    } else {
        DParams *p = nullptr;
        if ((rc = stage_params(ctx, m, D, S, st, &p, &slot))) return rc;
        dp = p;
    }
    if (m.ext) {
        if (full) plain_kernel<4, true>(st, precision, dp, O, npix);
        else plain_kernel<1, true>(st, precision, dp, O, npix);
    } else {
        if (full) plain_kernel<4, false>(st, precision, dp, O, npix);
        else plain_kernel<1, false>(st, precision, dp, O, npix);
    }
    HIPCHECK(ctx, hipGetLastError());
    return stage_done(ctx, slot, st);
}

static int launch_batch(jsx_ctx *ctx, hipStream_t st, int precision, int stokes, const std::vector<BatchItem> &items,
                        int order) {
    if (!items.empty()) {
        /* models with the further reference kinds run through the plain kernels,
           on the same stream, before the batched launch of the others */
        bool any_ext = false;
        for (const BatchItem &it : items) any_ext = any_ext || it.m.ext || walk_eligible(ctx, precision, it.m);
        if (any_ext) {
            std::vector<BatchItem> common;
            std::vector<const BatchItem *> walk;
            for (const BatchItem &it : items) {
                if (walk_eligible(ctx, precision, it.m)) { walk.push_back(&it); continue; }
                if (!it.m.ext) { common.push_back(it); continue; }
                int rc = launch_plain(ctx, st, precision, stokes == JSX_STOKES_FULL, it.m, it.im, it.S, &it.c, it.out,
                                      it.npix);
                if (rc) return rc;
            }
            int rc = launch_walk(ctx, st, stokes == JSX_STOKES_FULL, walk);
            if (rc) return rc;
            return launch_batch(ctx, st, precision, stokes, common, order);
        }
    }
    size_t n = items.size();
    if (!n) return JSX_OK;
    size_t esz = precision == JSX_PREC_MIXED ? sizeof(DParamsM) : sizeof(DParams);
    int64_t total_pix = 0;
    for (size_t i = 0; i < n; ++i) total_pix += items[i].npix;
    if (total_pix == 0) return JSX_OK;
    /* Launches on one stream follow each other without a host wait; device
       buffers are only reallocated after the device has drained them. */
    const bool triage_on = ctx->triage == 1 || (ctx->triage == 2 && precision == JSX_PREC_MIXED);
    const bool grow = !ctx->d_queues || ctx->reqs_cap < n * esz || ctx->n_cap < n ||
                      (triage_on && ctx->live_cap < (size_t)total_pix);
    if (grow) {
        HIPCHECK(ctx, hipStreamSynchronize(st));
        HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->batch_stream) HIPCHECK(ctx, hipStreamSynchronize(ctx->batch_stream));
    } else if (ctx->batch_ev && ctx->batch_stream != st) {
        HIPCHECK(ctx, hipStreamWaitEvent(st, ctx->batch_ev, 0));
    }
    if (!ctx->d_queues) HIPCHECK(ctx, hipMalloc(&ctx->d_queues, sizeof(BatchQueues)));
    if (ctx->reqs_cap < n * esz) {
        (void)hipFree(ctx->d_reqs);
        ctx->reqs_cap = n * esz;
        HIPCHECK(ctx, hipMalloc(&ctx->d_reqs, ctx->reqs_cap));
    }
    if (ctx->n_cap < n) {
        (void)hipFree(ctx->d_off); (void)hipFree(ctx->d_outs);
        ctx->n_cap = n;
        HIPCHECK(ctx, hipMalloc(&ctx->d_off, sizeof(int64_t) * (n + 1)));
        HIPCHECK(ctx, hipMalloc(&ctx->d_outs, sizeof(DOut) * n));
    }
    /* request table, offsets and outputs in a pinned buffer (the one used two
       launches ago: its copy has long completed) */
    const size_t b_req = (n * esz + 255) & ~(size_t)255, b_off = (sizeof(int64_t) * (n + 1) + 255) & ~(size_t)255;
    const size_t b_all = b_req + b_off + sizeof(DOut) * n;
    const int sk = ctx->stage_next;
    ctx->stage_next ^= 1;
    if (!ctx->stage_ev[sk]) HIPCHECK(ctx, hipEventCreateWithFlags(&ctx->stage_ev[sk], hipEventDisableTiming));
    HIPCHECK(ctx, hipEventSynchronize(ctx->stage_ev[sk]));
    if (ctx->h_stage_cap[sk] < b_all) {
        (void)hipHostFree(ctx->h_stage[sk]);
        ctx->h_stage[sk] = nullptr;
        ctx->h_stage_cap[sk] = 0;
        HIPCHECK(ctx, hipHostMalloc((void **)&ctx->h_stage[sk], b_all, hipHostMallocDefault));
        ctx->h_stage_cap[sk] = b_all;
    }
    unsigned char *hreq = ctx->h_stage[sk];
    int64_t *off = (int64_t *)(hreq + b_req);
    DOut *outs = (DOut *)(hreq + b_req + b_off);
    off[0] = 0;
    for (size_t i = 0; i < n; ++i) {
        const BatchItem &it = items[i];
        if (precision == JSX_PREC_MIXED) {
            DParamsM q; q.m = it.m; q.im = it.im; q.S = it.S; q.c = it.c;
            std::memcpy(hreq + i * esz, &q, esz);
        } else {
            DParams q; q.m = it.m; q.im = it.im; q.S = it.S;
            std::memcpy(hreq + i * esz, &q, esz);
        }
        off[i + 1] = off[i] + it.npix;
        outs[i] = it.out;
    }
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d_reqs, hreq, n * esz, hipMemcpyHostToDevice, st));
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d_off, off, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d_outs, outs, sizeof(DOut) * n, hipMemcpyHostToDevice, st));
    HIPCHECK(ctx, hipEventRecord(ctx->stage_ev[sk], st));
    int blocks = ctx->n_cu * ctx->persist_blocks_per_cu;
    int64_t total = off[n];   /* == total_pix */
    int nreq = (int)n;
    const bool full = stokes == JSX_STOKES_FULL;
    SplitWS ws{};
    const bool split = precision == JSX_PREC_MIXED && full && ctx->defer_rhs > 0;
    jsx_ctx::EvSet *ev = nullptr;
    if (ctx->timing && ctx->ev_n < jsx_ctx::kEvMax) {
        if ((int)ctx->evs.size() <= ctx->ev_n) {
            jsx_ctx::EvSet es;
            for (int k = 0; k < 7; ++k) HIPCHECK(ctx, hipEventCreate(&es.e[k]));
            ctx->evs.push_back(es);
        }
        ev = &ctx->evs[ctx->ev_n++];
    }
    auto mark = [&](int k) { if (ev) (void)hipEventRecord(ev->e[k], st); };
    if (split) {
        int rc = split_workspace(ctx, st);
        if (rc) return rc;
        ws = ctx->split;
    }
    /* triage list: one int64 per ray of the launch */
    int64_t *live = nullptr, *sorted = nullptr;
    unsigned char *bucket = nullptr, *lbucket = nullptr;
    const bool triage = ctx->triage == 1 || (ctx->triage == 2 && precision == JSX_PREC_MIXED);
    if (triage) {
        if (ctx->live_cap < (size_t)total) {
            (void)hipFree(ctx->d_live);
            (void)hipFree(ctx->d_sorted);
            (void)hipFree(ctx->d_bucket);
            (void)hipFree(ctx->d_sbucket);
            ctx->d_live = ctx->d_sorted = nullptr;
            ctx->d_bucket = ctx->d_sbucket = nullptr;
            ctx->live_cap = 0;
            HIPCHECK(ctx, hipMalloc(&ctx->d_live, sizeof(int64_t) * (size_t)total));
            HIPCHECK(ctx, hipMalloc(&ctx->d_sorted, sizeof(int64_t) * (size_t)total));
            HIPCHECK(ctx, hipMalloc(&ctx->d_bucket, (size_t)total));
            HIPCHECK(ctx, hipMalloc(&ctx->d_sbucket, (size_t)total));
            ctx->live_cap = (size_t)total;
        }
        live = ctx->d_live;
        if (ctx->lpt) {
            sorted = ctx->d_sorted;
            bucket = ctx->d_bucket;
        }
    }
    const int64_t tri_need = (total + 255) / 256;
    const int tri_blocks = (int)(tri_need < (int64_t)ctx->n_cu * 16 ? tri_need : (int64_t)ctx->n_cu * 16);
    /* default tail threshold: one wavefront per SIMD of the chip (half a
     * wavefront per SIMD left the long rays of a many-image sweep running
     * sequentially for longer: 11% slower on 512 images, 2% on one) */
    const unsigned park_lanes = ctx->park_lanes ? ctx->park_lanes : (unsigned)ctx->n_cu * 4u * 64u;
    /* the launch sequence: it depends on the request table only through
       device buffers, so while the key below is unchanged it is replayed as
       one HIP graph captured at its first use (no memset nodes: the queues
       and the split counters are reset by kernels) */
    auto enqueue = [&]() {
    hipLaunchKernelGGL(jsx_k_reset_queues, dim3(1), dim3(64), 0, st, ctx->d_queues);
    if (split) hipLaunchKernelGGL(jsx_k_split_reset, dim3(1), dim3(64), 0, st, ws);
    mark(5);
    if (precision == JSX_PREC_MIXED) {
        const DParamsM *R = (const DParamsM *)ctx->d_reqs;
        if (live) {
            if (full)
                hipLaunchKernelGGL(jsx_k_triage<PolMixed<4>>, dim3(tri_blocks), dim3(256), 0, st, R, ctx->d_off,
                                   ctx->d_outs, nreq, total, order, ctx->d_queues, live, bucket, ctx->triage_walk);
            else
                hipLaunchKernelGGL(jsx_k_triage<PolMixed<1>>, dim3(tri_blocks), dim3(256), 0, st, R, ctx->d_off,
                                   ctx->d_outs, nreq, total, order, ctx->d_queues, live, bucket, ctx->triage_walk);
        }
        if (sorted) {
            hipLaunchKernelGGL(jsx_k_lpt_scatter, dim3(tri_blocks), dim3(256), 0, st, ctx->d_queues, live, bucket, sorted,
                               ctx->d_sbucket);
            live = sorted;
            lbucket = ctx->d_sbucket;
        }
        mark(0);
        if (full)
            hipLaunchKernelGGL(jsx_k_batch<PolMixed<4>>, dim3(blocks), dim3(256), 0, st, R, ctx->d_off, ctx->d_outs, nreq,
                               total, ctx->d_queues, ctx->queue_block, order, ctx->steal, ws, ctx->defer_rhs, live,
                               park_lanes, ctx->park_wave, lbucket, ctx->park_bucket);
        else
            hipLaunchKernelGGL(jsx_k_batch<PolMixed<1>>, dim3(blocks), dim3(256), 0, st, R, ctx->d_off, ctx->d_outs, nreq,
                               total, ctx->d_queues, ctx->queue_block, order, ctx->steal, ws, 0u, live, park_lanes,
                               ctx->park_wave, lbucket, 0);
        mark(1);
        if (split) {
            const int cap = ws.cap;
            hipLaunchKernelGGL(jsx_k_split_sample, dim3((cap * 64 + 255) / 256), dim3(256), 0, st, R, ws);
            hipLaunchKernelGGL(jsx_k_split_alloc, dim3(1), dim3(1024), 0, st, ws);
            hipLaunchKernelGGL(jsx_k_split_cut, dim3(cap), dim3(256), 0, st, ws);
            mark(2);
            if (ctx->piece_layout == 0)
                hipLaunchKernelGGL(jsx_k_split_piece, dim3(blocks), dim3(256), 0, st, R, ws);
            else if (ctx->piece_layout != 2)
                hipLaunchKernelGGL(jsx_k_split_piece20, dim3(ctx->n_cu * 4), dim3(256), 0, st, R, ws);
            mark(6);
            if (ctx->piece_layout >= 2) {
                if (ctx->mg_part) {
                    /* one launch, two workgroups per CU (its residency) */
                    hipLaunchKernelGGL(jsx_k_split_magnus, dim3(ctx->n_cu * 2), dim3(256), 0, st, R, ws, 0u,
                                       ctx->mg_part);
                } else {
                    const int rounds = ctx->piece_budget ? 4 : 1;
                    for (int r = 0; r < rounds; ++r) {
                        if (r) hipLaunchKernelGGL(jsx_k_split_round, dim3(1), dim3(64), 0, st, ws);
                        hipLaunchKernelGGL(jsx_k_split_magnus, dim3(ctx->n_cu * 8), dim3(256), 0, st, R, ws,
                                           r + 1 < rounds ? ctx->piece_budget : 0u, 0u);
                    }
                }
            }
            mark(3);
            /* four chain-jumping rounds when the exponential pieces split into chains */
            const int jump = ctx->piece_layout >= 2 && (ctx->piece_budget || ctx->mg_part) ? 4 : 0;
            for (int r = 0; r < jump; ++r)
                hipLaunchKernelGGL(jsx_k_split_jump, dim3(ctx->n_cu * 8), dim3(256), 0, st, ws, r);
            hipLaunchKernelGGL(jsx_k_split_combine, dim3((cap + 255) / 256), dim3(256), 0, st, ctx->d_outs, ws,
                               jump ? 1 : 0);
        } else {
            mark(2);
            mark(6);
            mark(3);
        }
        mark(4);
    } else {
        const DParams *R = (const DParams *)ctx->d_reqs;
        if (live) {
            if (full)
                hipLaunchKernelGGL(jsx_k_triage<PolFP64<4>>, dim3(tri_blocks), dim3(256), 0, st, R, ctx->d_off,
                                   ctx->d_outs, nreq, total, order, ctx->d_queues, live, bucket, ctx->triage_walk);
            else
                hipLaunchKernelGGL(jsx_k_triage<PolFP64<1>>, dim3(tri_blocks), dim3(256), 0, st, R, ctx->d_off,
                                   ctx->d_outs, nreq, total, order, ctx->d_queues, live, bucket, ctx->triage_walk);
        }
        if (sorted) {
            hipLaunchKernelGGL(jsx_k_lpt_scatter, dim3(tri_blocks), dim3(256), 0, st, ctx->d_queues, live, bucket, sorted,
                               ctx->d_sbucket);
            live = sorted;
            lbucket = ctx->d_sbucket;
        }
        mark(0);
        if (full)
            hipLaunchKernelGGL(jsx_k_batch<PolFP64<4>>, dim3(blocks), dim3(256), 0, st, R, ctx->d_off, ctx->d_outs, nreq,
                               total, ctx->d_queues, ctx->queue_block, order, ctx->steal, ws, 0u, live, park_lanes,
                               ctx->park_wave, lbucket, 0);
        else
            hipLaunchKernelGGL(jsx_k_batch<PolFP64<1>>, dim3(blocks), dim3(256), 0, st, R, ctx->d_off, ctx->d_outs, nreq,
                               total, ctx->d_queues, ctx->queue_block, order, ctx->steal, ws, 0u, live, park_lanes,
                               ctx->park_wave, lbucket, 0);
        mark(1);
        mark(2);
        mark(6);
        mark(3);
        mark(4);
    }
    };
    BatchGraphKey key;
    std::memset(&key, 0, sizeof key);
    key.precision = precision; key.stokes = stokes; key.order = order; key.nreq = nreq; key.total = total;
    key.blocks = blocks; key.tri_blocks = tri_blocks; key.park_lanes = park_lanes;
    key.reqs = ctx->d_reqs; key.off = ctx->d_off; key.outs = ctx->d_outs; key.queues = ctx->d_queues;
    key.live = live; key.sorted = sorted; key.bucket = bucket; key.sbucket = ctx->d_sbucket;
    key.queue_block = ctx->queue_block; key.steal = ctx->steal; key.defer_rhs = ctx->defer_rhs;
    key.park_wave = ctx->park_wave; key.park_bucket = ctx->park_bucket; key.piece_layout = ctx->piece_layout;
    key.piece_budget = ctx->piece_budget; key.mg_part = ctx->mg_part; key.n_cu = ctx->n_cu; key.split = split;
    key.triage_walk = ctx->triage_walk;
    if (split) key.ws = ws;
    hipStreamCaptureStatus cap_st = hipStreamCaptureStatusNone;
    const bool use_graph = ctx->graph && !ev && hipStreamIsCapturing(st, &cap_st) == hipSuccess &&
                           cap_st == hipStreamCaptureStatusNone;
    /* a shape seen for the first time runs as it is (a server's launches
       change shape all the time); the second launch of a shape in a row is
       captured, later ones replay the graph */
    const bool same = ctx->bg_seen && std::memcmp(&key, &ctx->bg_key, sizeof key) == 0;
    if (!same && ctx->bg_exec) {
        /* the last replay may still run: wait for it before destroying */
        if (ctx->batch_ev) HIPCHECK(ctx, hipEventSynchronize(ctx->batch_ev));
        (void)hipGraphExecDestroy(ctx->bg_exec);
        ctx->bg_exec = nullptr;
    }
    if (use_graph) {
        ctx->bg_key = key;
        ctx->bg_seen = true;
    }
    if (!use_graph || !same) {
        enqueue();
    } else {
        if (!ctx->bg_exec) {
            hipGraph_t g = nullptr;
            HIPCHECK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            enqueue();
            hipError_t e = hipStreamEndCapture(st, &g);
            if (e == hipSuccess) e = hipGraphInstantiate(&ctx->bg_exec, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ctx->bg_exec = nullptr;
                return fail(ctx, JSX_EHIP, "batched launch graph: %s", hipGetErrorString(e));
            }
            ctx->graph_captures++;
        }
        HIPCHECK(ctx, hipGraphLaunch(ctx->bg_exec, st));
        ctx->graph_launches++;
    }
    HIPCHECK(ctx, hipGetLastError());
    if (!ctx->batch_ev) HIPCHECK(ctx, hipEventCreateWithFlags(&ctx->batch_ev, hipEventDisableTiming));
    HIPCHECK(ctx, hipEventRecord(ctx->batch_ev, st));
    ctx->batch_stream = st;
    return JSX_OK;
}

/* ---------------- observe ---------------- */
static int observe_locked(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_image_desc *image,
                          const jsx_solver_params *params, const jsx_outputs *out, hipStream_t st) {
    if (!jet || !image || !params || !out) return fail(ctx, JSX_EINVAL, "null argument");
    if (!out->I) return fail(ctx, JSX_EINVAL, "outputs.I must be set");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DModel m;
    DSolver S;
    DImage D;
    ImageRec *rec = nullptr;
    int rc;
    if ((rc = lower_model(ctx, *jet, params->flags, m))) return rc;
    if ((rc = lower_solver(ctx, *params, S))) return rc;
    if ((rc = lower_image(ctx, *image, D, &rec))) return rc;
    DOut O = lower_out(*out);
    int64_t npix = image->n_across * image->n_along;
    if (npix == 0) return JSX_OK;
    const bool full = params->stokes == JSX_STOKES_FULL;
    if (params->precision == JSX_PREC_MIXED) {
        DParamsM hm;
        hm.m = m; hm.im = D; hm.S = S;
        lower_mixed(*jet, *params, m, D, hm.c);
        if (params->flags & JSX_OPT_PERSISTENT) {
            std::vector<BatchItem> one(1);
            one[0].m = m; one[0].im = D; one[0].S = S; one[0].c = hm.c; one[0].out = O; one[0].npix = npix;
            return launch_batch(ctx, st, JSX_PREC_MIXED, params->stokes, one, (params->flags & 0x200u) ? 1 : 0);
        }
        return launch_plain(ctx, st, JSX_PREC_MIXED, full, m, D, S, &hm.c, O, npix);
    }
    if (params->flags & JSX_OPT_PERSISTENT) {
        std::vector<BatchItem> one(1);
        std::memset(&one[0].c, 0, sizeof(MConsts));
        one[0].m = m; one[0].im = D; one[0].S = S; one[0].out = O; one[0].npix = npix;
        return launch_batch(ctx, st, JSX_PREC_FP64, params->stokes, one, (params->flags & 0x200u) ? 1 : 0);
    }
    return launch_plain(ctx, st, JSX_PREC_FP64, full, m, D, S, nullptr, O, npix);
}

extern "C" int jsx_observe(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_image_desc *image,
                           const jsx_solver_params *params, const jsx_outputs *out_device, void *stream) {
    if (!ctx) return JSX_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return observe_locked(ctx, jet, image, params, out_device, st);
}

extern "C" int jsx_alloc_outputs(jsx_ctx *ctx, int64_t n, int32_t stokes, jsx_outputs *o) {
    if (!ctx || !o || n <= 0) return fail(ctx, JSX_EINVAL, "bad arguments");
    std::memset(o, 0, sizeof *o);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    size_t nd = (size_t)n * sizeof(double), nu = (size_t)n * sizeof(uint32_t);
    if (hipMalloc(&o->tau, nd) != hipSuccess || hipMalloc(&o->I, nd) != hipSuccess || hipMalloc(&o->l, nd) != hipSuccess ||
        hipMalloc(&o->flags, nu) != hipSuccess || hipMalloc(&o->n_rhs, nu) != hipSuccess ||
        hipMalloc(&o->n_steps, nu) != hipSuccess) {
        jsx_free_outputs(ctx, o);
        return fail(ctx, JSX_ENOMEM, "output allocation of %lld pixels failed", (long long)n);
    }
    if (stokes == JSX_STOKES_FULL) {
        if (hipMalloc(&o->Q, nd) != hipSuccess || hipMalloc(&o->U, nd) != hipSuccess || hipMalloc(&o->V, nd) != hipSuccess) {
            jsx_free_outputs(ctx, o);
            return fail(ctx, JSX_ENOMEM, "output allocation failed");
        }
    }
    return JSX_OK;
}

extern "C" int jsx_free_outputs(jsx_ctx *ctx, jsx_outputs *o) {
    if (!ctx || !o) return JSX_EINVAL;
    (void)hipSetDevice(ctx->device);
    void *ps[] = {o->tau, o->I, o->Q, o->U, o->V, o->l, o->flags, o->n_rhs, o->n_steps};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    std::memset(o, 0, sizeof *o);
    return JSX_OK;
}

extern "C" int jsx_observe_host(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_image_desc *image,
                                const jsx_solver_params *params, const jsx_outputs *out_host) {
    if (!ctx || !image || !params || !out_host) return JSX_EINVAL;
    int64_t n = image->n_across * image->n_along;
    if (n <= 0 || (image->n_across % 2)) return fail(ctx, JSX_EINVAL, "image needs n_along > 0 and even n_across > 0");
    jsx_outputs d;
    int rc = jsx_alloc_outputs(ctx, n, JSX_STOKES_FULL, &d);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rc = observe_locked(ctx, jet, image, params, &d, ctx->stream);
        if (rc == JSX_OK) {
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = fail(ctx, JSX_EHIP, "kernel failed: %s", hipGetErrorString(e));
        }
    }
    if (rc == JSX_OK) {
        size_t nd = (size_t)n * sizeof(double), nu = (size_t)n * sizeof(uint32_t);
        struct { void *h; void *dv; size_t b; } cp[] = {
            {out_host->tau, d.tau, nd}, {out_host->I, d.I, nd}, {out_host->Q, d.Q, nd}, {out_host->U, d.U, nd},
            {out_host->V, d.V, nd}, {out_host->l, d.l, nd}, {out_host->flags, d.flags, nu},
            {out_host->n_rhs, d.n_rhs, nu}, {out_host->n_steps, d.n_steps, nu}};
        for (auto &c : cp)
            if (c.h && hipMemcpy(c.h, c.dv, c.b, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = fail(ctx, JSX_EHIP, "copy back failed");
                break;
            }
    }
    jsx_free_outputs(ctx, &d);
    return rc;
}

extern "C" int jsx_hit_host(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_solver_params *params, const double *rays,
                            int64_t n, int32_t *n_seg, double *segments) {
    if (!ctx || !jet || !rays || !n_seg || !segments || n <= 0) return fail(ctx, JSX_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DModel m;
    int rc = lower_model(ctx, *jet, params ? params->flags : 0, m);
    if (rc) return rc;
    double *d_rays = nullptr, *d_segs = nullptr;
    int32_t *d_ns = nullptr;
    HIPCHECK(ctx, hipMalloc(&d_rays, sizeof(double) * 6 * n));
    HIPCHECK(ctx, hipMalloc(&d_segs, sizeof(double) * 12 * n));
    HIPCHECK(ctx, hipMalloc(&d_ns, sizeof(int32_t) * n));
    HIPCHECK(ctx, hipMemcpy(d_rays, rays, sizeof(double) * 6 * n, hipMemcpyHostToDevice));
    int threads = 256;
    hipLaunchKernelGGL(jsx_k_hit, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0, ctx->stream, m,
                       d_rays, n, d_ns, d_segs);
    HIPCHECK(ctx, hipGetLastError());
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHECK(ctx, hipMemcpy(n_seg, d_ns, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    HIPCHECK(ctx, hipMemcpy(segments, d_segs, sizeof(double) * 12 * n, hipMemcpyDeviceToHost));
    (void)hipFree(d_rays); (void)hipFree(d_segs); (void)hipFree(d_ns);
    return JSX_OK;
}

extern "C" int jsx_fields_host(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_solver_params *params,
                               const double *points, const double *n_los, int64_t n, double nu, double *out) {
    if (!ctx || !jet || !points || !n_los || !out || n <= 0) return fail(ctx, JSX_EINVAL, "bad arguments");
    if (!(nu > 0)) return fail(ctx, JSX_EINVAL, "nu must be > 0");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DModel m;
    int rc = lower_model(ctx, *jet, params ? params->flags : 0, m);
    if (rc) return rc;
    double *d_in = nullptr, *d_out = nullptr;
    const size_t b_in = sizeof(double) * 6 * (size_t)n, b_out = sizeof(double) * JSX_FIELDS_STRIDE * (size_t)n;
    hipError_t e = hipMalloc(&d_in, b_in);
    if (e == hipSuccess) e = hipMalloc(&d_out, b_out);
    if (e == hipSuccess) e = hipMemcpy(d_in, points, b_in / 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_in + 3 * n, n_los, b_in / 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(jsx_k_fields, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, m, d_in,
                           d_in + 3 * n, n, nu, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, b_out, hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(ctx, JSX_EHIP, "jsx_fields_host: %s", hipGetErrorString(e));
    return JSX_OK;
}

/* ---------------- tau stripe ---------------- */
extern "C" int jsx_tau_stripe(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_image_desc *image,
                              const jsx_solver_params *params, double *tau_host) {
    if (!ctx || !jet || !image || !params || !tau_host) return fail(ctx, JSX_EINVAL, "null argument");
    if (image->n_across <= 0 || image->n_along <= 0 || (image->n_across % 2))
        return fail(ctx, JSX_EINVAL, "image needs n_along > 0 and even n_across > 0");
    if (params->n <= 0) return fail(ctx, JSX_EINVAL, "n must be positive");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DModel m;
    DSolver S;
    DImage D;
    int rc;
    if ((rc = lower_model(ctx, *jet, params->flags, m))) return rc;
    if ((rc = lower_solver(ctx, *params, S))) return rc;
    if ((rc = lower_image(ctx, *image, D, nullptr))) return rc;
    const int64_t n_out = image->n_along - image->n_along / 2;
    DParams *dp = nullptr;
    int slot = 0;
    if ((rc = stage_params(ctx, m, D, S, ctx->stream, &dp, &slot))) return rc;
    double *d_tau = nullptr;
    HIPCHECK(ctx, hipMalloc(&d_tau, sizeof(double) * (size_t)n_out));
    if (m.ext)
        hipLaunchKernelGGL(jsx_k_tau_stripe<true>, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, ctx->stream, dp,
                           d_tau, n_out);
    else
        hipLaunchKernelGGL(jsx_k_tau_stripe<false>, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, ctx->stream,
                           dp, d_tau, n_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && (rc = stage_done(ctx, slot, ctx->stream))) {
        (void)hipFree(d_tau);
        return rc;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(tau_host, d_tau, sizeof(double) * (size_t)n_out, hipMemcpyDeviceToHost);
    (void)hipFree(d_tau);
    if (e != hipSuccess) return fail(ctx, JSX_EHIP, "tau stripe failed: %s", hipGetErrorString(e));
    return JSX_OK;
}

/* ---------------- visibilities (jsx_fourier.h) ----------------
 * Launch shape: JSX_VIS_BLOCK x JSX_VIS_R visibilities per block; the pixel
 * list is cut into chunks of at least 512 pixels until the grid has about
 * eight blocks per CU, and chunk partial sums are added by jsx_k_vis_reduce
 * in chunk order. */
static int vis_check(jsx_ctx *ctx, const double *x, const double *y, const double *c_re, int64_t n_src,
                     const double *u, const double *v, int64_t n_vis, int32_t precision, const double *vis) {
    if (!ctx) return JSX_EINVAL;
    if (n_src < 0 || n_vis < 0) return fail(ctx, JSX_EINVAL, "negative pixel or visibility count");
    if (precision != JSX_PREC_FP64 && precision != JSX_PREC_MIXED)
        return fail(ctx, JSX_EINVAL, "precision must be JSX_PREC_FP64 or JSX_PREC_MIXED");
    if (n_src > 0 && (!x || !y || !c_re)) return fail(ctx, JSX_EINVAL, "null pixel buffer");
    if (n_vis > 0 && (!u || !v || !vis)) return fail(ctx, JSX_EINVAL, "null visibility buffer");
    if ((n_vis + (int64_t)JSX_VIS_BLOCK * JSX_VIS_R - 1) / ((int64_t)JSX_VIS_BLOCK * JSX_VIS_R) > 0x7fffffffLL)
        return fail(ctx, JSX_EINVAL, "too many visibilities for one launch");
    return JSX_OK;
}

/* fixed: NULL -- partial sums in the context's buffer (grown with hipMalloc,
 * ordered across streams by an event); else the caller's fixed buffer of
 * fixed_cap entries, used as is (no allocation, no event) */
static int vis_launch(jsx_ctx *ctx, const double *x, const double *y, const double *cre, const double *cim,
                      int64_t n_src, const double *u, const double *v, int64_t n_vis, int32_t isign,
                      int32_t precision, double *vis, hipStream_t st, double2 *fixed = nullptr,
                      size_t fixed_cap = 0) {
    if (n_vis == 0) return JSX_OK;
    double2 *out = (double2 *)vis;
    if (n_src == 0) {
        hipLaunchKernelGGL(jsx_k_vis_zero, dim3((unsigned)((n_vis + 255) / 256)), dim3(256), 0, st, out, n_vis);
        HIPCHECK(ctx, hipGetLastError());
        return JSX_OK;
    }
    const int64_t per_block = (int64_t)JSX_VIS_BLOCK * JSX_VIS_R;
    const int64_t n_tb = (n_vis + per_block - 1) / per_block;
    const int64_t want = 8LL * ctx->n_cu;
    int64_t n_chunks = (want + n_tb - 1) / n_tb;
    const int64_t max_chunks = (n_src + 511) / 512;
    if (n_chunks > max_chunks) n_chunks = max_chunks;
    if (n_chunks > 4096) n_chunks = 4096;
    if (n_chunks < 1) n_chunks = 1;
    const int64_t chunk = (n_src + n_chunks - 1) / n_chunks;
    n_chunks = (n_src + chunk - 1) / chunk;
    double2 *dst = out;
    if (n_chunks > 1 && fixed) {
        const size_t need = (size_t)n_chunks * (size_t)n_vis;
        if (need > fixed_cap)
            return fail(ctx, JSX_EINVAL,
                        "%lld visibilities of %lld pixels need %zu partial sums; the no-allocation entry point "
                        "holds at most %zu (use jsx_visibilities_device)",
                        (long long)n_vis, (long long)n_src, need, fixed_cap);
        dst = fixed;
    } else if (n_chunks > 1) {
        if (!ctx->vis_ev) HIPCHECK(ctx, hipEventCreateWithFlags(&ctx->vis_ev, hipEventDisableTiming));
        const size_t need = (size_t)n_chunks * (size_t)n_vis;
        if (need > ctx->vis_part_cap) {
            HIPCHECK(ctx, hipEventSynchronize(ctx->vis_ev));
            (void)hipFree(ctx->d_vis_part);
            ctx->d_vis_part = nullptr;
            ctx->vis_part_cap = 0;
            if (hipMalloc(&ctx->d_vis_part, need * sizeof(double2)) != hipSuccess)
                return fail(ctx, JSX_ENOMEM, "visibility partial sums: %zu bytes", need * sizeof(double2));
            ctx->vis_part_cap = need;
        }
        /* the partial-sum buffer is shared by every launch of this context */
        HIPCHECK(ctx, hipStreamWaitEvent(st, ctx->vis_ev, 0));
        dst = ctx->d_vis_part;
    }
    const dim3 grid((unsigned)n_tb, (unsigned)n_chunks);
    const double sgn = isign >= 0 ? 1.0 : -1.0;
    const bool mixed = precision == JSX_PREC_MIXED;
    const bool hw = mixed && ctx->vis_trig == 1;
#define JSX_VIS_LAUNCH(M, C, H)                                                                                    \
    hipLaunchKernelGGL((jsx_k_vis<M, C, H>), grid, dim3(JSX_VIS_BLOCK), 0, st, x, y, cre, cim, u, v, dst, n_src, n_vis, \
                       chunk, sgn)
    if (!mixed) {
        if (cim) JSX_VIS_LAUNCH(false, true, false);
        else JSX_VIS_LAUNCH(false, false, false);
    } else if (hw) {
        if (cim) JSX_VIS_LAUNCH(true, true, true);
        else JSX_VIS_LAUNCH(true, false, true);
    } else {
        if (cim) JSX_VIS_LAUNCH(true, true, false);
        else JSX_VIS_LAUNCH(true, false, false);
    }
#undef JSX_VIS_LAUNCH
    HIPCHECK(ctx, hipGetLastError());
    if (n_chunks > 1) {
        hipLaunchKernelGGL(jsx_k_vis_reduce, dim3((unsigned)((n_vis + 255) / 256)), dim3(256), 0, st, dst, out, n_vis,
                           (int)n_chunks);
        HIPCHECK(ctx, hipGetLastError());
        if (!fixed) HIPCHECK(ctx, hipEventRecord(ctx->vis_ev, st));
    }
    return JSX_OK;
}

extern "C" int jsx_visibilities_device_noalloc(jsx_ctx *ctx, const double *x, const double *y, const double *c_re,
                                               const double *c_im, int64_t n_src, const double *u, const double *v,
                                               int64_t n_vis, int32_t isign, int32_t precision, double *vis,
                                               void *stream) {
    int rc = vis_check(ctx, x, y, c_re, n_src, u, v, n_vis, precision, vis);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    return vis_launch(ctx, x, y, c_re, c_im, n_src, u, v, n_vis, isign, precision, vis,
                      stream ? (hipStream_t)stream : ctx->stream, ctx->vis_scratch, (size_t)JSX_VIS_SCRATCH_MAX);
}

extern "C" int jsx_visibilities_device(jsx_ctx *ctx, const double *x, const double *y, const double *c_re,
                                       const double *c_im, int64_t n_src, const double *u, const double *v,
                                       int64_t n_vis, int32_t isign, int32_t precision, double *vis, void *stream) {
    int rc = vis_check(ctx, x, y, c_re, n_src, u, v, n_vis, precision, vis);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    return vis_launch(ctx, x, y, c_re, c_im, n_src, u, v, n_vis, isign, precision, vis,
                      stream ? (hipStream_t)stream : ctx->stream);
}

extern "C" int jsx_visibilities(jsx_ctx *ctx, const double *x, const double *y, const double *c_re,
                                const double *c_im, int64_t n_src, const double *u, const double *v, int64_t n_vis,
                                int32_t isign, int32_t precision, double *vis) {
    int rc = vis_check(ctx, x, y, c_re, n_src, u, v, n_vis, precision, vis);
    if (rc) return rc;
    if (n_vis == 0) return JSX_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    const size_t nsrc_arrays = c_im ? 4 : 3;
    const size_t bytes = sizeof(double) * ((size_t)n_src * nsrc_arrays + (size_t)n_vis * 4);
    char *mem = nullptr;
    if (hipMalloc(&mem, bytes) != hipSuccess) return fail(ctx, JSX_ENOMEM, "visibility buffers: %zu bytes", bytes);
    double *dx = (double *)mem, *dy = dx + n_src, *dcr = dy + n_src, *dci = c_im ? dcr + n_src : nullptr;
    double *du = dcr + n_src * (c_im ? 2 : 1), *dv = du + n_vis, *dvis = dv + n_vis;
    hipStream_t st = ctx->stream;
    hipError_t e = hipSuccess;
    const size_t sb = sizeof(double) * (size_t)n_src, vb = sizeof(double) * (size_t)n_vis;
    if (n_src > 0) {
        if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dy, y, sb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dcr, c_re, sb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && c_im) e = hipMemcpyAsync(dci, c_im, sb, hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(du, u, vb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dv, v, vb, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(mem);
        return fail(ctx, JSX_EHIP, "visibility upload: %s", hipGetErrorString(e));
    }
    rc = vis_launch(ctx, dx, dy, dcr, dci, n_src, du, dv, n_vis, isign, precision, dvis, st);
    if (rc == JSX_OK) {
        e = hipMemcpyAsync(vis, dvis, 2 * vb, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(ctx, JSX_EHIP, "visibilities: %s", hipGetErrorString(e));
    } else {
        (void)hipStreamSynchronize(st);
    }
    (void)hipFree(mem);
    return rc;
}

/* ---------------- cosmology (host, once per image) ----------------
 * comoving_transfer_distance2 + mas_to_pc (src/utils.cpp:186-208, 373-378):
 * controlled Dormand-Prince (abs = rel = 1.49e-8, dt0 = 1e-9) over the Ctd
 * integrand, with the same step controller the kernels use. */
static double ctd(double t) {
    const double omega_M = 0.2865, omega_V = 0.7134130719051658, gamma_nu = 8.69280948342326e-05;
    return pow((omega_M + (1. + t) * gamma_nu) * (1. + t) * (1. + t) * (1. + t) + omega_V, -0.5);
}

extern "C" double jsx_mas_to_pc(double z) {
    const double eps_abs = 1.49e-8, eps_rel = 1.49e-8, EPS = 2.220446049250313080847e-16;
    double x = 0.0, t = 0.0, dt = 1E-09, k1 = 0.0;
    int first = 1;
    while (z - t > EPS) {
        if ((t + dt) - z > EPS) dt = z - t;
        if (first) { k1 = ctd(t); first = 0; }
        for (int tries = 0; tries < 500; ++tries) {
            double k3 = ctd(t + dt * DP_A3);
            double k4 = ctd(t + dt * DP_A4);
            double k5 = ctd(t + dt * DP_A5);
            double k6 = ctd(t + dt);
            double xn = x + dt * DP_C1 * k1 + dt * DP_C3 * k3 + dt * DP_C4 * k4 + dt * DP_C5 * k5 + dt * DP_C6 * k6;
            double k7 = ctd(t + dt);
            double xe = dt * DP_DC1 * k1 + dt * DP_DC3 * k3 + dt * DP_DC4 * k4 + dt * DP_DC5 * k5 + dt * DP_DC6 * k6 +
                        dt * DP_DC7 * k7;
            double err = fabs(xe) / (eps_abs + eps_rel * (fabs(x) + dt * fabs(k1)));
            if (err > 1.0) {
                double fac = 0.9 * pow(err, -1.0 / 3);
                dt = dt * (fac > 0.2 ? fac : 0.2);
                continue;
            }
            t = t + dt;
            if (err < 0.5) {
                double e2 = pow(5.0, -5.0) > err ? pow(5.0, -5.0) : err;
                dt = dt * (0.9 * pow(e2, -1.0 / 5.0));
            }
            x = xn;
            k1 = k7;
            break;
        }
    }
    const double H0 = 69.32;
    double result = (299792.458 / H0) * pow(10., 6.) * x;
    double d_a = result / (1. + z);
    return 4.84813681109536e-09 * d_a;
}

/* ---------------- batch ----------------
 * Requests are taken in order into groups that share (precision, stokes) and
 * whose output planes fit the HBM budget; each group is one persistent launch
 * (apex columns first), then its planes are copied to the callers' buffers. */
static int batch_group(jsx_ctx *ctx, const jsx_request *reqs, int32_t i0, int32_t i1) {
    std::vector<BatchItem> items(i1 - i0);
    std::vector<jsx_outputs> dev(i1 - i0);
    int rc = JSX_OK;
    int32_t n_alloc = 0;
    for (int32_t i = i0; i < i1 && rc == JSX_OK; ++i) {
        const jsx_request &r = reqs[i];
        BatchItem &it = items[i - i0];
        it.npix = r.image.n_across * r.image.n_along;
        if ((rc = jsx_alloc_outputs(ctx, it.npix, JSX_STOKES_FULL, &dev[i - i0]))) break;
        n_alloc++;
        std::lock_guard<std::mutex> lk(ctx->mu);
        if ((rc = lower_model(ctx, r.jet, r.params.flags, it.m))) break;
        if ((rc = lower_solver(ctx, r.params, it.S))) break;
        if ((rc = lower_image(ctx, r.image, it.im, nullptr))) break;
        if (r.params.precision == JSX_PREC_MIXED) lower_mixed(r.jet, r.params, it.m, it.im, it.c);
        else std::memset(&it.c, 0, sizeof it.c);
        it.out = lower_out(dev[i - i0]);
    }
    if (rc == JSX_OK) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rc = launch_batch(ctx, ctx->stream, reqs[i0].params.precision, reqs[i0].params.stokes, items, 1);
        if (rc == JSX_OK) {
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = fail(ctx, JSX_EHIP, "batch kernel failed: %s", hipGetErrorString(e));
        }
    }
    for (int32_t i = i0; i < i0 + n_alloc; ++i) {
        const jsx_outputs &h = reqs[i].out_host, &d = dev[i - i0];
        size_t n = (size_t)items[i - i0].npix, nd = n * sizeof(double), nu = n * sizeof(uint32_t);
        struct { void *h; void *dv; size_t b; } cp[] = {
            {h.tau, d.tau, nd}, {h.I, d.I, nd}, {h.Q, d.Q, nd}, {h.U, d.U, nd}, {h.V, d.V, nd},
            {h.l, d.l, nd}, {h.flags, d.flags, nu}, {h.n_rhs, d.n_rhs, nu}, {h.n_steps, d.n_steps, nu}};
        if (rc == JSX_OK)
            for (auto &c : cp)
                if (c.h && hipMemcpy(c.h, c.dv, c.b, hipMemcpyDeviceToHost) != hipSuccess) {
                    rc = fail(ctx, JSX_EHIP, "copy back failed");
                    break;
                }
        jsx_free_outputs(ctx, &dev[i - i0]);
    }
    return rc;
}

extern "C" int jsx_observe_batch(jsx_ctx *ctx, const jsx_request *reqs, int32_t n_reqs) {
    if (!ctx || (!reqs && n_reqs) || n_reqs < 0) return JSX_EINVAL;
    /* device bytes per pixel of a group: the nine output planes, plus the
       triage pass's list, sorted list and two cost buckets (8 + 8 + 1 + 1) */
    const size_t per_pix = 6 * sizeof(double) + 3 * sizeof(uint32_t) + 18;
    size_t avail;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        const size_t held = ctx->used + split_workspace_bytes();   /* grids + the split workspace */
        avail = ctx->budget > held ? ctx->budget - held : 0;
    }
    for (int32_t i = 0; i < n_reqs; ++i) {
        const jsx_image_desc &im = reqs[i].image;
        if (im.n_across <= 0 || im.n_along <= 0 || (im.n_across % 2))
            return fail(ctx, JSX_EINVAL, "request %d: image needs n_along > 0 and even n_across > 0", i);
        if (!reqs[i].out_host.I) return fail(ctx, JSX_EINVAL, "request %d: out_host.I must be set", i);
        if ((size_t)(im.n_across * im.n_along) * per_pix > avail)
            return fail(ctx, JSX_EBUDGET, "request %d does not fit the HBM budget", i);
    }
    int32_t i0 = 0;
    while (i0 < n_reqs) {
        size_t bytes = 0;
        int32_t i1 = i0;
        while (i1 < n_reqs && reqs[i1].params.precision == reqs[i0].params.precision &&
               reqs[i1].params.stokes == reqs[i0].params.stokes) {
            size_t b = (size_t)(reqs[i1].image.n_across * reqs[i1].image.n_along) * per_pix;
            if (i1 > i0 && bytes + b > avail) break;
            bytes += b;
            ++i1;
        }
        int rc = batch_group(ctx, reqs, i0, i1);
        if (rc) return rc;
        i0 = i1;
    }
    return JSX_OK;
}

extern "C" int jsx_observe_batch_device(jsx_ctx *ctx, const jsx_request *reqs, int32_t n_reqs, void *stream) {
    if (!ctx || (!reqs && n_reqs) || n_reqs < 0) return JSX_EINVAL;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int32_t i0 = 0;
    while (i0 < n_reqs) {
        int32_t i1 = i0;
        std::vector<BatchItem> items;
        while (i1 < n_reqs && reqs[i1].params.precision == reqs[i0].params.precision &&
               reqs[i1].params.stokes == reqs[i0].params.stokes) {
            const jsx_request &r = reqs[i1];
            if (r.image.n_across <= 0 || r.image.n_along <= 0 || (r.image.n_across % 2))
                return fail(ctx, JSX_EINVAL, "request %d: image needs n_along > 0 and even n_across > 0", i1);
            if (!r.out_host.I) return fail(ctx, JSX_EINVAL, "request %d: outputs.I must be set", i1);
            BatchItem it;
            it.npix = r.image.n_across * r.image.n_along;
            int rc;
            if ((rc = lower_model(ctx, r.jet, r.params.flags, it.m))) return rc;
            if ((rc = lower_solver(ctx, r.params, it.S))) return rc;
            if ((rc = lower_image(ctx, r.image, it.im, nullptr))) return rc;
            if (r.params.precision == JSX_PREC_MIXED) lower_mixed(r.jet, r.params, it.m, it.im, it.c);
            else std::memset(&it.c, 0, sizeof it.c);
            it.out = lower_out(r.out_host);
            items.push_back(it);
            ++i1;
        }
        int rc = launch_batch(ctx, st, reqs[i0].params.precision, reqs[i0].params.stokes, items, 1);
        if (rc) return rc;
        i0 = i1;
    }
    return JSX_OK;
}

extern "C" size_t jsx_sizeof(int which) {
    switch (which) {
    case 0: return sizeof(jsx_jet_desc);
    case 1: return sizeof(jsx_image_desc);
    case 2: return sizeof(jsx_solver_params);
    case 3: return sizeof(jsx_outputs);
    case 4: return sizeof(jsx_grid_desc);
    case 5: return sizeof(jsx_request);
    case 6: return sizeof(jsx_grid_payload);
    case 7: return sizeof(jsx_request_timing);
    default: return 0;
    }
}

#include "jsx_serve.h"
