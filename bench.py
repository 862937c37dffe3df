"""Benchmark: integrated pixels/s of the per-pixel radiative-transfer solver.

Workload (one step = one full image through the hot path, inputs resident in
HBM): full-Stokes IQUV on the paraboloid jet (r0 = 0.1 pc at z = 1 pc, helical
B 0.2 G (r/pc)^-1 pitch 45 deg, N = 500 (r/pc)^-2 with s = 2.5 plus cold
thermal electrons at 0.1 x, Gamma 2 -> 15 over z in [0, 5] pc, LOS 0.05 rad,
512 x 512 pixels of 0.01 pc, 15.4 GHz observed at z = 0.1).

Timed: the batched mixed-precision solver (persistent work-queue kernel, FP32
coefficients and stages, FP64 roots / step control / tau placement).  Beside
it: the plain FP64 kernel (one thread per pixel), the error of the mixed
result against FP64, per-image p50/p99 latency, the roofline of the dominant
kernel and the CPU oracle on a bounded pixel sample.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

OPT_PERSISTENT = 0x100
OPT_APEX_FIRST = 0x200

# FP64 / FP32 VALU operations per right-hand-side evaluation (per lane, one
# loop trip of the kernel), measured with rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA,
# TRANS}_{F64,F32} on single-wavefront launches (bench.py --calibrate; FMA = 2);
# the counter table is in profiles/valu_ops_per_rhs_calibration.md.
OPS_PER_RHS = {
    "mixed_full": {"f64": 83.5, "f32": 312.0},
    "fp64_full": {"f64": 1290.1, "f32": 0.0},
    "fp64_I_grid": {"f64": 906.3, "f32": 0.0},
    "mixed_I_grid": {"f64": 477.8, "f32": 99.2},
    "fp64_I_cone": {"f64": 906.3, "f32": 0.0},
    # per piece right-hand side (20 components, one lane) and per coefficient
    # evaluation of an exponential piece (2 per step: coefficients, the two
    # exponentials, error estimate and propagator update)
    "piece20": {"f64": 146.9, "f32": 590.1},
    "exp": {"f64": 726.9, "f32": 236.0},
}
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
# HBM bytes per image of a kernel on the default headline image (library
# defaults), from rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a copy of
# known size (profiles/c1_hbm_traffic_counters.md); reported as roofline
# "traffic" when that kernel dominates the default headline run.
TRAFFIC_BYTES_C1 = {"jsx_k_split_magnus": 299.2e6, "jsx_k_batch<PolMixed<4>>": 13.7e6}
# the same for the c2 workload's FP64 batch kernel, per launch of both images
# (profiles/c2_grid_fields_fp64_hbm_traffic_counters.md)
TRAFFIC_BYTES_C2_FP64 = 82.4e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-side", type=int, default=512)
    ap.add_argument("--rel-tol", type=float, default=1e-7, help="mixed-precision Stokes rel. tolerance")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c2", "c3", "c4", "vis"],
                    help="c0: conical Stokes-I image, FP64 (the CPU-runnable case); "
                         "c1: paraboloid full-Stokes image (headline); c2: grid-defined fields, 1024^2 at two "
                         "frequencies; c3: batched ABC sweep of cone models; c4: request server under Poisson "
                         "arrivals; vis: visibilities of the c1 Stokes-I image at scattered (u, v)")
    ap.add_argument("--n-vis", type=int, default=131072, help="vis: visibilities per step")
    ap.add_argument("--vis-precision", default="mixed", choices=["fp64", "mixed"], help="vis: precision")
    ap.add_argument("--rate", type=float, default=200.0, help="c4: mean arrivals per second")
    ap.add_argument("--duration", type=float, default=10.0, help="c4: seconds of arrivals")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"], help="c2: solver precision")
    ap.add_argument("--models", type=int, default=256, help="c3: models in the sweep (x2 frequencies)")
    ap.add_argument("--tune", type=float, default=0.0, metavar="SECONDS",
                    help="c3: autotune the launch shape (tune.sweep) on the first 128 models for this many seconds "
                         "before timing")
    ap.add_argument("--set", action="append", default=[], metavar="OPTION=VALUE",
                    help="solver option applied to the context before the run (jsx_ctx_set_option), repeatable")
    ap.add_argument("--traffic", action="store_true",
                    help="one untimed launch of the headline image, then a 256 MiB device copy as the reference "
                         "for rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs")
    ap.add_argument("--calibrate", action="store_true",
                    help="single-wavefront launches (2 rays) of the FP64 and mixed full-Stokes kernels for "
                         "rocprofv3 instruction counters; prints the RHS evaluations per launch")
    a = ap.parse_args()
    a.options = {}
    for kv in a.set:
        k, v = kv.split("=", 1)
        a.options[k] = float(v)
    return a


def new_context(a):
    """Context with the options of --set applied (they appear in the JSON config)."""
    from jetshow_amd import host
    ctx = host.Context(0)
    for k, v in a.options.items():
        ctx.set_option(k, v)
    return ctx


def calibrate():
    """One wavefront per launch: every loop trip of that wavefront is one RHS
    evaluation for each active lane, so counters / max(n_rhs) = VALU operations
    per RHS evaluation per lane.  Launches, in order: FP64 full Stokes, mixed
    full Stokes (paraboloid, analytic fields), FP64 Stokes I, mixed Stokes I
    (grid-defined fields)."""
    from jetshow_amd import abi, grf, host, models
    ctx = host.Context(0)
    j, im, p = models.paraboloid_helical_model(n_side=2, pixel_pc=0.3)
    im.n_along = 1
    runs = [("fp64_full", j, im, p), ("mixed_full", j, im, models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6))]
    (B, N, G), geo = grf.jet_grids(96, 24)
    hs = [ctx.upload_grid(v, *geo, nan_value=nv) for v, nv in ((B, 0.0), (N, 0.0), (G, 1.0))]
    jg, img, pg = models.grid_jet_model(hs, n_side=2, pixel_pc=0.3)
    img.n_along = 1
    runs += [("fp64_I_grid", jg, img, pg), ("mixed_I_grid", jg, img, models.with_precision(pg, abi.PREC_MIXED))]
    # FP64 Stokes I on the conical workload (its 0.05 pc pixels): 2 x 32 rays,
    # one wavefront, the two middle rows, some of which cross the jet
    jc, imc, pc = models.conical_toroidal_model(n_side=2, pixel_pc=0.05)
    imc.n_along = 32
    runs += [("fp64_I_cone", jc, imc, pc)]
    for name, jj, ii, q in runs:
        r = ctx.observe(jj, ii, q)
        print(json.dumps({"calibration": name, "max_n_rhs": int(r["n_rhs"].max()),
                          "n_rhs": r["n_rhs"].reshape(-1).tolist()}), flush=True)
    # The piece kernels: the same two rays, each parked after 64 right-hand
    # sides into a single piece (two lanes of one wavefront do all the work),
    # by 20-component Dormand-Prince and by exponential propagators; the
    # longest piece's right-hand sides are the wavefront's loop trips.  The
    # other wavefronts of those launches only find the queue empty.
    qm = models.with_precision(p, abi.PREC_MIXED, rel_tol=1e-6)
    qm.flags |= OPT_PERSISTENT
    for name, layout in (("mixed_full_piece20", 1), ("mixed_full_exp", 2)):
        for k, v in (("defer_rhs", 64), ("park_lanes", 4e9 - 1), ("piece_layout", layout), ("piece_budget", 0),
                     ("piece_target", 1)):
            ctx.set_option(k, v)
        r = ctx.observe(j, im, qm)
        rhs = ctx.get_option("stat_rhs_pieces" if layout == 1 else "stat_rhs_exp")
        print(json.dumps({"calibration": name, "pieces": int(ctx.get_option("stat_pieces")),
                          "piece_rhs": int(rhs), "max_piece_rhs": int(ctx.get_option("stat_piece_steps_max")),
                          "n_rhs": r["n_rhs"].reshape(-1).tolist()}), flush=True)
    ctx.close()


def traffic_probe(a):
    """HBM traffic counters need a launch of known bytes to scale them: the
    headline image once (library defaults), then torch's copy of 256 MiB
    (268435456 bytes read and written)."""
    import torch

    from jetshow_amd import abi, models
    ctx = new_context(a)
    j, im, p64 = models.paraboloid_helical_model(n_side=a.n_side)
    pm = models.with_precision(p64, abi.PREC_MIXED, rel_tol=a.rel_tol)
    pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
    ctx.observe(j, im, pm)
    x = torch.ones(1 << 25, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"traffic_probe": "done", "copy_bytes": 1 << 28, "rhs_exp": ctx.get_option("stat_rhs_exp"),
                      "pieces": ctx.get_option("stat_pieces")}), flush=True)
    ctx.close()


def _log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


KERNEL_STATS = ("stat_batch_ms", "stat_piece_ms", "stat_piece_dp_ms", "stat_piece_exp_ms", "stat_split_ms",
                "stat_launches", "stat_rhs_batch", "stat_rhs_pieces", "stat_rhs_exp", "stat_parked",
                "stat_parked_exp", "stat_pieces")


def _peak(ops):
    f = ops["f64"] + ops["f32"]
    return f / (ops["f64"] / PEAK_TFLOPS["f64"] + ops["f32"] / PEAK_TFLOPS["f32"])


def kernel_rooflines(kern, batch_ops="mixed_full", batch_name="jsx_k_batch<PolMixed<4>>"):
    """Per kernel of the batched launch: algorithmic VALU operations (right-hand
    sides it evaluated x operations per right-hand side of that kernel) over
    its mean time by HIP events, against the FP64/FP32 vector peaks weighted
    by its op mix.  Returns (table, name of the longest kernel)."""
    rows = {
        batch_name: (kern["stat_rhs_batch"], kern["stat_batch_ms"], batch_ops),
        "jsx_k_split_piece20": (kern.get("stat_rhs_pieces", 0.0), kern.get("stat_piece_dp_ms", 0.0), "piece20"),
        "jsx_k_split_magnus": (kern.get("stat_rhs_exp", 0.0), kern.get("stat_piece_exp_ms", 0.0), "exp"),
    }
    table = {}
    for name, (rhs, ms, key) in rows.items():
        ops = OPS_PER_RHS[key]
        f = ops["f64"] + ops["f32"]
        ach = f * rhs / (ms / 1e3) / 1e12 if ms > 0 else 0.0
        table[name] = {"ms": ms, "rhs": rhs, "ops_per_rhs": f, "achieved_TFLOPs": ach, "peak_TFLOPs": _peak(ops),
                       "frac": ach / _peak(ops)}
    dom = max(table, key=lambda k: table[k]["ms"])
    return table, dom


def main():
    a = parse()
    if a.calibrate:
        calibrate()
        return
    if a.traffic:
        traffic_probe(a)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (every ray is independent, so images "
              "and models would shard across GPUs by whole requests). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "c3":
        sweep_bench(a)
        return
    if a.workload == "c2":
        grid_bench(a)
        return
    if a.workload == "c4":
        serve_bench(a)
        return
    if a.workload == "vis":
        vis_bench(a)
        return
    if a.workload == "c0":
        cone_bench(a)
        return
    import ctypes as C

    import torch

    from jetshow_amd import abi, host, models

    lib = host.lib()
    dev = torch.device("cuda:0")
    torch.cuda.init()
    ctx = new_context(a)
    j, im, p64 = models.paraboloid_helical_model(n_side=a.n_side)
    npix = im.n_across * im.n_along
    pm = models.with_precision(p64, abi.PREC_MIXED, rel_tol=a.rel_tol)
    pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST

    def planes():
        t = {k: torch.zeros(npix, dtype=torch.float64, device=dev) for k in ("tau", "I", "Q", "U", "V", "l")}
        for k in ("flags", "n_rhs", "n_steps"):
            t[k] = torch.zeros(npix, dtype=torch.int32, device=dev)
        o = abi.Outputs()
        for k in ("tau", "I", "Q", "U", "V", "l"):
            setattr(o, k, C.cast(C.c_void_p(t[k].data_ptr()), C.POINTER(C.c_double)))
        for k in ("flags", "n_rhs", "n_steps"):
            setattr(o, k, C.cast(C.c_void_p(t[k].data_ptr()), C.POINTER(C.c_uint32)))
        return t, o

    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)

    def run(params, o):
        rc = lib.jsx_observe(ctx.handle, C.byref(j), C.byref(im), C.byref(params), C.byref(o), sh)
        if rc:
            raise RuntimeError(lib.jsx_last_error(ctx.handle).decode())

    def timed(params, steps, warmup, tag):
        t, o = planes()
        for i in range(warmup):
            _log("%s warmup %d/%d" % (tag, i + 1, warmup))
            run(params, o)
            stream.synchronize()
        torch.cuda.synchronize()
        ctx.set_option("timing", 1)
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        with torch.cuda.stream(stream):
            for e0, e1 in evs:
                e0.record(stream)
                run(params, o)
                e1.record(stream)
        # the launches are asynchronous: report progress while a long one runs
        t_last = t0
        while not evs[-1][1].query():
            time.sleep(0.5)
            if time.perf_counter() - t_last >= 20.0:
                t_last = time.perf_counter()
                _log("%s: running, %.0f s" % (tag, t_last - t0))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
        _log("%s: %d steps, %.1f ms mean" % (tag, steps, lat.mean()))
        return t, lat, wall

    # headline: batched mixed-precision full-Stokes solver (long rays split in time)
    tm, lat_m, wall_m = timed(pm, a.steps, a.warmup, "mixed batched")
    kern = {k: ctx.get_option(k) for k in KERNEL_STATS}
    ctx.set_option("timing", 0)
    ms_step = float(lat_m.mean())
    value = npix / (ms_step / 1e3)
    res = {k: v.cpu().numpy() for k, v in tm.items()}
    n_rhs_tot = float(res["n_rhs"].astype(np.float64).sum())

    # plain FP64 kernel (one thread per pixel) and the accuracy of the mixed
    # result against it; its time is set by the slowest (Faraday-thick) ray
    plain = {}
    if not a.no_plain:
        _log("FP64 plain kernel, %d image(s): the slowest ray takes minutes" % max(1, a.plain_steps))
        t64, lat64, _ = timed(p64, max(1, a.plain_steps), 0, "fp64 plain")
        ctx.set_option("timing", 0)
        r64 = {k: v.cpu().numpy() for k, v in t64.items()}
        hit = (r64["flags"] & abi.PIX_HIT) != 0
        assert np.array_equal(hit, (res["flags"] & abi.PIX_HIT) != 0), "hit masks differ between FP64 and mixed"
        I64, Im = r64["I"][hit], res["I"][hit]
        pos = I64 > 0
        rel = np.abs(Im[pos] - I64[pos]) / I64[pos]
        peak = I64.max()
        qerr = max(np.abs(res[k][hit] - r64[k][hit]).max() for k in ("Q", "U", "V")) / peak
        evpa = 0.5 * np.degrees(np.arctan2(res["U"][hit], res["Q"][hit]) - np.arctan2(r64["U"][hit], r64["Q"][hit]))
        evpa = (evpa + 90.0) % 180.0 - 90.0
        pol = np.hypot(r64["Q"][hit], r64["U"][hit]) > 1e-3 * peak
        plain = {
            "fp64_plain_pixels_per_s": npix / (lat64.mean() / 1e3),
            "fp64_plain_ms_per_image": float(lat64.mean()),
            "speedup_vs_fp64_plain": float(lat64.mean() / lat_m.mean()),
            "fp64_plain_max_n_rhs": int(r64["n_rhs"].max()),
            "max_rel_err_I": float(rel.max()) if rel.size else 0.0,
            "p99_rel_err_I": float(np.percentile(rel, 99)) if rel.size else 0.0,
            "max_err_QUV_over_peak_I": float(qerr),
            "max_evpa_err_deg": float(np.abs(evpa[pol]).max()) if pol.any() else 0.0,
            "hit_pixels": int(hit.sum()),
            "hit_mask_identical": True,
            "meets_error_bar": bool((rel.max() if rel.size else 0.0) <= 1e-5),
        }
        if not plain["meets_error_bar"]:
            _log("WARNING: max relative Stokes-I error %.3g exceeds the 1e-5 bar of the metric" % rel.max())

    # roofline of the dominant kernel: algorithmic VALU operations (right-hand
    # sides evaluated in that kernel x operations per right-hand side) over
    # its mean launch time, both measured in this run
    table, dom = kernel_rooflines(kern)
    t = table[dom]
    default_run = a.n_side == 512 and not a.options and abs(a.rel_tol - 1e-7) < 1e-12
    traffic = TRAFFIC_BYTES_C1.get(dom) if default_run else None
    roofline = {"bound": "valu", "pipe": "valu", "kernel": dom, "achieved": t["achieved_TFLOPs"],
                "peak": t["peak_TFLOPs"], "unit": "TFLOP/s", "frac": t["frac"], "traffic": traffic,
                "traffic_unit": "HBM bytes per image in this kernel (all its launches)",
                "hbm_GBps": traffic / (t["ms"] / 1e3) / 1e9 if traffic else None,
                "note": "VALU-bound transcendental code, no MFMA: peak = FP64/FP32 vector peaks weighted by the "
                        "kernel's op mix; ops per RHS from rocprofv3 VALU counters; traffic from rocprofv3 "
                        "FETCH_SIZE/WRITE_SIZE runs of the default image (profiles/c1_hbm_traffic_counters.md)"}
    kernels = {"batch_ms": kern["stat_batch_ms"], "split_total_ms": kern["stat_split_ms"],
               "piece_ms": kern["stat_piece_ms"], "rhs_batch": kern["stat_rhs_batch"],
               "rhs_pieces": kern["stat_rhs_pieces"], "rhs_exp": kern["stat_rhs_exp"],
               "rays_split": kern["stat_parked"], "rays_exp": kern["stat_parked_exp"], "pieces": kern["stat_pieces"],
               "per_kernel": table}

    cpu = None
    if not a.no_cpu:
        _log("CPU oracle sample (%.0f s)" % a.cpu_seconds)
        cpu = cpu_baseline(j, im, p64, a.cpu_seconds, r64["n_rhs"] if not a.no_plain else None)

    out = {
        "metric": "integrated pixels/s at <=1e-5 rel Stokes-I err vs FP64; p50/p99 solve latency",
        "value": value,
        "unit": "pixels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (analytic jet model, deterministic)",
        "config": {"workload": "paraboloid helical-B full-Stokes IQUV, %dx%d pixels of 0.01 pc, 15.4 GHz; "
                               "batched FP32+compensated solver (FP64 roots/step control) vs plain FP64 kernel"
                               % (a.n_side, a.n_side),
                   "n_pixels": npix, "rel_tol_mixed": max(a.rel_tol, 1e-7), "rel_tol_fp64": p64.stokes_rel_tol},
        "latency_ms": {"p50": float(np.percentile(lat_m, 50)), "p99": float(np.percentile(lat_m, 99))},
        "rhs_evaluations": n_rhs_tot,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(plain)
    if a.options:
        out["config"]["options"] = a.options
    line = json.dumps(out)
    print(line, flush=True)
    if a.json_out:
        with open(a.json_out, "w") as f:
            f.write(line + "\n")
    ctx.close()


def fp64_error_sample(ctx, items, mixed_planes, label, tight=None, probe=True):
    """Per-image error of mixed results against FP64 on a sample of images:
    items = [(jet, image, params)] (the mixed request's params), mixed_planes
    = their planes ("I", "flags"; "Q", "U", "V" when full Stokes).  FP64 runs
    as one batched launch at relative tolerance 1e-8 (tight: that tolerance
    instead, for grid fields, whose kinks the error estimate undershoots).
    Pixels whose FP64 value is not determined -- not a number, or moving by
    more than 2e-5 when the electron density is raised by 1e-5 (the
    exponentially growing full-Stokes modes the reference's coefficients
    allow where the field lies near the line of sight) -- are counted and
    left out.  Returns the worst errors over the sample."""
    from jetshow_amd import abi, models
    fp = []
    for j, im, p in items:
        q = models.with_precision(p, abi.PREC_FP64, rel_tol=tight or 1e-8)
        q.stokes_abs_tol = 1e-20
        if tight:
            q.tau_abs_tol, q.tau_rel_tol = 1e-20, tight
        q.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
        fp.append((j, im, q))
    _log("%s: FP64 error sample over %d images" % (label, len(fp)))
    g64 = ctx.observe_batch(fp)
    dens = []
    for j, im, q in fp:
        jj = models.copy_struct(j)
        jj.n0 *= 1.0 + 1e-5
        dens.append((jj, im, q))
    gd = ctx.observe_batch(dens) if probe else g64
    worst_I, worst_quv, n_hit, n_und, p99 = 0.0, 0.0, 0, 0, []
    for a, d, m in zip(g64, gd, mixed_planes):
        a = {k: v.reshape(-1) for k, v in a.items()}
        d = {k: v.reshape(-1) for k, v in d.items()}
        hit = (a["flags"] & abi.PIX_HIT) != 0
        if not np.array_equal(hit, (m["flags"] & abi.PIX_HIT) != 0):
            raise AssertionError("%s: hit masks differ between FP64 and mixed" % label)
        with np.errstate(invalid="ignore", divide="ignore"):
            und = ((a["flags"] | m["flags"]) & abi.PIX_NAN) != 0
            und |= ~np.isfinite(a["I"]) | ~np.isfinite(m["I"])
            und |= ~(np.abs(d["I"] - a["I"]) <= 2e-5 * np.abs(a["I"]))
        ok = hit & ~und & (a["I"] > 0)
        n_hit += int(hit.sum())
        n_und += int((hit & und).sum())
        if not ok.any():
            continue
        rel = np.abs(m["I"][ok] - a["I"][ok]) / a["I"][ok]
        worst_I = max(worst_I, float(rel.max()))
        p99.append(float(np.percentile(rel, 99)))
        if "Q" in m:
            peak = a["I"][ok].max()
            worst_quv = max(worst_quv, max(float(np.abs(m[k][ok] - a[k][ok]).max()) for k in ("Q", "U", "V")) / peak)
    out = {"images": len(fp), "hit_pixels": n_hit, "undetermined_pixels": n_und, "max_rel_err_I": worst_I,
           "p99_rel_err_I": max(p99) if p99 else 0.0, "max_err_QUV_over_peak_I": worst_quv,
           "fp64_rel_tol": tight or 1e-8, "meets_error_bar": worst_I <= 1e-5}
    if not out["meets_error_bar"]:
        _log("WARNING: %s: max relative Stokes-I error %.3g exceeds the 1e-5 bar" % (label, worst_I))
    return out


def cone_bench(a):
    """The conical Stokes-I image (Cone half-angle 0.01 rad clipped to z <= 10
    pc, toroidal B 0.1 G (r/pc)^-1, N 500 (r/pc)^-2 with s = 2.5, Gamma 10,
    128^2 pixels of 0.05 pc, LOS 0.1, tau_max 1e4, I tolerance 1e-10/1e-8):
    the FP64 persistent launch timed per image (device outputs), checked bit
    for bit against the CPU oracle in its shared-math mode; the whole image on
    the CPU oracle with C-library math (the configuration the reference runs
    on its CPU path) timed beside it."""
    import ctypes as C

    import torch

    from jetshow_amd import abi, host, models
    from oracle import oracle

    lib = host.lib()
    dev = torch.device("cuda:0")
    torch.cuda.init()
    ctx = new_context(a)
    j, im, p = models.conical_toroidal_model()
    q = models.copy_struct(p)
    q.flags |= OPT_PERSISTENT
    npix = im.n_across * im.n_along
    t = {k: torch.zeros(npix, dtype=torch.float64, device=dev) for k in ("tau", "I", "l")}
    for k in ("flags", "n_rhs", "n_steps"):
        t[k] = torch.zeros(npix, dtype=torch.int32, device=dev)
    o = abi.Outputs()
    for k in ("tau", "I", "l"):
        setattr(o, k, C.cast(C.c_void_p(t[k].data_ptr()), C.POINTER(C.c_double)))
    for k in ("flags", "n_rhs", "n_steps"):
        setattr(o, k, C.cast(C.c_void_p(t[k].data_ptr()), C.POINTER(C.c_uint32)))
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)

    def run():
        rc = lib.jsx_observe(ctx.handle, C.byref(j), C.byref(im), C.byref(q), C.byref(o), sh)
        if rc:
            raise RuntimeError(lib.jsx_last_error(ctx.handle).decode())

    for _ in range(max(1, a.warmup)):
        run()
    stream.synchronize()
    ctx.set_option("timing", 1)
    steps = max(a.steps, 20)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for e0, e1 in evs:
        e0.record(stream)
        run()
        e1.record(stream)
    torch.cuda.synchronize()
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    kern = {k: ctx.get_option(k) for k in KERNEL_STATS}
    ctx.set_option("timing", 0)
    g = {k: v.cpu().numpy().view(np.uint32) if k in ("flags", "n_rhs", "n_steps") else v.cpu().numpy()
         for k, v in t.items()}
    ref = oracle.observe(j, im, p, None, math=oracle.MATH_JSX)
    identical = all(np.array_equal(g[k], ref[k]) for k in ("tau", "I", "l", "flags", "n_rhs"))
    if not identical:
        _log("WARNING: FP64 GPU image differs from the oracle")
    # CPU: the whole image, repeated to a measurable time
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ncores = max(1, min(ncores, 16))
    reps, t_cpu = 0, 0.0
    while t_cpu < 2.0 and reps < 10000:
        t0 = time.perf_counter()
        oracle.observe(j, im, p, None, math=oracle.MATH_LIBM, threads=ncores)
        t_cpu += time.perf_counter() - t0
        reps += 1
    table, dom = kernel_rooflines(kern, batch_ops="fp64_I_cone", batch_name="jsx_k_batch<PolFP64<1>>")
    dom = "jsx_k_batch<PolFP64<1>>"
    tr = table[dom]
    ms = float(lat.mean())
    out = {
        "metric": "integrated pixels/s at <=1e-5 rel Stokes-I err vs FP64; p50/p99 solve latency",
        "value": npix / (ms / 1e3), "unit": "pixels/s", "n_gpus": 1, "steps": steps, "warmup": max(1, a.warmup),
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (analytic jet model, deterministic)",
        "config": {"workload": "conical toroidal-B Stokes I, 128x128 pixels of 0.05 pc, 15.4 GHz, FP64 "
                               "(persistent launch) vs the CPU oracle on the whole image",
                   "n_pixels": npix, "hit_pixels": int(((g["flags"] & abi.PIX_HIT) != 0).sum())},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "identical_to_oracle": bool(identical),
        "rhs_evaluations": float(g["n_rhs"].astype(np.float64).sum()),
        "kernels": {"batch_ms": kern["stat_batch_ms"], "per_kernel": {dom: tr}},
        "roofline": {"bound": "valu", "kernel": dom, "achieved": tr["achieved_TFLOPs"], "peak": tr["peak_TFLOPs"],
                     "unit": "TFLOP/s", "frac": tr["frac"], "traffic": None,
                     "note": "a 128^2 image with 46 hit pixels: the launch is latency-bound, not VALU-bound"},
        "cpu_baseline": {"value": npix * reps / t_cpu, "unit": "pixels/s", "cores": ncores, "kind": "port",
                         "sample": "the whole image %d times, FP64 oracle with glibc libm, %.2f s" % (reps, t_cpu)},
    }
    line = json.dumps(out)
    print(line, flush=True)
    if a.json_out:
        with open(a.json_out, "w") as f:
            f.write(line + "\n")
    ctx.close()


def sweep_bench(a):
    """Batched ABC-style sweep: `models` cone models x 2 frequencies, full
    Stokes, mixed precision, all images in one persistent launch per step
    (device output planes, no host copies in the timed region)."""
    import ctypes as C

    import torch

    from jetshow_amd import abi, host, models

    lib = host.lib()
    dev = torch.device("cuda:0")
    torch.cuda.init()
    ctx = new_context(a)
    reqs_py = models.abc_sweep_models(a.models)
    n = len(reqs_py)
    sizes = [im.n_across * im.n_along for _, im, _, _ in reqs_py]
    total = int(sum(sizes))
    buf = {k: torch.zeros(total, dtype=torch.float64, device=dev) for k in ("tau", "I", "Q", "U", "V", "l")}
    for k in ("flags", "n_rhs", "n_steps"):
        buf[k] = torch.zeros(total, dtype=torch.int32, device=dev)
    reqs = (abi.Request * n)()
    off = 0
    for i, (j, im, p, _) in enumerate(reqs_py):
        p.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
        reqs[i].jet, reqs[i].image, reqs[i].params = j, im, p
        o = abi.Outputs()
        for k in ("tau", "I", "Q", "U", "V", "l"):
            setattr(o, k, C.cast(C.c_void_p(buf[k].data_ptr() + 8 * off), C.POINTER(C.c_double)))
        for k in ("flags", "n_rhs", "n_steps"):
            setattr(o, k, C.cast(C.c_void_p(buf[k].data_ptr() + 4 * off), C.POINTER(C.c_uint32)))
        reqs[i].out_host = o
        off += sizes[i]
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)

    def run():
        rc = lib.jsx_observe_batch_device(ctx.handle, reqs, n, sh)
        if rc:
            raise RuntimeError(lib.jsx_last_error(ctx.handle).decode())

    tuned = None
    if a.tune > 0:
        from jetshow_amd import tune
        sub = [(j, im, p) for j, im, p, _ in reqs_py[:256]]
        _log("autotuning on %d images for %.0f s" % (len(sub), a.tune))
        tuned = tune.tune(ctx, sub, reps=1, budget_s=a.tune, cache=None, force=True, log=_log)
        _log("tuned: %s" % tuned)

    for i in range(a.warmup):
        _log("sweep warmup %d/%d" % (i + 1, a.warmup))
        run()
        stream.synchronize()
    torch.cuda.synchronize()
    ctx.set_option("timing", 1)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    for e0, e1 in evs:
        e0.record(stream)
        run()
        e1.record(stream)
    torch.cuda.synchronize()
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    kern = {k: ctx.get_option(k) for k in KERNEL_STATS}
    ctx.set_option("timing", 0)
    flags = buf["flags"].cpu().numpy()
    nrhs = buf["n_rhs"].cpu().numpy().astype(np.int64)
    ms = float(lat.mean())
    table, dom = kernel_rooflines(kern)
    # FP64 error on the images of the first 8 models (only their planes come to the host)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    sample = [i for i, r in enumerate(reqs_py) if r[3] < 8]
    mixed = []
    for i in sample:
        d = {}
        for k in ("I", "Q", "U", "V", "flags"):
            v = buf[k][int(offs[i]):int(offs[i + 1])].cpu().numpy()
            d[k] = v.view(np.uint32) if k == "flags" else v
        mixed.append(d)
    err = fp64_error_sample(ctx, [reqs_py[i][:3] for i in sample], mixed, "sweep")
    cpu = None
    if not a.no_cpu:
        cpu = sweep_cpu_baseline(reqs_py, a.cpu_seconds)
    out = {
        "metric": "integrated pixels/s at <=1e-5 rel Stokes-I err vs FP64; p50/p99 solve latency",
        "value": total / (ms / 1e3), "unit": "pixels/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (cone models drawn from the priors, seed 20261203)",
        "config": {"workload": "ABC sweep: %d cone models x 2 frequencies (8.1, 15.4 GHz), full Stokes, log pixel "
                               "grids, sides in {64,128,256,512}, one persistent launch" % a.models,
                   "n_images": n, "n_pixels": total, "hit_pixels": int(((flags & 1) != 0).sum())},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "rhs_evaluations": float(nrhs.sum()),
        "fp64_error_sample": err,
        "kernels": {"batch_ms": kern["stat_batch_ms"], "piece_ms": kern["stat_piece_ms"],
                    "rhs_batch": kern["stat_rhs_batch"], "rhs_pieces": kern["stat_rhs_pieces"],
                    "rhs_exp": kern["stat_rhs_exp"], "rays_split": kern["stat_parked"],
                    "rays_exp": kern["stat_parked_exp"], "pieces": kern["stat_pieces"], "per_kernel": table},
        "roofline": {"bound": "valu", "pipe": "valu", "kernel": dom, "achieved": table[dom]["achieved_TFLOPs"],
                     "peak": table[dom]["peak_TFLOPs"], "unit": "TFLOP/s", "frac": table[dom]["frac"],
                     "traffic": None},
        "cpu_baseline": cpu,
    }
    if a.options:
        out["config"]["options"] = a.options
    if tuned is not None:
        out["config"]["tuned"] = {"options": tuned, "on": "first 128 models x 2 frequencies",
                                  "seconds": a.tune}
    line = json.dumps(out)
    print(line, flush=True)
    if a.json_out:
        with open(a.json_out, "w") as f:
            f.write(line + "\n")
    ctx.close()


def _device_requests(items, dev):
    """Requests with device output planes: one torch buffer per plane for all
    images; returns (Request array, buffers, sizes)."""
    import ctypes as C

    import torch

    from jetshow_amd import abi
    n = len(items)
    sizes = [im.n_across * im.n_along for _, im, _ in items]
    total = int(sum(sizes))
    buf = {k: torch.zeros(total, dtype=torch.float64, device=dev) for k in ("tau", "I", "Q", "U", "V", "l")}
    for k in ("flags", "n_rhs", "n_steps"):
        buf[k] = torch.zeros(total, dtype=torch.int32, device=dev)
    reqs = (abi.Request * n)()
    off = 0
    for i, (j, im, p) in enumerate(items):
        reqs[i].jet, reqs[i].image, reqs[i].params = j, im, p
        o = abi.Outputs()
        for k in ("tau", "I", "Q", "U", "V", "l"):
            setattr(o, k, C.cast(C.c_void_p(buf[k].data_ptr() + 8 * off), C.POINTER(C.c_double)))
        for k in ("flags", "n_rhs", "n_steps"):
            setattr(o, k, C.cast(C.c_void_p(buf[k].data_ptr() + 4 * off), C.POINTER(C.c_uint32)))
        reqs[i].out_host = o
        off += sizes[i]
    return reqs, buf, sizes


def grid_bench(a):
    """Grid-defined fields: B, N, Gamma on 2048 x 512 (z, r) node grids of
    seeded k^-11/3 random fields around the conical profiles, resident in HBM;
    Stokes I on 1024^2 pixels at 8.1 and 15.4 GHz, both images in one
    persistent launch per step."""
    import ctypes as C

    import torch

    from jetshow_amd import abi, grf, host, models

    lib = host.lib()
    dev = torch.device("cuda:0")
    torch.cuda.init()
    ctx = new_context(a)
    (B, N, G), geo = grf.jet_grids(2048, 512)
    hs = [ctx.upload_grid(v, *geo, nan_value=nv) for v, nv in ((B, 0.0), (N, 0.0), (G, 1.0))]
    prec = abi.PREC_FP64 if a.precision == "fp64" else abi.PREC_MIXED
    items = []
    for nu in (8.1, 15.4):
        j, im, p = models.grid_jet_model(hs, nu_ghz=nu, precision=prec)
        p.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
        items.append((j, im, p))
    reqs, buf, sizes = _device_requests(items, dev)
    total = int(sum(sizes))
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)

    def run():
        rc = lib.jsx_observe_batch_device(ctx.handle, reqs, len(items), sh)
        if rc:
            raise RuntimeError(lib.jsx_last_error(ctx.handle).decode())

    for i in range(a.warmup):
        _log("grid warmup %d/%d" % (i + 1, a.warmup))
        run()
        stream.synchronize()
    torch.cuda.synchronize()
    ctx.set_option("timing", 1)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    for e0, e1 in evs:
        e0.record(stream)
        run()
        e1.record(stream)
    torch.cuda.synchronize()
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    kern = {k: ctx.get_option(k) for k in ("stat_batch_ms", "stat_rhs_batch")}
    ctx.set_option("timing", 0)
    flags = buf["flags"].cpu().numpy()
    ms = float(lat.mean())
    ms_step = ms
    key = "fp64_I_grid" if prec == abi.PREC_FP64 else "mixed_I_grid"
    ops = OPS_PER_RHS.get(key, OPS_PER_RHS["fp64_full" if prec == abi.PREC_FP64 else "mixed_full"])
    f_rhs = ops["f64"] + ops["f32"]
    peak = f_rhs / (ops["f64"] / PEAK_TFLOPS["f64"] + (ops["f32"] / PEAK_TFLOPS["f32"] if ops["f32"] else 0.0))
    kname = "jsx_k_batch"
    if not kern["stat_batch_ms"] and prec == abi.PREC_MIXED:
        # mixed requests on grid fields run the cell-by-cell walk kernel, the
        # whole step: its coefficient evaluations over the step's time (ops
        # per evaluation from the mixed grid lane's calibration, an estimate)
        kname = "jsx_k_walk"
        kern = {"stat_batch_ms": ms_step, "stat_rhs_batch": float(buf["n_rhs"].cpu().numpy().astype(np.int64).sum())}
    ach = f_rhs * kern["stat_rhs_batch"] / (kern["stat_batch_ms"] / 1e3) / 1e12 if kern["stat_batch_ms"] else 0.0
    grid_gbs = 72.0 * kern["stat_rhs_batch"] / (kern["stat_batch_ms"] / 1e3) / 1e9 if kern["stat_batch_ms"] else 0.0
    err = None
    if prec == abi.PREC_MIXED:
        hp = {k: v.cpu().numpy() for k, v in buf.items()}
        offs = np.concatenate([[0], np.cumsum(sizes)])
        err = fp64_error_sample(ctx, items, [{"I": hp["I"][offs[i]:offs[i + 1]],
                                              "flags": hp["flags"][offs[i]:offs[i + 1]].view(np.uint32)}
                                             for i in range(len(items))], "grid", tight=1e-12, probe=False)
    cpu = None
    if not a.no_cpu:
        cpu = grid_cpu_baseline(items, (B, N, G), geo, a.cpu_seconds)
    out = {
        "metric": "integrated pixels/s at <=1e-5 rel Stokes-I err vs FP64; p50/p99 solve latency",
        "value": total / (ms / 1e3), "unit": "pixels/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == abi.PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded k^-11/3 random-field grids, seeds 1, 2, 3)",
        "config": {"workload": "grid-defined fields (2048x512 node grids in HBM), cylinder r 0.5 pc z<=20 pc, "
                               "Stokes I, 1024x1024 pixels of 0.002 pc at 8.1 and 15.4 GHz",
                   "n_pixels": total, "hit_pixels": int(((flags & 1) != 0).sum()),
                   "rel_tol": items[0][2].stokes_rel_tol},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "fp64_error_sample": err,
        "roofline": {"bound": "valu", "pipe": "valu", "kernel": kname, "achieved": ach, "peak": peak,
                     "unit": "TFLOP/s", "frac": ach / peak,
                     "traffic": TRAFFIC_BYTES_C2_FP64 if prec == abi.PREC_FP64 and not a.options else None,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE of the default c2 run)",
                     "hbm_GBps": (TRAFFIC_BYTES_C2_FP64 / (kern["stat_batch_ms"] / 1e3) / 1e9
                                  if prec == abi.PREC_FP64 and not a.options and kern["stat_batch_ms"] else None),
                     "grid_lookup_GBps_algorithmic": grid_gbs},
        "cpu_baseline": cpu,
    }
    if a.options:
        out["config"]["options"] = a.options
    line = json.dumps(out)
    print(line, flush=True)
    if a.json_out:
        with open(a.json_out, "w") as f:
            f.write(line + "\n")
    ctx.close()


def serve_bench(a):
    """Request server under Poisson arrivals (seed 4): half the requests are
    cone models drawn as in the ABC sweep, half grid-defined models with their
    own B/N/Gamma grids (sizes 128x512, 256x1024 or 512x2048 nodes, uploaded
    on admission); one or two frequencies each (one request per frequency);
    full Stokes, mixed precision; HBM admission budget 288 GB.  Afterwards the
    first 8 cone and 4 grid requests are checked against FP64."""
    import torch

    from jetshow_amd import abi, grf, host, models

    torch.cuda.init()
    ctx = new_context(a)
    rng = np.random.default_rng(4)
    n_arr = rng.poisson(a.rate * a.duration)
    t_arr = np.sort(rng.uniform(0.0, a.duration, n_arr))
    cones = models.abc_sweep_models(max(64, n_arr), seed=20261203, freqs_ghz=(8.1,))
    grid_sets = {}
    for nz, nr in ((512, 128), (1024, 256), (2048, 512)):
        (B, N, G), geo = grf.jet_grids(nz, nr)
        descs = []
        for nv in (0.0, 0.0, 1.0):
            d = abi.GridDesc()
            d.nz, d.nr = nz, nr
            d.z0, d.dz, d.r0, d.dr = geo
            d.nan_value = nv
            descs.append(d)
        grid_sets[(nz, nr)] = list(zip((abi.FIELD_B, abi.FIELD_N, abi.FIELD_V), descs, (B, N, G)))
    plan = []
    for i in range(n_arr):
        nf = 1 + int(rng.integers(0, 2))
        freqs = (8.1, 15.4)[:nf]
        if rng.uniform() < 0.5:
            j, im, p, _ = cones[i % len(cones)]
            for f in freqs:
                q = models.copy_struct(p)
                q.nu = f * 1e9 * 1.1
                plan.append((t_arr[i], j, im, q, (), None))
        else:
            key = list(grid_sets)[int(rng.integers(0, 3))]
            side = int((64, 128, 256)[int(rng.integers(0, 3))])
            for f in freqs:
                j, im, p = models.grid_jet_model((0, 0, 0), n_side=side, pixel_pc=2.0 / side, nu_ghz=f,
                                                 precision=abi.PREC_MIXED, stokes=abi.STOKES_FULL)
                plan.append((t_arr[i], j, im, p, grid_sets[key], key))
    # requests whose planes are kept for the FP64 check
    check_cone = [k for k, r in enumerate(plan) if not r[4]][:8]
    check_grid = [k for k, r in enumerate(plan) if r[4]][:4]
    checked = set(check_cone) | set(check_grid)
    srv = host.Server(ctx, abi.PREC_MIXED, abi.STOKES_FULL, max_active=64, hbm_budget=288 * 10 ** 9)
    ids = []
    _log("serving %d requests over %.0f s" % (len(plan), a.duration))
    t0 = time.perf_counter()
    for k, (t_rel, j, im, p, grids, _) in enumerate(plan):
        now = time.perf_counter() - t0
        if t_rel > now:
            time.sleep(t_rel - now)
        keep = ("I", "Q", "U", "V", "flags") if k in checked else ("I", "flags")
        ids.append((srv.submit(j, im, p, grids=grids, keep=keep), im.n_across * im.n_along))
    lats, px, kept = [], 0, {}
    for k, (rid, n) in enumerate(ids):
        planes, t = srv.wait(rid)
        lats.append(t["latency_s"])
        px += n
        if k in checked:
            kept[k] = {kk: v.reshape(-1) for kk, v in planes.items()}
        if (k + 1) % 500 == 0:
            _log("%d of %d requests done" % (k + 1, len(ids)))
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lats = np.array(lats) * 1e3
    err_cone = fp64_error_sample(ctx, [plan[k][1:4] for k in check_cone], [kept[k] for k in check_cone],
                                 "serve cones")
    handles = {}
    items_g = []
    for k in check_grid:
        _, j, im, p, grids, key = plan[k]
        if key not in handles:
            handles[key] = [ctx.upload_grid(v, d.z0, d.dz, d.r0, d.dr, nan_value=d.nan_value) for _, d, v in grids]
        jj = models.copy_struct(j)
        jj.grid_b, jj.grid_n, jj.grid_v = handles[key]
        items_g.append((jj, im, p))
    err_grid = fp64_error_sample(ctx, items_g, [kept[k] for k in check_grid], "serve grids", tight=1e-12,
                                 probe=False)
    out = {
        "metric": "integrated pixels/s at <=1e-5 rel Stokes-I err vs FP64; p50/p99 solve latency",
        "value": px / wall, "unit": "pixels/s", "n_gpus": 1, "steps": len(plan), "warmup": 0,
        "ms_per_step": float(lats.mean()), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (Poisson arrivals seed 4; cone priors seed 20261203; grid seeds 1-3)",
        "config": {"workload": "request server: Poisson %.0f req/s for %.0f s, half cone models (sides 64-512), "
                               "half grid models (grids 128x512..512x2048, sides 64-256), 1-2 frequencies, "
                               "full Stokes, HBM budget 288 GB" % (a.rate, a.duration),
                   "requests": len(plan), "arrivals": int(n_arr), "pixels": int(px), "epochs": st["epochs"]},
        "latency_ms": {"p50": float(np.percentile(lats, 50)), "p99": float(np.percentile(lats, 99)),
                       "max": float(lats.max())},
        "offered_pixels_per_s": px / a.duration,
        "wall_s": wall,
        "fp64_error_sample": {"cones": err_cone, "grids": err_grid},
    }
    print(json.dumps(out), flush=True)
    if a.json_out:
        with open(a.json_out, "w") as f:
            f.write(json.dumps(out) + "\n")
    ctx.close()


def grid_cpu_baseline(items, grids, geo, seconds):
    """CPU oracle on interleaved pixel sets of the two grid images."""
    from jetshow_amd import abi, models
    from oracle import oracle
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ncores = max(1, min(ncores, 16))
    table = []
    for v, nv in zip(grids, (0.0, 0.0, 1.0)):
        d = abi.GridDesc()
        d.nz, d.nr = v.shape
        d.z0, d.dz, d.r0, d.dr = geo
        d.nan_value = nv
        table.append((d, v))
    t_used, count, off, stride = 0.0, 0, 0, 997
    while t_used < seconds and off < stride:
        for j, im, p in items:
            q = models.copy_struct(p)
            q.precision = abi.PREC_FP64
            idx = np.arange(off, im.n_across * im.n_along, stride, dtype=np.int64)
            t0 = time.perf_counter()
            oracle.observe(j, im, q, idx, math=oracle.MATH_LIBM, grids=table, threads=ncores)
            t_used += time.perf_counter() - t0
            count += len(idx)
        off += 13
    return {"value": count / t_used, "unit": "pixels/s", "cores": ncores, "kind": "port",
            "sample": "every 997th pixel of both images from offsets 0, 13, ...: %d pixels, FP64 oracle with glibc "
                      "libm, %.1f s" % (count, t_used)}


# VALU floating-point operations per pixel-visibility pair of jsx_k_vis (FMA =
# 2, a hardware sine or cosine = 1), from rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA,
# TRANS}_{F64,F32} of the c1-image launch (262144 pixels x 131072 visibilities)
# divided by its pairs per lane (profiles/vis_c1_image_kernel_trace.md):
# FP64 sine/cosine; mixed with the hardware sine/cosine (default) and with the
# FP32 polynomials.
OPS_PER_PAIR = {"fp64": {"f64": 43.0, "f32": 0.0}, "mixed_hw": {"f64": 4.0, "f32": 6.0},
                "mixed_poly": {"f64": 6.0, "f32": 24.0}}
# all VALU instructions per pair (SQ_INSTS_VALU of the same launches) and the
# issue peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz lane-instructions/s
VALU_INSTR_PER_PAIR = {"fp64": 53.5, "mixed_hw": 9.48, "mixed_poly": 33.9}
VALU_ISSUE_PEAK_T = 256 * 4 * 16 * 2.4e9 / 1e12


def vis_pixels_and_uv(a):
    """The c1 image's pixel centres in radians (x across, y along the jet, as
    JetModelZoom's d and r_ob) and n_vis (u, v) points in wavelengths: 32
    baselines' elliptical tracks of an Earth-rotation synthesis out to
    5e8 wavelengths (VLBA at 2 cm), deterministic."""
    from jetshow_amd import abi, host, models
    j, im, p64 = models.paraboloid_helical_model(n_side=a.n_side)
    sizes = host.pixel_sizes(im)                         # [n_across, n_along] cm
    s_pc = sizes / abi.PC
    n_ac = im.n_across
    along = np.cumsum(s_pc, axis=1) - s_pc / 2
    across = (n_ac / 2 - np.arange(n_ac)[:, None] - 0.5) * s_pc
    rad_per_pc = abi.MAS_TO_RAD / host.mas_to_pc(0.1)
    x = (across * rad_per_pc).ravel()
    y = (along * rad_per_pc).ravel()
    r = np.random.default_rng(5)
    nb = 32
    per = -(-a.n_vis // nb)
    h = np.linspace(-np.pi / 2, np.pi / 2, per)
    uu, vv = [], []
    for b in range(nb):
        L = r.uniform(0.05, 1.0) * 5e8
        ang = r.uniform(0, np.pi)
        dec = r.uniform(0.3, 1.2)
        uu.append(L * np.cos(h + ang))
        vv.append(L * np.sin(h + ang) * np.sin(dec))
    u = np.concatenate(uu)[:a.n_vis]
    v = np.concatenate(vv)[:a.n_vis]
    return (j, im, p64), x, y, u, v


def vis_bench(a):
    """Visibilities of the c1 Stokes-I image (262144 scattered pixels) at
    n_vis (u, v) points: one jsx_visibilities_device launch per step, inputs
    resident in HBM; the image comes from one untimed mixed-precision solve."""
    import torch

    from jetshow_amd import abi, models
    dev = torch.device("cuda:0")
    torch.cuda.init()
    ctx = new_context(a)
    (j, im, p64), x, y, u, v = vis_pixels_and_uv(a)
    pm = models.with_precision(p64, abi.PREC_MIXED, rel_tol=a.rel_tol)
    pm.flags |= OPT_PERSISTENT | OPT_APEX_FIRST
    img = ctx.observe(j, im, pm)["I"].ravel()
    n_src, n_vis = img.size, u.size
    d = {k: torch.from_numpy(np.ascontiguousarray(t)).to(dev) for k, t in
         (("x", x), ("y", y), ("c", img), ("u", u), ("v", v))}
    prec = abi.PREC_MIXED if a.vis_precision == "mixed" else abi.PREC_FP64
    stream = torch.cuda.Stream()
    outs = {p_: torch.zeros(2 * n_vis, dtype=torch.float64, device=dev) for p_ in (abi.PREC_MIXED, abi.PREC_FP64)}

    def run(precision):
        ctx.visibilities_device(d["x"].data_ptr(), d["y"].data_ptr(), d["c"].data_ptr(), None, n_src,
                                d["u"].data_ptr(), d["v"].data_ptr(), n_vis, outs[precision].data_ptr(),
                                precision=precision, stream=stream.cuda_stream)

    for i in range(a.warmup):
        _log("vis warmup %d/%d" % (i + 1, a.warmup))
        run(prec)
    stream.synchronize()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        run(prec)
        e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    ms = float(lat.mean())
    _log("vis: %.3f ms per launch (events), %.3f ms wall per step" % (ms, wall / a.steps * 1e3))
    # accuracy of the timed precision against FP64 on the GPU and the oracle on a sample
    other = abi.PREC_FP64 if prec == abi.PREC_MIXED else abi.PREC_MIXED
    run(other)
    stream.synchronize()
    o = outs[prec].cpu().numpy()
    vis = o[0::2] + 1j * o[1::2]
    o64 = outs[abi.PREC_FP64].cpu().numpy()
    vis64 = o64[0::2] + 1j * o64[1::2]
    flux = float(np.abs(img).sum())
    acc = {"max_err_vs_fp64_over_flux": float(np.max(np.abs(vis - vis64)) / flux),
           "rel_l2_vs_fp64": float(np.linalg.norm(vis - vis64) / np.linalg.norm(vis64))}
    trig = int(ctx.get_option("vis_trig"))
    variant = "fp64" if prec == abi.PREC_FP64 else ("mixed_hw" if trig == 1 else "mixed_poly")
    ops = OPS_PER_PAIR[variant]
    pairs = float(n_src) * n_vis
    ach = (ops["f64"] + ops["f32"]) * pairs / (ms / 1e3) / 1e12
    peak = _peak(ops)
    cpu = None
    if not a.no_cpu:
        from threadpoolctl import threadpool_limits
        from oracle import fourier as ofour
        # the CPU leg also checks the FP64 visibilities on a sample
        sample = np.arange(0, n_vis, max(1, n_vis // 256))
        ref = ofour.direct_sum(x, y, img, u[sample], v[sample])
        acc["fp64_vs_oracle_sample_over_flux"] = float(np.max(np.abs(vis64[sample] - ref)) / flux)
        acc["oracle_sample"] = int(sample.size)
        t_used, done = 0.0, 0
        with threadpool_limits(1):
            while t_used < min(a.cpu_seconds, 10.0):
                k = np.arange(done * 7 % n_vis, n_vis, 4099)[:16]
                t1 = time.perf_counter()
                ofour.direct_sum(x, y, img, u[k], v[k])
                t_used += time.perf_counter() - t1
                done += len(k)
        cpu = {"value": done / t_used, "unit": "visibilities/s", "cores": 1, "kind": "port",
               "sample": "%d visibilities (every 4099th, several offsets) of the same image, all %d pixels, FP64 "
                         "numpy direct sum, %.1f s" % (done, n_src, t_used)}
    out = {
        "metric": "visibilities/s of a model image (direct sum over its pixels)",
        "value": n_vis / (ms / 1e3), "unit": "visibilities/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if prec == abi.PREC_MIXED else "f64", "data": "synthetic (c1 model image, "
        "Earth-rotation uv tracks)",
        "config": {"workload": "visibilities of the c1 paraboloid Stokes-I image (%d pixels) at %d scattered (u, v), "
                               "|uv| <= 5e8 wavelengths" % (n_src, n_vis),
                   "n_pixels": n_src, "n_vis": n_vis, "pairs_per_s": pairs / (ms / 1e3)},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "accuracy": acc,
        "roofline": {"bound": "valu", "pipe": "valu", "kernel": "jsx_k_vis", "achieved": ach, "peak": peak,
                     "unit": "TFLOP/s", "frac": ach / peak, "traffic": None,
                     "valu_issue": {"variant": variant, "instr_per_pair": VALU_INSTR_PER_PAIR[variant],
                                    "achieved_T_per_s": VALU_INSTR_PER_PAIR[variant] * pairs / (ms / 1e3) / 1e12,
                                    "peak_T_per_s": VALU_ISSUE_PEAK_T,
                                    "frac": VALU_INSTR_PER_PAIR[variant] * pairs / (ms / 1e3) / 1e12 /
                                    VALU_ISSUE_PEAK_T},
                     "note": "VALU-bound sine/cosine per pixel-visibility pair (no MFMA shape: the exponent couples "
                             "every pixel with every visibility); events span jsx_k_vis and its chunk reduction"},
        "cpu_baseline": cpu,
    }
    if a.options:
        out["config"]["options"] = a.options
    line = json.dumps(out)
    print(line, flush=True)
    if a.json_out:
        with open(a.json_out, "w") as f:
            f.write(line + "\n")
    ctx.close()


def sweep_cpu_baseline(reqs_py, seconds):
    """CPU oracle (FP64 counterpart of each request) on interleaved pixel
    subsets of the sweep's first 64 models, ~`seconds` of CPU work."""
    from jetshow_amd import abi, models
    from oracle import oracle
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ncores = max(1, min(ncores, 16))
    sub = [r for r in reqs_py if r[3] < 64]
    t_used, count, rounds = 0.0, 0, 0
    stride = 61
    off = 0
    while t_used < seconds and off < stride:
        for j, im, p, _ in sub:
            q = models.copy_struct(p)
            q.precision = abi.PREC_FP64
            q.stokes_rel_tol = 1e-8
            idx = np.arange(off, im.n_across * im.n_along, stride, dtype=np.int64)
            t0 = time.perf_counter()
            oracle.observe(j, im, q, idx, math=oracle.MATH_LIBM, threads=ncores)
            t_used += time.perf_counter() - t0
            count += len(idx)
        off += 7
        rounds += 1
    return {"value": count / t_used, "unit": "pixels/s", "cores": ncores, "kind": "port",
            "sample": "%d round(s) of every %dth pixel of the 128 images of the first 64 models, FP64 oracle with "
                      "glibc libm, %.1f s" % (rounds, stride, t_used)}


def cpu_baseline(j, im, p64, seconds, n_rhs_all=None):
    """The CPU oracle (C-library math, FP64, OpenMP dynamic schedule over
    pixels) timed on interleaved pixel sets of the same image with the host's
    cores: set k holds every 61st pixel from offset 7k (61 is coprime with the
    row length, so every column -- the Faraday-thick apex columns included --
    is sampled in proportion).

    The cost of a pixel is set by its right-hand-side count, which spans seven
    orders of magnitude on this image, so a uniform sample extrapolates
    poorly.  When the FP64 GPU run's per-pixel counts are at hand (the same
    algorithm: its counts equal the oracle's in the shared-math mode), the
    whole-image CPU time is the sample's time per right-hand side times the
    image's total: value = pixels / (t_sample * R_image / R_sample).  The
    plain sample rate (pixels done / time) is reported beside it."""
    from oracle import oracle
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ncores = max(1, min(ncores, 16))
    npix = im.n_across * im.n_along
    stride = 61
    done, t_used, count, r_used = 0, 0.0, 0, 0.0
    off = 0
    while t_used < seconds and off < stride:
        idx = np.arange(off, npix, stride, dtype=np.int64)
        t0 = time.perf_counter()
        o = oracle.observe(j, im, p64, idx, math=oracle.MATH_LIBM, threads=ncores)
        t_used += time.perf_counter() - t0
        r_used += float(o["n_rhs"].astype(np.float64).sum())
        count += len(idx)
        off += 7
        done += 1
    sample_rate = count / t_used
    out = {"value": sample_rate, "unit": "pixels/s", "cores": ncores, "kind": "port",
           "sample": "%d interleaved pixel sets (every %dth pixel from offsets 0, 7, ...) = %d pixels of the same "
                     "image, FP64 oracle with glibc libm, %.1f s" % (done, stride, count, t_used)}
    if n_rhs_all is not None and r_used > 0:
        r_all = float(np.asarray(n_rhs_all, dtype=np.float64).sum())
        t_image = t_used * r_all / r_used
        out["value"] = npix / t_image
        out["sample_rate_pixels_per_s"] = sample_rate
        out["sample"] += ("; whole-image time = sample time x (image right-hand sides %.4g / sample's %.4g) = "
                          "%.1f s" % (r_all, r_used, t_image))
    return out


if __name__ == "__main__":
    main()
