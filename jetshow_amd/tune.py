"""Autotuner for the batched solver's launch shape and scheduling.

Knobs (jsx_ctx_set_option): persistent workgroups per CU ("blocks_per_cu",
i.e. waves per SIMD the launch asks for), ray ids per queue block
("queue_block", the work-stealing chunk), per-XCD queues with stealing vs one
global queue ("steal"), the right-hand-side budget after which a full-Stokes
ray is split into time-parallel pieces ("defer_rhs"), how few lanes may still
hold a ray before splitting starts ("park_lanes"), and the Stokes-state layout
of a piece and its method ("piece_layout": Dormand-Prince in one lane with all
20 components, or the per-ray choice between that and exponential propagators)
with the exponential pieces' per-round budget ("piece_budget").

A sweep times each candidate on a workload (a list of (jet, image, params)
sent through one batched launch), keeps the fastest, and caches it in a JSON
file keyed by device name and a workload signature, so later contexts apply
the tuned shape without re-timing.  Results are identical across launch
shapes for FP64 and for mixed precision without splitting; with splitting
they agree within the integration tolerance.
"""
import hashlib
import itertools
import json
import os
import time

DEFAULT_SPACE = {
    "blocks_per_cu": [1, 2, 4, 8],
    "queue_block": [16, 64, 256],
    "steal": [0, 1],
    "defer_rhs": [512, 2048, 8192],
    "piece_layout": [1, 3],
    "park_lanes": [0, 8192],
    "piece_budget": [64, 128],
}
CACHE_PATH = os.environ.get("JSX_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "jetshow_amd",
                                                            "tune.json"))


def signature(items):
    """Workload signature: shapes, precisions and model kinds, not values."""
    h = hashlib.sha1()
    for jet, image, params in items:
        h.update(("%d,%d,%d,%d|%d,%d|%d,%d" % (jet.geometry, jet.bfield, jet.nfield, jet.vfield, image.n_across,
                                                image.n_along, params.precision, params.stokes)).encode())
    return h.hexdigest()[:16]


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # pragma: no cover - torch without a device
        return "unknown"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, data):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _time(ctx, items, reps):
    ctx.observe_batch(items)          # warm: code objects, workspaces, caches
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.observe_batch(items)
        best = min(best, time.perf_counter() - t0)
    return best


def sweep(ctx, items, space=None, reps=2, budget_s=120.0, log=None):
    """Coordinate descent over `space` (dict of option -> values) on `items`,
    from the context's current settings: each option in turn is set to each of
    its values with the others at the best found so far, and the fastest value
    kept; passes repeat until one changes nothing.  A truncated search still
    visits every option near the defaults first (a full product visited in
    sorted order would spend its budget on the first option's corner).
    Returns (best_options, table, complete): complete is False when budget_s
    ran out before a pass without change."""
    space = dict(DEFAULT_SPACE if space is None else space)
    keys = sorted(space)
    saved = {k: ctx.get_option(k) for k in keys}
    table = []
    seen = {}
    t_start = time.perf_counter()

    def run(opts):
        key = tuple(sorted(opts.items()))
        if key in seen:
            return seen[key]
        for k, v in opts.items():
            ctx.set_option(k, v)
        t = _time(ctx, items, reps)
        seen[key] = t
        table.append((t, dict(opts)))
        if log:
            log("%.4f s  %s" % (t, opts))
        return t

    complete = False
    try:
        best = {k: saved[k] for k in keys}
        best_t = run(best)
        while time.perf_counter() - t_start <= budget_s:
            changed = False
            for k in keys:
                for v in space[k]:
                    if time.perf_counter() - t_start > budget_s:
                        break
                    trial = dict(best)
                    trial[k] = v
                    t = run(trial)
                    if t < best_t:
                        best, best_t, changed = trial, t, True
            if not changed and time.perf_counter() - t_start <= budget_s:
                complete = True
                break
    finally:
        for k, v in saved.items():
            ctx.set_option(k, v)
    table.sort(key=lambda r: r[0])
    return (table[0][1] if table else {}), table, complete


def tune(ctx, items, space=None, reps=2, budget_s=120.0, cache=CACHE_PATH, force=False, log=None):
    """Apply the cached best shape for this device and workload, searching
    first when there is none, when the cached search was cut short by its
    budget, or when force=True.  Returns the options applied."""
    key = "%s|%s" % (_device_name(), signature(items))
    data = _load(cache) if cache else {}
    entry = data.get(key)
    if force or entry is None or not entry.get("complete", False):
        best, table, complete = sweep(ctx, items, space, reps, budget_s, log)
        data[key] = {"options": best, "seconds": table[0][0] if table else None, "tried": len(table),
                     "complete": complete, "when": time.strftime("%Y-%m-%d %H:%M:%S")}
        if cache:
            _save(cache, data)
    best = data[key]["options"]
    for k, v in best.items():
        ctx.set_option(k, v)
    return best


def apply_cached(ctx, items, cache=CACHE_PATH):
    """Apply a cached shape if one exists; returns it or None."""
    key = "%s|%s" % (_device_name(), signature(items))
    entry = _load(cache).get(key)
    if not entry:
        return None
    for k, v in entry["options"].items():
        ctx.set_option(k, v)
    return entry["options"]
