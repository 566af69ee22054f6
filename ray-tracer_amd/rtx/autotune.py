"""Autotuner for librtx launch shapes.

Every knob of rtx_tuning (include/rtx.h) changes only how work is scheduled, never a
result: LDS stack depth and loop structure of the traversal, the persistent grid, shade
block size and the wavefront batch size. The tuner times a caller-supplied workload under
candidate settings by coordinate descent (one knob at a time, keep the fastest), and can
cache the winner per GPU name in tuning.json beside this file.
"""
import json
import os
import time

import rtx

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

SPACE = {
    # trace entry points
    "trace_lds_depth": [8, 12, 16],
    "trace_loop": [0, 1, 2],
    "ray_sort": [0, 1],
    "wait_lanes": [0, 4, 8, 16, 24, 32],
    "refill_lanes": [12, 18, 24, 32, 40],
    "trace_grid_blocks": [1024, 2048, 4096, 8192],
    "leaf_coop": [0, 1],
    "trace_tree": [0, 1],
    # renderer
    "render_lds_depth": [8, 12, 16],
    "render_loop": [0, 1, 2],
    "render_wait_lanes": [0, 4, 8, 16, 24, 32],
    "render_leaf_coop": [0, 1],
    "grid_blocks": [1024, 2048, 4096, 8192],
    "shade_threads": [64, 128, 256, 512, 1024],
    "batch_paths": [1 << 20, 1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26],
}
# the refill threshold before stack depth and leaf-phase wait: their best values depend on it
TRACE_KEYS = ["trace_tree", "trace_loop", "leaf_coop", "refill_lanes", "trace_lds_depth", "ray_sort", "wait_lanes",
              "trace_grid_blocks"]
# loop structure first: the best block and batch sizes depend on which kernel runs
RENDER_KEYS = ["render_loop", "render_leaf_coop", "grid_blocks", "shade_threads", "batch_paths", "render_lds_depth",
               "render_wait_lanes"]


def _time(make_step, reps):
    import torch
    step = make_step()
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune(make_step, keys=None, reps=3, rounds=1, log=None):
    """Coordinate descent. make_step() builds a zero-argument callable that runs one step
    with the current tuning (rebuild anything a tuning bakes in, e.g. a captured graph).
    Returns (best tuning dict, best seconds per step, list of (tuning, seconds))."""
    keys = keys or list(SPACE)
    best = rtx.get_tuning()
    best_t = _time(make_step, reps)
    history = [(dict(best), best_t)]
    for _ in range(rounds):
        for k in keys:
            for v in SPACE[k]:
                if v == best[k]:
                    continue
                cand = dict(best, **{k: v})
                rtx.set_tuning(**cand)
                t = _time(make_step, reps)
                history.append((cand, t))
                if log:
                    log(f"  {k}={v}: {t * 1e3:.3f} ms")
                if t < best_t:
                    best, best_t = cand, t
            rtx.set_tuning(**best)
    return best, best_t, history


def device_name():
    import torch
    return torch.cuda.get_device_name(0)


def load_cached():
    if not os.path.exists(CACHE):
        return {}
    with open(CACHE) as f:
        return json.load(f)


def save(tuning, workload):
    data = load_cached()
    data[device_name()] = {"tuning": tuning, "workload": workload}
    with open(CACHE, "w") as f:
        json.dump(data, f, indent=1)


def apply_cached():
    """Apply the cached tuning for this GPU, if any; returns it (or None)."""
    entry = load_cached().get(device_name())
    if entry:
        rtx.set_tuning(**entry["tuning"])
        return entry["tuning"]
    return None
