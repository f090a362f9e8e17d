"""Autotuner of the mixed-order batched elemental launch.

For one batch of element segments of several orders p the tuner builds a
launch plan in every candidate configuration -- one launch per order in
sequence, one launch per order on concurrent streams, or one persistent launch
per kernel over all orders with several grid sizes -- captures each into a
hipGraph, times graph replays with events on the launch stream, and keeps the
fastest.  Decisions are cached per batch shape (the element count of each
order) for the life of the process.
"""
import numpy as np

from .runtime import _torch

PLAN_PER_ORDER, PLAN_FUSED, PLAN_CONCURRENT = 0, 1, 2
MODE_NAMES = {PLAN_PER_ORDER: "per_order", PLAN_FUSED: "fused_persistent", PLAN_CONCURRENT: "concurrent_streams"}

_cache = {}


def candidates(n_cu):
    out = [(PLAN_PER_ORDER, 0), (PLAN_CONCURRENT, 0)]
    out += [(PLAN_FUSED, k * n_cu) for k in (1, 2, 4, 8)]
    return out


def time_plan(ctx, plan, reps=3, warmup=1):
    torch = _torch()
    for _ in range(warmup):
        plan.launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(ctx.stream)
    for _ in range(reps):
        plan.launch()
    e1.record(ctx.stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def shape_key(masters, segments):
    counts = {}
    for m, seg in zip(masters, segments):
        counts[m.porder] = counts.get(m.porder, 0) + int(seg[1].shape[2])
    app = segments[0][0]
    return (app.physics, masters[0].nd, tuple(sorted(counts.items())))


def tune(ctx, masters, segments, reps=3, cands=None, log=None):
    """Returns (best plan, {candidate name: ms}); the plan is captured."""
    torch = _torch()
    n_cu = torch.cuda.get_device_properties(ctx.device).multi_processor_count
    key = shape_key(masters, segments)
    cands = cands or candidates(n_cu)
    if key in _cache:
        mode, grid = _cache[key][0]
        plan = ctx.plan(masters, segments, mode, grid)
        plan.capture()
        return plan, _cache[key][1]
    timings = {}
    best, best_ms = None, np.inf
    for mode, grid in cands:
        plan = ctx.plan(masters, segments, mode, grid)
        plan.capture()
        ms = time_plan(ctx, plan, reps)
        name = MODE_NAMES[mode] + (f"_grid{grid}" if mode == PLAN_FUSED else "")
        timings[name] = ms
        if log:
            log(f"[autotune] {name}: {ms:.3f} ms")
        if ms < best_ms:
            if best is not None:
                best[0].close()
            best, best_ms = (plan, (mode, grid)), ms
        else:
            plan.close()
    _cache[key] = (best[1], timings)
    return best[0], timings
