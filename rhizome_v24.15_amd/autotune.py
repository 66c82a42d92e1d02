"""Autotuner for the batched bake: times candidate launch configurations
(concurrent lanes, sweep load-group length and waves per block) on a sample
of the batch and keeps the fastest.  Every configuration computes identical
results; only the time differs.

    best, table = autotune_bake(rhz, torch, items, params)
"""
from __future__ import annotations

import time

LANES = (1, 2, 4, 8)
SWEEP_SHAPES = ((8, 4), (4, 8), (2, 4))  # (sweep_step, sweep_rows)


def _time_bake(rhz, torch, items, params, n_lanes):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc, _, _ = rhz.bake(items, params, n_lanes)
    torch.cuda.synchronize()
    if rc != rhz.RHZ_OK:
        raise RuntimeError(f"bake failed during tuning ({rc})")
    return (time.perf_counter() - t0) * 1e3


def autotune_bake(rhz, torch, items, params, lanes=LANES, sweep_shapes=SWEEP_SHAPES, sample=16, log=None):
    """Returns ({"n_lanes", "sweep_step", "sweep_rows", "ms"}, [timings]) and
    leaves the library's sweep tuning set to the winner."""
    sub = list(items[:sample])
    table = []
    _time_bake(rhz, torch, sub, params, max(lanes))  # warm the lanes' workspaces
    for step, rows in sweep_shapes:
        rhz.tuning_set("sweep_step", step)
        rhz.tuning_set("sweep_rows", rows)
        for nl in lanes:
            ms = _time_bake(rhz, torch, sub, params, nl)
            table.append({"n_lanes": nl, "sweep_step": step, "sweep_rows": rows, "ms": ms})
            if log:
                log(f"autotune: lanes {nl} step {step} rows {rows}: {ms:.1f} ms")
    best = min(table, key=lambda r: r["ms"])
    rhz.tuning_set("sweep_step", best["sweep_step"])
    rhz.tuning_set("sweep_rows", best["sweep_rows"])
    return dict(best), table
