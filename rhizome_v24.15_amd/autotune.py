"""Autotuner for the batched bake: times candidate launch configurations on a
sample of the batch and keeps the fastest.  Every configuration computes
identical results; only the time differs.

Searched, coordinate-wise: the bake mode (the eikonal rounds of all
lattices as batched launches, or each item's whole build on a lane); the
number of concurrent lanes; the sweep kernel
shape (the tile per sweep orientation, steps whose loads are issued
together); whether the
FSM rounds of small lattices are enqueued in one go and replayed as a
hipGraph.

    best, table = autotune_bake(rhz, torch, items, params)
"""
from __future__ import annotations

import time

BAKE_MODES = (1, 0)  # bake_batched
LANES = (1, 2, 4, 8)
# (sweep_shape_x, sweep_shape_y, sweep_shape_z, sweep_step)
SWEEP_SHAPES = ((14, 21, 21, 2), (14, 21, 22, 2), (12, 12, 14, 2), (1, 1, 1, 8), (14, 1, 21, 8))
FSM_MODES = ((16 << 20, 0), (16 << 20, 1), (0, 0))  # (fsm_batch_rounds_max_voxels, fsm_graph)


def _time_bake(rhz, torch, items, params, n_lanes):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc, _, _ = rhz.bake(items, params, n_lanes)
    torch.cuda.synchronize()
    if rc != rhz.RHZ_OK:
        raise RuntimeError(f"bake failed during tuning ({rc})")
    return (time.perf_counter() - t0) * 1e3


def apply_config(rhz, cfg):
    """Sets the library's tuning keys from a configuration (its n_lanes is the
    caller's to pass to bake)."""
    for k in ("sweep_shape_x", "sweep_shape_y", "sweep_shape_z", "sweep_step"):
        rhz.tuning_set(k, cfg[k])
    rhz.tuning_set("fsm_batch_rounds_max_voxels", cfg["fsm_batch_rounds_max_voxels"])
    rhz.tuning_set("fsm_graph", cfg["fsm_graph"])
    if "bake_batched" in cfg:
        rhz.tuning_set("bake_batched", cfg["bake_batched"])


def autotune_bake(rhz, torch, items, params, sample=16, log=None):
    """Returns (best config with its time, [all timings]) and leaves the
    library tuned to the winner."""
    sub = list(items[:sample])
    cfg = {"n_lanes": 4, **{k: rhz.tuning_get(k) for k in ("sweep_shape_x", "sweep_shape_y", "sweep_shape_z",
                                                           "sweep_step")},
           "fsm_batch_rounds_max_voxels": rhz.tuning_get("fsm_batch_rounds_max_voxels"),
           "fsm_graph": rhz.tuning_get("fsm_graph"), "bake_batched": rhz.tuning_get("bake_batched")}
    table = []

    def trial(c):
        apply_config(rhz, c)
        ms = _time_bake(rhz, torch, sub, params, c["n_lanes"])
        row = dict(c, ms=ms)
        table.append(row)
        if log:
            log(f"autotune: {row}")
        return ms

    apply_config(rhz, cfg)
    _time_bake(rhz, torch, sub, params, max(LANES))  # warm the lanes' workspaces
    best_ms = None
    for mode in BAKE_MODES:
        c = dict(cfg, bake_batched=mode)
        ms = trial(c)
        if best_ms is None or ms < best_ms:
            best_ms, cfg = ms, c
    for nl in LANES:
        c = dict(cfg, n_lanes=nl)
        ms = trial(c)
        if ms < best_ms:
            best_ms, cfg = ms, c
    for sx, sy, sz, step in SWEEP_SHAPES:
        c = dict(cfg, sweep_shape_x=sx, sweep_shape_y=sy, sweep_shape_z=sz, sweep_step=step)
        ms = trial(c)
        if ms < best_ms:
            best_ms, cfg = ms, c
    for mv, g in FSM_MODES:
        c = dict(cfg, fsm_batch_rounds_max_voxels=mv, fsm_graph=g)
        ms = trial(c)
        if ms < best_ms:
            best_ms, cfg = ms, c
    apply_config(rhz, cfg)
    return dict(cfg, ms=best_ms), table
