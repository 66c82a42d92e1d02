"""Autotuners.  Every configuration computes identical results; only the
time differs.

Per lattice shape (``tune_sweep_table`` / ``ensure_sweep_table``): for each
lattice-size class of the sweep tuning (``rhz.sweep_shape_class``: up to
2^21, 2^24, 2^27 voxels, larger) a build of the vessel mesh on a lattice of
that class is timed (the builds' own sweep time, ``Stats.ms_sweeps``) for
every tile shape of each sweep orientation and every load-group depth,
coordinate-wise; the winners go to the per-class tuning keys
("sweep_shape_x@C", ...), which the single builds and the batched bake's
per-class sweep launches read.  The table is cached as JSON per device.

For the batched bake (``autotune_bake``): times candidate launch
configurations on a sample of the batch and keeps the fastest, including the
batched sweeps' own shape per size class present ("batch_shape_x@C", ...):
they are a stage of their own, sharing the GPU with the sign stages.

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

import json
import math
import os
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


# per size class of the batched bake's sweeps: (x, y, z shapes, step) tried
# against the class-free shapes (0, 0, 0, 0)
BATCH_CLASS_CHOICES = ((1, 1, 1, 8), (12, 12, 14, 2))
BATCH_KEYS = ("batch_shape_x", "batch_shape_y", "batch_shape_z", "batch_step")


def set_batch_class(rhz, cls, choice):
    for k, v in zip(BATCH_KEYS, choice):
        rhz.tuning_set(f"{k}@{cls}", v)


def apply_config(rhz, cfg):
    """Sets the library's tuning keys from a configuration (its n_lanes is the
    caller's to pass to bake)."""
    for k in ("sweep_shape_x", "sweep_shape_y", "sweep_shape_z", "sweep_step"):
        rhz.tuning_set(k, cfg[k])
    rhz.tuning_set("fsm_batch_rounds_max_voxels", cfg["fsm_batch_rounds_max_voxels"])
    rhz.tuning_set("fsm_graph", cfg["fsm_graph"])
    if "bake_batched" in cfg:
        rhz.tuning_set("bake_batched", cfg["bake_batched"])
    for c in range(4):
        set_batch_class(rhz, c, tuple(cfg.get("batch_classes", {}).get(str(c), (0, 0, 0, 0))))


def autotune_bake(rhz, torch, items, params, sample=16, log=None, shapes=True):
    """Returns (best config with its time, [all timings]) and leaves the
    library tuned to the winner.  shapes=False leaves the sweep shapes alone
    (a per-class table is in effect)."""
    sub = list(items[:sample])
    cfg = {"n_lanes": 4, **{k: rhz.tuning_get(k) for k in ("sweep_shape_x", "sweep_shape_y", "sweep_shape_z",
                                                           "sweep_step")},
           "fsm_batch_rounds_max_voxels": rhz.tuning_get("fsm_batch_rounds_max_voxels"),
           "fsm_graph": rhz.tuning_get("fsm_graph"), "bake_batched": rhz.tuning_get("bake_batched"),
           "batch_classes": {}}
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
    for sx, sy, sz, step in (SWEEP_SHAPES if shapes else ()):
        c = dict(cfg, sweep_shape_x=sx, sweep_shape_y=sy, sweep_shape_z=sz, sweep_step=step)
        ms = trial(c)
        if ms < best_ms:
            best_ms, cfg = ms, c
    for mv, g in FSM_MODES:
        c = dict(cfg, fsm_batch_rounds_max_voxels=mv, fsm_graph=g)
        ms = trial(c)
        if ms < best_ms:
            best_ms, cfg = ms, c
    if shapes and cfg["bake_batched"]:
        # the batched sweeps' own shape per size class present in the batch
        classes = sorted({rhz.sweep_shape_class([it[4][i] - it[3][i] + 1 for i in range(3)]) for it in sub})
        for cls in classes:
            for choice in BATCH_CLASS_CHOICES:
                c = dict(cfg, batch_classes=dict(cfg["batch_classes"], **{str(cls): choice}))
                ms = trial(c)
                if ms < best_ms:
                    best_ms, cfg = ms, c
    apply_config(rhz, cfg)
    return dict(cfg, ms=best_ms), table


# ---------------------------------------------------------------------------
# per-class sweep shapes
# ---------------------------------------------------------------------------
SHAPE_KEYS = ("sweep_shape_x", "sweep_shape_y", "sweep_shape_z")
TILE_SHAPES = (14, 21, 22, 12, 1)  # 10 NF + NC tiles, and 1 = the one-line walker
KEEP_MARGIN = 0.005  # relative: 0.5% of a build's sweep time
LOAD_STEPS = (2, 4, 8)
# vessel lattice spacings (mm) whose lattices fall in classes 0..3
CLASS_SPACINGS = {0: 0.9, 1: 0.5, 2: 0.25, 3: 0.10}  # 1.6M, 7.2M, 48M, 671M voxels
TABLE_VERSION = 1


def default_table_path() -> str:
    return os.path.join(os.path.expanduser("~"), ".cache", "rhizome_v24.15_amd", "sweep_table.json")


def device_key(torch) -> str:
    """Architecture and compute-unit count of the current device."""
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{getattr(p, 'gcnArchName', p.name)} cu{p.multi_processor_count}"


def apply_sweep_table(rhz, table) -> None:
    """Sets the per-class tuning keys from a table (``tune_sweep_table``)."""
    for c, cfg in table["classes"].items():
        for k in SHAPE_KEYS + ("sweep_step",):
            rhz.tuning_set(f"{k}@{int(c)}", int(cfg[k]))


def clear_sweep_table(rhz) -> None:
    """Every class back to the class-free keys (single builds and the batched
    bake's sweeps)."""
    for c in range(4):
        for k in SHAPE_KEYS + ("sweep_step",) + BATCH_KEYS:
            rhz.tuning_set(f"{k}@{c}", 0)


def _class_case(rhz, geom, torch, cls):
    v, f = geom.vessel()
    tris = rhz.build_triangles(v, f)
    h = CLASS_SPACINGS[cls]
    mn, mx = geom.frame_grid(v, h)
    dims = [int(mx[i] - mn[i] + 1) for i in range(3)]
    if rhz.sweep_shape_class(dims) != cls:
        raise RuntimeError(f"lattice {dims} at h = {h} is not in size class {cls}")
    d_tris = torch.from_numpy(tris).cuda()
    d_out = torch.empty(math.prod(dims), dtype=torch.float32, device="cuda")
    return d_tris, len(tris), h, mn, mx, d_out, dims


def _sweeps_ms(rhz, torch, case, params, reps):
    d_tris, nt, h, mn, mx, d_out, _ = case
    best = math.inf
    for _ in range(reps):
        rc, st = rhz.compute_sdf_device(d_tris, nt, h, mn, mx, d_out, params)
        torch.cuda.synchronize()
        if rc != rhz.RHZ_OK:
            raise RuntimeError(f"build failed during tuning ({rc})")
        best = min(best, st.ms_sweeps)
    return best


def tune_sweep_class(rhz, torch, geom, cls, params=None, reps=3, log=None):
    """Coordinate search over the tile shape of each orientation, then the
    load-group depth, for one size class.  Returns (config, trials)."""
    p = params if params is not None else rhz.default_params(method=rhz.METHOD_FAST_LEARNED)
    case = _class_case(rhz, geom, torch, cls)
    cfg = {k: rhz.tuning_get(k) for k in SHAPE_KEYS + ("sweep_step",)}
    trials = []

    def trial(c):
        for k, v in c.items():
            rhz.tuning_set(f"{k}@{cls}", v)
        ms = _sweeps_ms(rhz, torch, case, p, reps)
        trials.append(dict(c, cls=cls, ms=ms))
        if log:
            log(f"sweep table: class {cls} {c} sweeps {ms:.3f} ms")
        return ms

    _sweeps_ms(rhz, torch, case, p, 1)  # warm the workspace
    best = trial(cfg)
    # a candidate replaces the incumbent only when faster by more than the
    # spread between repeated timings
    for k in SHAPE_KEYS:
        for shape in TILE_SHAPES:
            if shape == cfg[k]:
                continue
            c = dict(cfg, **{k: shape})
            ms = trial(c)
            if ms < best * (1.0 - KEEP_MARGIN):
                best, cfg = ms, c
    for step in LOAD_STEPS:
        if step == cfg["sweep_step"]:
            continue
        c = dict(cfg, sweep_step=step)
        ms = trial(c)
        if ms < best * (1.0 - KEEP_MARGIN):
            best, cfg = ms, c
    for k, v in cfg.items():
        rhz.tuning_set(f"{k}@{cls}", v)
    return dict(cfg, ms=best, dims=case[6], h=case[2]), trials


def tune_sweep_table(rhz, torch, geom, classes=(0, 1, 2, 3), reps=3, log=None):
    """Tunes every class; returns the table (also applied)."""
    saved = {k: rhz.tuning_get(k) for k in ("verify_overlap", "fsm_graph")}
    rhz.tuning_set("verify_overlap", 0)  # the side stream would share the GPU with the timed sweeps
    rhz.tuning_set("fsm_graph", 0)  # captured rounds record no per-sweep events
    table = {"version": TABLE_VERSION, "device": device_key(torch), "workload": "vessel mesh, learned method",
             "classes": {}, "trials": []}
    try:
        for cls in classes:
            cfg, trials = tune_sweep_class(rhz, torch, geom, cls, reps=reps, log=log)
            table["classes"][str(cls)] = cfg
            table["trials"] += trials
    finally:
        for k, v in saved.items():
            rhz.tuning_set(k, v)
    return table


def load_sweep_table(path, torch=None):
    """The cached table, or None when absent, unreadable, of another version
    or (with torch) of another device."""
    try:
        with open(path) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        return None
    if not isinstance(table, dict) or table.get("version") != TABLE_VERSION or "classes" not in table:
        return None
    if torch is not None and table.get("device") != device_key(torch):
        return None
    for cfg in table["classes"].values():
        if not all(k in cfg for k in SHAPE_KEYS + ("sweep_step",)):
            return None
    return table


def ensure_sweep_table(rhz, torch, geom, path=None, retune=False, log=None):
    """Applies the cached per-class table of this device, tuning and writing
    it first when there is none (or ``retune``).  Returns (table, cached)."""
    path = path or default_table_path()
    table = None if retune else load_sweep_table(path, torch)
    cached = table is not None
    if table is None:
        table = tune_sweep_table(rhz, torch, geom, log=log)
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(table, fh, indent=1)
        os.replace(tmp, path)
    apply_sweep_table(rhz, table)
    return table, cached
