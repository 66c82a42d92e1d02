"""Per-class sweep shapes change no bit of a field: single builds on lattices
of three size classes, and a batched bake whose lattices span two classes,
each class with its own tile shapes (single-build and batched-stage keys);
and the autotuner's table round trip."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# per class: x, y, z tile shapes (1 = the one-line walker) and load-group depth
EXOTIC = {0: (1, 22, 12, 8), 1: (12, 1, 1, 4), 2: (22, 12, 14, 8)}


@pytest.fixture()
def tune(rhz):
    import __graft_entry__ as entry

    t = importlib.import_module(entry.PKG_NAME + ".autotune")
    t.clear_sweep_table(rhz)
    yield t
    t.clear_sweep_table(rhz)


def _set_exotic(rhz):
    for c, (sx, sy, sz, step) in EXOTIC.items():
        for pre in ("sweep", "batch"):
            rhz.tuning_set(f"{pre}_shape_x@{c}", sx)
            rhz.tuning_set(f"{pre}_shape_y@{c}", sy)
            rhz.tuning_set(f"{pre}_shape_z@{c}", sz)
            rhz.tuning_set(f"{pre}_step@{c}", step)


@pytest.mark.parametrize("h", [0.9, 0.5, 0.25])
def test_single_build_unchanged_by_class_shapes(rhz, geom, torch_gpu, tune, h):
    v, f = geom.vessel()
    tris = rhz.build_triangles(v, f)
    mn, mx = geom.frame_grid(v, h)
    cls = rhz.sweep_shape_class([mx[i] - mn[i] + 1 for i in range(3)])
    assert cls in EXOTIC
    p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED)
    rc, ref, st0 = rhz.compute_sdf(tris, h, mn, mx, p)
    assert rc == rhz.RHZ_OK
    _set_exotic(rhz)
    rc, got, st1 = rhz.compute_sdf(tris, h, mn, mx, p)
    assert rc == rhz.RHZ_OK
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert st1.eikonal_rounds == st0.eikonal_rounds
    print("class", cls, "sweeps ms", round(st0.ms_sweeps, 3), "->", round(st1.ms_sweeps, 3))


def test_mixed_class_bake_matches_single_builds(rhz, geom, torch_gpu, tune):
    torch = torch_gpu
    items, refs = [], []
    p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED)
    for k, (seed, h) in enumerate([(3, 0.9), (4, 0.5), (5, 0.25), (6, 0.7), (7, 0.45)]):
        v, f = geom.random_vessel(seed, 20_000)
        tris = rhz.build_triangles(v, f)
        mn, mx = geom.frame_grid(v, h)
        dims = [mx[i] - mn[i] + 1 for i in range(3)]
        n = int(np.prod(dims))
        d_t = torch.from_numpy(tris).cuda()
        d_o = torch.empty(n, dtype=torch.float32, device="cuda")
        rc, st = rhz.compute_sdf_device(d_t, len(tris), h, mn, mx, d_o, p)
        assert rc == rhz.RHZ_OK
        refs.append(d_o.cpu().numpy().copy())
        items.append((d_t, len(tris), h, mn, mx, torch.empty(n, dtype=torch.float32, device="cuda")))
    classes = sorted({rhz.sweep_shape_class([it[4][i] - it[3][i] + 1 for i in range(3)]) for it in items})
    print("classes", classes)
    assert len(classes) >= 2
    _set_exotic(rhz)
    saved = rhz.tuning_get("bake_batched")
    try:
        for mode in (1, 0):
            rhz.tuning_set("bake_batched", mode)
            for it in items:
                it[5].fill_(7.0)
            torch.cuda.synchronize()
            rc, stats, _ = rhz.bake(items, p, 3)
            torch.cuda.synchronize()
            assert rc == rhz.RHZ_OK
            for it, ref in zip(items, refs):
                assert np.array_equal(it[5].cpu().numpy().view(np.uint32), ref.view(np.uint32)), mode
    finally:
        rhz.tuning_set("bake_batched", saved)


def test_tune_cache_and_reload(rhz, geom, torch_gpu, tune, tmp_path, monkeypatch):
    torch = torch_gpu
    # two classes, one timing each, to keep the test short
    real = tune.tune_sweep_table
    monkeypatch.setattr(tune, "tune_sweep_table",
                        lambda rhz_, torch_, geom_, log=None: real(rhz_, torch_, geom_, classes=(0, 1), reps=1, log=log))
    path = str(tmp_path / "sweep_table.json")
    table, cached = tune.ensure_sweep_table(rhz, torch, geom, path=path, log=print)
    assert not cached and set(table["classes"]) == {"0", "1"}
    for c, cfg in table["classes"].items():
        assert cfg["sweep_shape_x"] in tune.TILE_SHAPES and cfg["sweep_step"] in tune.LOAD_STEPS
        assert rhz.tuning_get(f"sweep_step@{c}") == cfg["sweep_step"]
    tune.clear_sweep_table(rhz)
    again, cached = tune.ensure_sweep_table(rhz, torch, geom, path=path)
    assert cached and again["classes"] == table["classes"]
    assert rhz.tuning_get("sweep_shape_y@1") == table["classes"]["1"]["sweep_shape_y"]
