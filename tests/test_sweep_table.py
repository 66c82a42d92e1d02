"""Per-class sweep tuning keys and the autotuner's cached table (host side:
no kernel runs)."""
import json

import pytest


@pytest.fixture()
def tune(rhz):
    import importlib

    import __graft_entry__ as entry

    t = importlib.import_module(entry.PKG_NAME + ".autotune")
    yield t
    t.clear_sweep_table(rhz)


def test_size_classes(rhz):
    assert [rhz.sweep_shape_class(d) for d in [(128, 128, 128), (129, 128, 128), (256, 256, 256),
                                                (512, 512, 512), (513, 512, 512), (883, 823, 923)]] == [0, 1, 1, 2, 3, 3]
    with pytest.raises(rhz.RhzError):
        rhz.sweep_shape_class((0, 4, 4))


def test_class_keys(rhz, tune):
    base = rhz.tuning_get("sweep_shape_x")
    rhz.tuning_set("sweep_shape_x@2", 22)
    rhz.tuning_set("sweep_step@1", 8)
    assert rhz.tuning_get("sweep_shape_x@2") == 22 and rhz.tuning_get("sweep_step@1") == 8
    assert rhz.tuning_get("sweep_shape_x") == base  # the class-free key is untouched
    assert rhz.tuning_get("sweep_shape_y@0") == 0  # unset: follows the class-free key
    rhz.tuning_set("batch_shape_z@3", 1)
    rhz.tuning_set("batch_step@3", 8)
    assert rhz.tuning_get("batch_shape_z@3") == 1 and rhz.tuning_get("sweep_shape_z@3") == 0
    for bad, v in [("sweep_shape_x@4", 21), ("sweep_shape_x@", 21), ("sweep_step@1x", 2), ("fsm_graph@1", 1),
                   ("sweep_step@1", 3), ("sweep_shape_z@0", 13)]:
        with pytest.raises(rhz.RhzError):
            rhz.tuning_set(bad, v)
    with pytest.raises(rhz.RhzError):
        rhz.tuning_set("batch_step", 2)  # the batched stage has only per-class keys
    tune.clear_sweep_table(rhz)
    assert rhz.tuning_get("sweep_shape_x@2") == 0 and rhz.tuning_get("sweep_step@1") == 0
    assert rhz.tuning_get("batch_shape_z@3") == 0 and rhz.tuning_get("batch_step@3") == 0


def _table():
    return {"version": 1, "device": "gfx950 cu256", "workload": "test",
            "classes": {"0": {"sweep_shape_x": 12, "sweep_shape_y": 22, "sweep_shape_z": 1, "sweep_step": 4},
                        "3": {"sweep_shape_x": 14, "sweep_shape_y": 21, "sweep_shape_z": 21, "sweep_step": 2}},
            "trials": []}


def test_table_load_and_apply(rhz, tune, tmp_path):
    path = tmp_path / "t.json"
    path.write_text(json.dumps(_table()))
    t = tune.load_sweep_table(str(path))
    assert t is not None
    tune.apply_sweep_table(rhz, t)
    assert [rhz.tuning_get(f"{k}@0") for k in ("sweep_shape_x", "sweep_shape_y", "sweep_shape_z", "sweep_step")] \
        == [12, 22, 1, 4]
    assert rhz.tuning_get("sweep_step@3") == 2 and rhz.tuning_get("sweep_shape_x@1") == 0
    # unusable tables are ignored
    bad = _table()
    bad["version"] = 0
    path.write_text(json.dumps(bad))
    assert tune.load_sweep_table(str(path)) is None
    bad = _table()
    del bad["classes"]["0"]["sweep_step"]
    path.write_text(json.dumps(bad))
    assert tune.load_sweep_table(str(path)) is None
    path.write_text("{not json")
    assert tune.load_sweep_table(str(path)) is None
    assert tune.load_sweep_table(str(tmp_path / "absent.json")) is None
