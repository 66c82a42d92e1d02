"""rhz_orient_mesh / rhz_signed_volume (the reference's orient_mesh_gpu and
compute_signed_volume_gpu) against the oracle: identical indices and flags on
orientable meshes, identical flags on a Moebius band, volumes to binary64
rounding."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _flip(f, rows):
    g = np.array(f, dtype=np.uint32, copy=True)
    g[rows] = g[rows][:, [0, 2, 1]]
    return g


def _meshes(geom):
    v, f = geom.icosphere(4)
    v = (v * 20.0).astype(np.float32)
    rng = np.random.default_rng(4)
    out = {
        "seeded_flips": (v, _flip(f, rng.choice(np.arange(1, len(f)), len(f) // 10, replace=False))),
        "first_flipped": (v, _flip(f, np.concatenate([[0], rng.choice(np.arange(1, len(f)), 200, replace=False)]))),
    }
    s, sf = geom.uv_sphere(15.0, 48, 48)
    out["inside_out_sphere"] = (s, sf[:, ::-1].copy())
    vi = (v * 0.5).astype(np.float32)
    out["nested_inside_out_shell"] = (np.concatenate([v, vi]), np.concatenate([f, f[:, ::-1] + len(v)]).astype(np.uint32))
    vv, fv = geom.vessel()
    out["vessel_1pct_flips"] = (vv, _flip(fv, rng.choice(np.arange(1, len(fv)), len(fv) // 100, replace=False)))
    return out


@pytest.mark.parametrize("name", ["seeded_flips", "first_flipped", "inside_out_sphere", "nested_inside_out_shell",
                                  "vessel_1pct_flips"])
def test_orient_matches_oracle(rhz, geom, orc, torch_gpu, name):
    v, f = _meshes(geom)[name]
    got = rhz.orient_mesh(v, f)
    ref = orc.orient_mesh(v, f)
    assert got[1:4] == ref[1:4], (got[1:4], ref[1:4])
    assert got[1]  # orientable
    assert np.array_equal(got[0], ref[0])
    assert got[4] == pytest.approx(ref[4], rel=1e-12)
    assert got[4] > 0
    assert rhz.signed_volume(v, got[0]) == pytest.approx(got[4], rel=1e-12)


def test_orient_moebius_band(rhz, geom, orc, torch_gpu):
    v, f = geom.moebius()
    got = rhz.orient_mesh(v, f)
    ref = orc.orient_mesh(v, f)
    assert got[1:3] == ref[1:3] == (False, True)


def test_orient_mug_specimen_is_consistent(rhz, orc, torch_gpu):
    import __graft_entry__ as entry

    v, f = entry.load_obj().load_obj(os.path.join(HERE, "golden", "mug.obj"))
    got = rhz.orient_mesh(v, f)
    assert got[1:4] == orc.orient_mesh(v, f)[1:4]
    assert got[1] and not got[2]
    assert abs(got[4] - 75913) < 1.0


def test_orient_empty_and_bad_index(rhz, torch_gpu):
    out = rhz.orient_mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    assert out[1:] == (True, False, False, 0.0)
    with pytest.raises(rhz.RhzError):
        rhz.orient_mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 3]], np.uint32))
