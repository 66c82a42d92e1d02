"""Marching cubes and the interior-voxel count on the GPU against the oracle:
bit-exact vertex and triangle arrays (the order is deterministic on both
sides), through the Python binding, the device entry point and the C++
drop-in; the surface of a built SDF is a closed manifold whose volume agrees
with the interior count."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cxx", "dropin_main")


def _same(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("shape,seed", [((2, 2, 2), 0), ((5, 4, 3), 1), ((17, 9, 33), 2), ((70, 65, 40), 3)])
def test_iso_surface_random_fields_bit_exact(rhz, orc, shape, seed):
    rng = np.random.default_rng(seed)
    grid = rng.normal(size=int(np.prod(shape))).astype(np.float32)
    h, origin, iso = 0.5, (-1.0, 2.0, 0.25), 0.1
    v, t = rhz.iso_surface(grid, shape, h, origin, iso)
    rv, rt = orc.mc_extract(grid, shape, h, origin, iso)
    assert (len(t) > 0) == (min(shape) > 2)  # a 2-cell axis is all border, sealed outside
    assert _same(v, rv) and _same(t, rt)


def test_iso_surface_sphere_nan_and_flat_values(rhz, orc):
    """A sphere field with NaN voxels and exact zeros (v0 - v1 below 1e-7)."""
    n, h, r = 64, 0.25, 5.0
    c = (n - 1) / 2
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    grid = (np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) * h - r).astype(np.float32).reshape(-1)
    rng = np.random.default_rng(5)
    grid[rng.choice(grid.size, 40, replace=False)] = np.nan
    grid[rng.choice(grid.size, 400, replace=False)] = 0.0
    origin = (-c * h,) * 3
    v, t = rhz.iso_surface(grid, (n, n, n), h, origin, 0.0)
    rv, rt = orc.mc_extract(grid, (n, n, n), h, origin, 0.0)
    assert _same(v, rv) and _same(t, rt)


def test_iso_surface_empty_cases(rhz):
    assert rhz.iso_surface(np.zeros(8, np.float32), (8, 1, 1), 1.0, (0, 0, 0))[0].shape == (0, 3)
    v, t = rhz.iso_surface(np.full(5 ** 3, 3.0, np.float32), (5, 5, 5), 1.0, (0, 0, 0))
    assert v.shape == (0, 3) and t.shape == (0, 3)


def test_iso_surface_device_and_volume_of_built_sdf(rhz, orc, geom, torch_gpu):
    """The learned build of a vessel at 0.5 mm, extracted on the device: equal
    to the oracle on the same field, a closed oriented manifold, and its
    enclosed volume within 2% of (interior voxel count) h^3."""
    torch = torch_gpu
    h = 0.5
    v, f = geom.vessel()
    mn, mx = geom.frame_grid(v, h)
    tris = rhz.build_triangles(v, f)
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    V = int(np.prod(dims))
    d_tris = torch.from_numpy(tris).to("cuda")
    d_out = torch.empty(V, dtype=torch.float32, device="cuda")
    rc, st = rhz.compute_sdf_device(d_tris, len(tris), h, mn, mx, d_out, method=rhz.METHOD_FAST_LEARNED)
    assert rc == rhz.RHZ_OK
    origin = [mn[i] * h for i in range(3)]
    nv, nt = rhz.iso_surface_device(d_out, dims, h, origin, 0.0, None, 0, None, 0)
    d_v = torch.empty(3 * nv, dtype=torch.float32, device="cuda")
    d_t = torch.empty(3 * nt, dtype=torch.int32, device="cuda")
    assert rhz.iso_surface_device(d_out, dims, h, origin, 0.0, d_v, nv, d_t, nt) == (nv, nt)
    torch.cuda.synchronize()
    field = d_out.cpu().numpy()
    verts = d_v.cpu().numpy().reshape(-1, 3)
    tri = d_t.cpu().numpy().view(np.uint32).reshape(-1, 3)
    rv, rt = orc.mc_extract(field, dims, h, origin, 0.0)
    assert _same(verts, rv) and _same(tri, rt)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]).astype(np.int64)
    key = e[:, 0] * (nv + 1) + e[:, 1]
    rkey = e[:, 1] * (nv + 1) + e[:, 0]
    assert len(np.unique(key)) == len(key)                 # every directed edge once
    assert np.isin(rkey, key).all()                        # and its reverse too
    p = verts.astype(np.float64)[tri.astype(np.int64)]
    vol = np.einsum("ij,ij->i", p[:, 0], np.cross(p[:, 1], p[:, 2])).sum() / 6
    count = rhz.interior_count_device(d_out, V)
    assert count == orc.interior_count(field)
    assert vol == pytest.approx(count * h ** 3, rel=0.02)


def test_interior_count_matches_oracle(rhz, orc, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(8)
    g = rng.normal(size=5_000_003).astype(np.float32)
    g[rng.choice(g.size, 1000, replace=False)] = np.nan
    g[:7] = np.array([0.0, -0.0, 1e-45, -1e-45, np.inf, -np.inf, 0.5], np.float32)
    want = orc.interior_count(g)
    assert rhz.interior_count(g) == want
    assert rhz.interior_count_device(torch.from_numpy(g).to("cuda"), g.size) == want
    assert rhz.interior_count(g[:0]) == 0


def _grid_file(path, grid, dims, origin, h, iso):
    with open(path, "wb") as fh:
        fh.write(np.array(dims, np.int32).tobytes())
        fh.write(np.array(origin, np.float32).tobytes())
        fh.write(np.array([h, iso], np.float32).tobytes())
        fh.write(np.ascontiguousarray(grid, np.float32).tobytes())


def test_dropin_extract_and_volume_match_binding(rhz, tmp_path):
    assert os.path.exists(BIN), "tests/cxx/dropin_main missing: run __graft_entry__.build()"
    shape, h, origin, iso = (33, 21, 17), 0.5, (1.0, -2.0, 3.0), 0.05
    grid = np.random.default_rng(9).normal(size=int(np.prod(shape))).astype(np.float32)
    gf = tmp_path / "g.bin"
    _grid_file(gf, grid, shape, origin, h, iso)
    out = tmp_path / "m.bin"
    r = subprocess.run([BIN, "mc", str(gf), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    nv, nt = np.frombuffer(raw[:16], np.int64)
    verts = np.frombuffer(raw[16:16 + 12 * nv], np.float32).reshape(-1, 3)
    tri = np.frombuffer(raw[16 + 12 * nv:], np.uint32).reshape(-1, 3)
    v, t = rhz.iso_surface(grid, shape, h, origin, iso)
    assert _same(verts, v) and _same(tri, t)
    r = subprocess.run([BIN, "vol", str(gf), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vol = np.fromfile(out, np.float32)[0]
    assert vol == np.float32(np.float32(rhz.interior_count(grid)) * np.float32(np.float32(h * h) * h))
