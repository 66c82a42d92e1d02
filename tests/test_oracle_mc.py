"""Pins the oracle's marching cubes and interior count (CPU only): the
triangle table by its structure, the extraction against an independent
per-cube restatement, and closed forms (sphere, plane, sealed borders)."""
import math

import numpy as np
import pytest

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def test_table_structure(orc):
    """Every case uses exactly its cut edges; inside the cube every segment is
    shared by two triangles in opposite directions, on a cube face by one; all
    triangles of a case face the same way (towards the inside corners in table
    order)."""
    for c in range(256):
        tri = orc.mc_table(c)
        inside = [(c >> i) & 1 for i in range(8)]
        cut = {e for e, (a, b) in enumerate(EDGE) if inside[a] != inside[b]}
        assert len(tri) % 3 == 0 and set(tri) == cut, c
        segs = {}
        for k in range(0, len(tri), 3):
            a, b, d = tri[k:k + 3]
            assert len({a, b, d}) == 3
            for s in ((a, b), (b, d), (d, a)):
                segs[s] = segs.get(s, 0) + 1

        def faces(e):
            p, q = (CORNER[i] for i in EDGE[e])
            return {(ax, p[ax]) for ax in range(3) if p[ax] == q[ax]}

        for (a, b), n in segs.items():
            if faces(a) & faces(b):
                assert n == 1 and (b, a) not in segs, (c, a, b)
            else:
                assert n == 1 and segs.get((b, a)) == 1, (c, a, b)

        def mid(e):
            p, q = (np.array(CORNER[i], float) for i in EDGE[e])
            return (p + q) / 2

        for k in range(0, len(tri), 3):
            p = [mid(e) for e in tri[k:k + 3]]
            n = np.cross(p[1] - p[0], p[2] - p[0])
            g = sum(np.array(CORNER[EDGE[e][1]], float) - np.array(CORNER[EDGE[e][0]], float)
                    if inside[EDGE[e][0]] else
                    np.array(CORNER[EDGE[e][0]], float) - np.array(CORNER[EDGE[e][1]], float) for e in tri[k:k + 3])
            assert n @ g < 0, (c, tri[k:k + 3])
    assert orc.mc_table(0) == [] and orc.mc_table(255) == []


def per_cube_mc(grid, dims, h, origin, iso):
    """Independent restatement: every cube on its own (no shared vertices),
    float32 arithmetic, triangles as vertex-coordinate triples."""
    dx, dy, dz = dims
    g = np.asarray(grid, np.float32).reshape(dz, dy, dx).copy()
    seal = np.float32(iso) + np.float32(h)
    for sl in (np.s_[0], np.s_[-1]):
        for view in (g[sl, :, :], g[:, sl, :], g[:, :, sl]):
            view[view < seal] = seal
    f32 = np.float32
    out = []
    for z in range(dz - 1):
        for y in range(dy - 1):
            for x in range(dx - 1):
                vals = [g[z + c[2], y + c[1], x + c[0]] - f32(iso) for c in CORNER]
                c = sum(1 << k for k in range(8) if vals[k] <= 0)
                tri = orc_table[c]
                for k in range(0, len(tri), 3):
                    pts = []
                    for e in (tri[k], tri[k + 2], tri[k + 1]):
                        a, b = EDGE[e]
                        lo, hi = (a, b) if CORNER[a] < CORNER[b] else (b, a)  # lattice edge from its lower end
                        v0, v1 = vals[lo], vals[hi]
                        den = f32(v0 - v1)
                        t = (min(max(f32(v0 / den), f32(0.001)), f32(0.999)) if abs(den) > f32(1e-7) else f32(0.5))
                        cl = CORNER[lo]
                        p = [f32(f32(origin[i]) + f32(f32([x, y, z][i] + cl[i]) * f32(h))) for i in range(3)]
                        ax = [i for i in range(3) if CORNER[lo][i] != CORNER[hi][i]][0]
                        p[ax] = f32(p[ax] + f32(t * f32(h)))
                        pts.append(tuple(float(q) for q in p))
                    out.append(pts)
    return out


def canon(tris):
    """Triangles as a sorted list of orientation-preserving canonical rotations."""
    res = []
    for t in tris:
        k = min(range(3), key=lambda i: t[i])
        res.append((t[k], t[(k + 1) % 3], t[(k + 2) % 3]))
    return sorted(res)


orc_table = None


@pytest.fixture(autouse=True)
def _table(orc):
    global orc_table
    orc_table = [orc.mc_table(c) for c in range(256)]


@pytest.mark.parametrize("shape,seed", [((2, 2, 2), 0), ((5, 4, 3), 1), ((9, 7, 6), 2), ((12, 3, 8), 3)])
def test_extraction_equals_per_cube_restatement(orc, shape, seed):
    rng = np.random.default_rng(seed)
    grid = rng.normal(size=int(np.prod(shape))).astype(np.float32)
    h, origin, iso = 0.5, (-1.0, 2.0, 0.25), 0.1
    verts, tris = orc.mc_extract(grid, shape, h, origin, iso)
    mine = canon([[tuple(float(q) for q in verts[i]) for i in t] for t in tris])
    ref = canon(per_cube_mc(grid, shape, h, origin, iso))
    assert mine == ref
    # welding: every vertex is used and no two vertices coincide on one edge
    assert len(np.unique(tris)) == len(verts)


def _sphere(n, h, r):
    c = (n - 1) / 2
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    return (np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) * h - r).astype(np.float32), c


def test_sphere_closed_genus0_volume(orc):
    n, h, r = 40, 0.5, 7.0
    grid, c = _sphere(n, h, r)
    verts, tris = orc.mc_extract(grid, (n, n, n), h, (-c * h,) * 3, 0.0)
    # closed, consistently oriented 2-manifold: every directed edge once, its reverse once
    directed = {}
    for a, b, d in tris:
        for e in ((a, b), (b, d), (d, a)):
            directed[e] = directed.get(e, 0) + 1
    assert all(v == 1 for v in directed.values())
    assert all((b, a) in directed for (a, b) in directed)
    V, E, F = len(verts), len(directed) // 2, len(tris)
    assert V - E + F == 2
    rad = np.linalg.norm(verts.astype(np.float64), axis=1)
    assert np.abs(rad - r).max() < 0.05 * h
    p = verts.astype(np.float64)[tris.astype(np.int64)]
    vol = np.einsum("ij,ij->i", p[:, 0], np.cross(p[:, 1], p[:, 2])).sum() / 6
    assert vol == pytest.approx(4 / 3 * math.pi * r ** 3, rel=0.01)  # outward normals: positive


def test_plane_vertices_exact(orc):
    """A linear field x - 3.3 crosses the x-edges at 3.3 exactly (up to
    rounding); the sealed border closes the slab (x = 0 is raised to h, so a
    second wall sits at x = 1 / 3.3) into a closed surface."""
    n, h = 10, 1.0
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float32),) * 3, indexing="ij")
    grid = (x - np.float32(3.3)).astype(np.float32)
    verts, tris = orc.mc_extract(grid, (n, n, n), h, (0.0, 0.0, 0.0), 0.0)
    inner = (verts[:, 0] > 1) & (verts[:, 1] > 1) & (verts[:, 1] < n - 2) & (verts[:, 2] > 1) & (verts[:, 2] < n - 2)
    assert inner.sum() == (n - 4) ** 2
    assert np.allclose(verts[inner, 0], 3.3, atol=1e-5)
    wall = (verts[:, 0] < 1) & (verts[:, 1] > 1) & (verts[:, 1] < n - 2) & (verts[:, 2] > 1) & (verts[:, 2] < n - 2)
    assert np.allclose(verts[wall, 0], 1 / 3.3, atol=1e-6)
    directed = {}
    for a, b, d in tris:
        for e in ((a, b), (b, d), (d, a)):
            directed[e] = directed.get(e, 0) + 1
    assert all(v == 1 for v in directed.values()) and all((b, a) in directed for (a, b) in directed)


def test_all_inside_is_sealed_box(orc):
    """Everything inside: the seal raises the border to iso + h, so the surface
    is a closed box half a voxel inside the border."""
    n, h = 6, 2.0
    grid = np.full(n ** 3, -5.0, np.float32)
    verts, tris = orc.mc_extract(grid, (n, n, n), h, (0.0, 0.0, 0.0), 0.0)
    assert len(tris) > 0
    # the walls: t = 2 / 7 from a border voxel (2) inward, 5 / 7 from an interior voxel (-5) outward
    f = np.float32
    lo = f(f(0.0) + f(f(2.0) / f(7.0)) * f(h))
    hi = f(f(f(n - 2) * f(h)) + f(f(-5.0) / f(-7.0)) * f(h))
    assert set(np.unique(verts[:, 0])) <= {f(0.0), f(h), f(2 * h), f(3 * h), f(4 * h), f(5 * h), lo, hi}
    assert lo in set(np.unique(verts[:, 0])) and hi in set(np.unique(verts[:, 0]))
    assert orc.mc_extract(grid, (n, n, 1), h, (0, 0, 0), 0.0)[1].shape == (0, 3)


def test_interior_count(orc):
    g = np.array([-1.0, 0.0, 1.0, np.nan, -0.0, 2.0], np.float32)
    assert orc.interior_count(g) == 3
