"""Pins the oracle's L0 math (point-triangle distance, solid angle, Godunov)
with closed forms, brute force and known answers (CPU only)."""
import math

import numpy as np
import pytest


def f32(*a):
    return np.array(a, dtype=np.float32)


def f64(*a):
    return np.array(a, dtype=np.float64)


def brute_dist2(p, a, b, c, n=400):
    """Min over a dense barycentric grid plus the three edges sampled finely
    (FP64): an upper bound that converges to the exact distance."""
    p, a, b, c = (np.asarray(x, np.float64) for x in (p, a, b, c))
    u = np.linspace(0, 1, n)
    U, W = np.meshgrid(u, u, indexing="ij")
    m = U + W <= 1
    pts = a + U[m, None] * (b - a) + W[m, None] * (c - a)
    best = ((pts - p) ** 2).sum(1).min()
    t = np.linspace(0, 1, 20001)[:, None]
    for q0, q1 in ((a, b), (b, c), (c, a)):
        e = q0 + t * (q1 - q0)
        best = min(best, ((e - p) ** 2).sum(1).min())
    return best


def test_distance_regions_closed_form(orc):
    L = orc.lib()
    a, b, c = f32(0, 0, 0), f32(4, 0, 0), f32(0, 3, 0)
    # face region: plane distance
    assert L.orc_point_triangle_distance_sq(f32(1, 1, 2.5), a, b, c) == pytest.approx(6.25, rel=1e-6)
    # vertex regions
    assert L.orc_point_triangle_distance_sq(f32(-1, -2, 0), a, b, c) == pytest.approx(5.0, rel=1e-6)
    assert L.orc_point_triangle_distance_sq(f32(6, -1, 1), a, b, c) == pytest.approx(4 + 1 + 1, rel=1e-6)
    assert L.orc_point_triangle_distance_sq(f32(-1, 5, 0), a, b, c) == pytest.approx(1 + 4, rel=1e-6)
    # edge regions: distance to the segment
    assert L.orc_point_triangle_distance_sq(f32(2, -3, 4), a, b, c) == pytest.approx(9 + 16, rel=1e-6)
    assert L.orc_point_triangle_distance_sq(f32(-2, 1, 1), a, b, c) == pytest.approx(4 + 1, rel=1e-6)
    # hypotenuse: line 3x + 4y = 12, point (4, 3, 0) -> (12+12-12)/5 = 2.4
    assert L.orc_point_triangle_distance_sq(f32(4, 3, 0), a, b, c) == pytest.approx(2.4 ** 2, rel=1e-6)


def test_distance_vs_brute_force(orc):
    L = orc.lib()
    rng = np.random.default_rng(0)
    for _ in range(60):
        a, b, c = (rng.normal(size=3) for _ in range(3))
        p = rng.normal(size=3) * 2
        got = L.orc_point_triangle_distance_sq_f64(*(np.ascontiguousarray(x, np.float64) for x in (p, a, b, c)))
        bf = brute_dist2(p, a, b, c)
        assert got <= bf + 1e-12
        assert got >= bf - 2e-3 * max(bf, 1e-3)
        g32 = L.orc_point_triangle_distance_sq(*(np.ascontiguousarray(x, np.float32) for x in (p, a, b, c)))
        assert g32 == pytest.approx(got, rel=2e-4, abs=1e-6)
        # invariant: never above the nearest vertex
        assert got <= min(((p - q) ** 2).sum() for q in (a, b, c)) + 1e-12


def test_solid_angle_known_answers(orc):
    L = orc.lib()
    o = f32(0, 0, 0)
    # octant triangle seen from the origin: pi/2 (sign by orientation)
    om = L.orc_triangle_solid_angle(o, f32(1, 0, 0), f32(0, 1, 0), f32(0, 0, 1))
    assert abs(om) == pytest.approx(math.pi / 2, rel=1e-6)
    om2 = L.orc_triangle_solid_angle(o, f32(1, 0, 0), f32(0, 0, 1), f32(0, 1, 0))
    assert om2 == pytest.approx(-om, rel=1e-6)  # flip negates
    # regular tetrahedron centred at the origin: each face is 4 pi / 4
    s = 1 / math.sqrt(3)
    V = [f64(s, s, s), f64(s, -s, -s), f64(-s, s, -s), f64(-s, -s, s)]
    faces = [(0, 1, 2), (0, 3, 1), (0, 2, 3), (1, 3, 2)]
    tot = 0.0
    for i, j, k in faces:
        w = L.orc_triangle_solid_angle_f64(f64(0, 0, 0), V[i], V[j], V[k])
        assert abs(w) == pytest.approx(math.pi, rel=1e-12)
        tot += w
    assert abs(tot) == pytest.approx(4 * math.pi, rel=1e-12)
    # vertex guard: p on a vertex contributes 0
    assert L.orc_triangle_solid_angle(f32(1, 0, 0), f32(1, 0, 0), f32(0, 1, 0), f32(0, 0, 1)) == 0.0


def test_winding_closed_mesh(orc, geom):
    v, f = geom.icosphere(2)
    tris = orc.build_triangles(v * 10, f)
    mn, dims = (-20, -20, -20), np.array([41, 41, 41], np.int32)
    for mode in (0, 1, 2, 3):
        w_in = orc.winding_at_voxels(tris, 1.0, mn, dims, [20 + 41 * 20 + 41 * 41 * 20], mode)[0]
        w_out = orc.winding_at_voxels(tris, 1.0, mn, dims, [0], mode)[0]
        assert w_in == pytest.approx(1.0, abs=1e-5)
        assert w_out == pytest.approx(0.0, abs=1e-5)
    flipped = orc.build_triangles(v * 10, f[:, [0, 2, 1]])
    assert orc.winding_at_voxels(flipped, 1.0, mn, dims, [20 + 41 * 20 + 41 * 41 * 20], 2)[0] == \
        pytest.approx(-1.0, abs=1e-9)


def test_godunov_closed_forms(orc):
    L = orc.lib()
    h = np.float32(0.5)
    S = np.float32(1e30)
    for s in (0.0, 0.25, 1.0, 7.5):
        s = np.float32(s)
        # one finite axis: 1-D update
        assert L.orc_godunov(s, S, S, h) == pytest.approx(s + h, rel=1e-6)
        # three equal axes: s + h / sqrt(3)
        assert L.orc_godunov(s, s, s, h) == pytest.approx(s + h / math.sqrt(3), rel=1e-6)
        # two equal axes, third far but finite: s + h / sqrt(2)
        assert L.orc_godunov(s, s, np.float32(1e3), h) == pytest.approx(s + h / math.sqrt(2), rel=1e-6)
    # argument order does not matter for finite values
    for perm in ((0.1, 0.3, 0.4), (0.3, 0.1, 0.4), (0.4, 0.3, 0.1)):
        assert L.orc_godunov(*map(np.float32, perm), h) == pytest.approx(L.orc_godunov(0.1, 0.3, 0.4, h), rel=1e-6)


def test_godunov_reference_quirks(orc):
    """Two properties of the reference update that the restatement keeps:
    the 3-D discriminant halves the squared differences (cuda_rasterizer.cu:
    416-418), and the middle value is recovered as sum - min - max, which
    cancels to 0 when exactly one axis holds the 1e30 sentinel (:409)."""
    L = orc.lib()
    # 3-D branch, h = 1: (0 + 0 + 0.5 + sqrt(3 - 0.5 * 0.5)) / 3
    assert L.orc_godunov(0.0, 0.0, 0.5, 1.0) == pytest.approx((0.5 + math.sqrt(2.75)) / 3, rel=1e-6)
    # the textbook Godunov would give (0.5 + sqrt(2.5)) / 3
    assert abs(L.orc_godunov(0.0, 0.0, 0.5, 1.0) - (0.5 + math.sqrt(2.5)) / 3) > 1e-2
    # sentinel cancellation: (h, h, 1e30) evaluates the 2-D root with s1 = 0
    h = 1.0
    got = L.orc_godunov(1.0, 1.0, 1e30, h)
    assert got == pytest.approx((1.0 + 0.0 + math.sqrt(2 - 1.0)) / 2, rel=1e-6)
    assert got < 1.0 + h / math.sqrt(2)


def _cube_mesh():
    v = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    # 12 triangles, two per face (orientation does not matter for distances)
    f = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                  [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.uint32)
    return v, f


def test_query_distances_closed_forms(orc):
    """Unit cube surface: face, edge and vertex regions, inside and outside,
    and the 1e18 start for an empty mesh."""
    v, f = _cube_mesh()
    p = np.array([[2.0, 0.5, 0.5],      # face region, outside: 1
                  [0.5, 0.5, 0.2],      # inside, nearest face z = 0: 0.2
                  [2.0, 2.0, 0.5],      # edge x = y = 1: sqrt(2)
                  [2.0, 2.0, 2.0],      # vertex (1, 1, 1): sqrt(3)
                  [0.25, 0.5, 0.5],     # inside, nearest face x = 0: 0.25
                  [1.0, 1.0, 1.0]],     # on a vertex: 0
                 np.float32)
    d = orc.query_distances(v, f, p)
    want = [1.0, 0.2, math.sqrt(2), math.sqrt(3), 0.25, 0.0]
    assert np.allclose(d, want, rtol=1e-6, atol=1e-7)
    assert orc.query_distances(v, f[:0], p[:1])[0] == np.float32(1e9)  # sqrt(1e18)


def test_query_distances_vs_fp64_min(orc, geom):
    """Every distance equals the FP64 minimum over all triangles to FP32
    accuracy, and degenerate triangles (a point, a segment) take part."""
    L = orc.lib()
    v, f = geom.perturbed_icosphere(seed=3, subdiv=2)
    v = np.concatenate([v, np.array([[0.0, 0.0, 40.0], [5.0, 5.0, 40.0], [10.0, 10.0, 40.0]], np.float32)])
    n = len(v)
    f = np.concatenate([f, np.array([[n - 3, n - 3, n - 3], [n - 3, n - 2, n - 1]], np.uint32)])
    rng = np.random.default_rng(4)
    p = rng.uniform(-35, 45, size=(300, 3)).astype(np.float32)
    d = orc.query_distances(v, f, p)
    v64 = v.astype(np.float64)
    for k in range(len(p)):
        pk = np.ascontiguousarray(p[k], np.float64)
        best = min(L.orc_point_triangle_distance_sq_f64(pk, *(np.ascontiguousarray(v64[i]) for i in tri))
                   for tri in f)
        assert d[k] == pytest.approx(math.sqrt(best), rel=2e-5, abs=1e-5)
    # the point-triangle at (0, 0, 40): a query right above it is 1 away
    assert orc.query_distances(v, f, np.array([[0.0, 0.0, 41.0]], np.float32))[0] == pytest.approx(1.0, abs=1e-6)
