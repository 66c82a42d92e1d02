"""The reference's analytic acceptance suite, replayed through this library.

The reference verifies its SDF path end to end: analytic mesh ->
``NarrowBandSdf::build_from_mesh`` (bbox framing + the SDF build,
narrow_band_sdf.cpp:9-35) -> ``extract_mesh`` (marching cubes at an iso
value) -> deviation of the extracted surface from the analytic one
(harness.cpp:38-215, harness_eikonal_cuda_tests.cpp:55-285,
cuda_verification.cu:108-195 and 393-509).  Here the same chain runs on
``build_from_mesh`` and ``iso_surface`` of this library, with the
reference's shapes, voxel sizes and gates:

* Tests 1-3: sphere R 20 at 0.40 mm, torus 25 / 5 at 0.35 mm, filleted box
  40 x 40 x 20 (fillet 2) at 0.35 mm; vertex and midpoint max deviation
  <= 0.100 mm.
* Test 10 case (a): |grad f| of the trilinear interpolant of the built field
  (sphere R 20 at 0.35 mm), six-point central differences with spacing
  0.5 h, 55,000 samples in the +-3 h shell drawn with the reference's hash
  generator (seed 1337); the case passes when >= 85% are within 0.02.  The
  evaluator's own self-test on the analytic field (>= 99.9% within 0.001,
  RMS <= 0.001) pins the numpy restatement.
* Test 11: iso surfaces at t = +-2 mm of the sphere (R 22 / 18) and of the
  torus (minor 7 / 3) at 0.35 mm; vertex and midpoint max <= 0.100 mm.

The input meshes are this repository's generators (UV sphere and torus,
96 x 96; the filleted box is the marching-cubes mesh of its analytic SDF at
0.2 mm, from the oracle's extraction); the analytic surfaces and gates are
the reference's (cuda_verification.cu:108-131).
"""
import math

import numpy as np
import pytest


def _methods(rhz):
    # "auto" is the reference policy: its fast path behind its own gates, the
    # legacy exact kernel where they decline (cuda_rasterizer.cu:1163-1265)
    return {"learned": rhz.METHOD_FAST_LEARNED, "auto": rhz.METHOD_AUTO}


def _build(rhz, v, f, h, method):
    mn, mx, field, st = rhz.build_from_mesh(v, f, h, rhz.default_params(method=method))
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    assert st.status == 0
    return mn, dims, field.reshape(-1)


def _extract(rhz, mn, dims, field, h, iso):
    origin = [np.float32(mn[i]) * np.float32(h) for i in range(3)]
    return rhz.iso_surface(field, dims, h, origin, iso)


def _sdf_sphere(p, R):
    return np.linalg.norm(p, axis=-1) - R


def _sdf_torus(p, R, r):
    q = np.sqrt(p[..., 0] ** 2 + p[..., 1] ** 2) - R
    return np.sqrt(q * q + p[..., 2] ** 2) - r


def _sdf_filleted_box(p, half, r):
    q = np.abs(p) - (np.asarray(half, dtype=np.float64) - r)
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(q.max(axis=-1), 0.0)
    return outside + inside - r


def _deviations(verts, tris, sdf):
    """Vertex max and midpoint max (three edge midpoints and the centroid of
    every triangle), cuda_verification.cu:134-195, in float64."""
    v = verts.astype(np.float64)
    t = tris.astype(np.int64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    mids = np.concatenate([(a + b) * 0.5, (b + c) * 0.5, (c + a) * 0.5, (a + b + c) / 3.0])
    return float(np.abs(sdf(v)).max()), float(np.abs(sdf(mids)).max())


def _filleted_box_mesh(orc, half=(20.0, 20.0, 10.0), r=2.0, step=0.2):
    n = [int(math.ceil((half[i] + 1.0) / step)) for i in range(3)]
    axes = [np.arange(-n[i], n[i] + 1, dtype=np.float64) * step for i in range(3)]
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    grid = _sdf_filleted_box(np.stack([X, Y, Z], -1), half, r).astype(np.float32)
    dims = [2 * n[i] + 1 for i in range(3)]
    origin = [-n[i] * step for i in range(3)]
    v, f = orc.mc_extract(grid.reshape(-1), dims, step, origin, 0.0)
    assert len(f) > 0
    return v, f


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["learned", "auto"])
@pytest.mark.parametrize("case", ["sphere", "torus", "filleted_box"])
def test_reference_tests_1_to_3(rhz, geom, orc, torch_gpu, case, method):
    """harness.cpp:38-165 / :173-183: extracted surface within 0.100 mm of the
    analytic one, at vertices and at triangle midpoints."""
    if case == "sphere":
        v, f = geom.uv_sphere(20.0, 96, 96)
        h, sdf = 0.40, lambda p: _sdf_sphere(p, 20.0)
    elif case == "torus":
        v, f = geom.torus(25.0, 5.0, center=(0.0, 0.0, 0.0), axis="z", nu=96, nv=96)
        h, sdf = 0.35, lambda p: _sdf_torus(p, 25.0, 5.0)
    else:
        v, f = _filleted_box_mesh(orc)
        h, sdf = 0.35, lambda p: _sdf_filleted_box(p, (20.0, 20.0, 10.0), 2.0)
    mn, dims, field = _build(rhz, v, f, h, _methods(rhz)[method])
    verts, tris = _extract(rhz, mn, dims, field, h, 0.0)
    assert len(verts) > 0 and len(tris) > 0
    vmax, mmax = _deviations(verts, tris, sdf)
    print(f"{case} h={h} ({method}): {len(verts)} verts, vertex max {vmax:.4f} mm, midpoint max {mmax:.4f} mm")
    assert vmax <= 0.100 and mmax <= 0.100


@pytest.mark.gpu
@pytest.mark.parametrize("case,t,p0,p1", [("sphere", 2.0, 22.0, 0.0), ("sphere", -2.0, 18.0, 0.0),
                                          ("torus", 2.0, 25.0, 7.0), ("torus", -2.0, 25.0, 3.0)])
def test_reference_test_11_offsets(rhz, geom, torch_gpu, case, t, p0, p1):
    """harness_eikonal_cuda_tests.cpp:218-285: the iso surface at t = +-2 mm
    is the analytic offset surface within 0.100 mm."""
    if case == "sphere":
        v, f = geom.uv_sphere(20.0, 96, 96)
        sdf = lambda p: _sdf_sphere(p, p0)  # noqa: E731
    else:
        v, f = geom.torus(25.0, 5.0, center=(0.0, 0.0, 0.0), axis="z", nu=96, nv=96)
        sdf = lambda p: _sdf_torus(p, p0, p1)  # noqa: E731
    h = 0.35
    mn, dims, field = _build(rhz, v, f, h, rhz.METHOD_FAST_LEARNED)
    verts, tris = _extract(rhz, mn, dims, field, h, t)
    assert len(verts) > 0, "no surface at the offset"
    vmax, mmax = _deviations(verts, tris, sdf)
    print(f"{case} t={t:+.1f}: {len(verts)} verts, vertex max {vmax:.4f} mm, midpoint max {mmax:.4f} mm")
    assert vmax <= 0.100 and mmax <= 0.100


# ---- Test 10: the eikonal property of the stored field -------------------------
def _hash01(x):
    """dev_hash_01 (math_utils.cuh:219-226) on uint32 arrays."""
    x = np.asarray(x, dtype=np.uint32).copy()
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return (x & np.uint32(0x00FFFFFF)).astype(np.float64) / 16777216.0


def _shell_samples(seed, n, R, h, mode):
    """Sample modes 0 and 4 of gpu_evaluate_eikonal_field_kernel
    (cuda_verification.cu:413-424): radius R + U(-3h, 3h), direction from
    the hash; mode 0 puts the polar axis on y, mode 4 on z."""
    i = np.arange(n, dtype=np.int64)
    s0 = (np.uint64(seed) + (i * 19997 + 31).astype(np.uint64)) & np.uint64(0xFFFFFFFF)
    th = 6.2831853 * _hash01(s0 + 1)
    cp = -1.0 + 2.0 * _hash01(s0 + 2)
    sp = np.sqrt(np.maximum(0.0, 1.0 - cp * cp))
    r = R + (-3.0 * h + 6.0 * h * _hash01(s0 + 3))
    if mode == 4:
        return np.stack([r * sp * np.cos(th), r * sp * np.sin(th), r * cp], -1)
    return np.stack([r * sp * np.cos(th), r * cp, r * sp * np.sin(th)], -1)


def _trilinear(grid, mn, dims, h, p):
    """dev_sample_sdf_trilinear (math_utils.cuh:344-388); every sample here
    lies inside the lattice, so the nearest-voxel fallback is not needed."""
    g = (p / h) - np.asarray(mn, dtype=np.float64)
    i0 = np.floor(g).astype(np.int64)
    assert np.all(i0 >= 0) and np.all(i0 < np.asarray(dims) - 1)
    fr = g - i0
    G = grid.reshape(dims[2], dims[1], dims[0]).astype(np.float64)
    out = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = ((fr[:, 0] if dx else 1 - fr[:, 0]) * (fr[:, 1] if dy else 1 - fr[:, 1]) *
                     (fr[:, 2] if dz else 1 - fr[:, 2]))
                out = out + w * G[i0[:, 2] + dz, i0[:, 1] + dy, i0[:, 0] + dx]
    return out


def _gradient_error(grid, mn, dims, h, p):
    s = 0.5 * h
    g = np.zeros_like(p)
    for a in range(3):
        e = np.zeros(3)
        e[a] = s
        g[:, a] = (_trilinear(grid, mn, dims, h, p + e) - _trilinear(grid, mn, dims, h, p - e)) / (2 * s)
    return np.abs(np.linalg.norm(g, axis=1) - 1.0)


def test_hash_restatement_matches_oracle(orc):
    xs = np.array([0, 1, 42, 1337, 9001, 0x7FFFFFFF, 0xFFFFFFFF, 123456789], dtype=np.uint64)
    ours = _hash01(xs)
    ref = [orc.lib().orc_hash_01(int(x)) for x in xs]
    assert np.array_equal(ours.astype(np.float32), np.asarray(ref, dtype=np.float32))


def test_reference_test_10_self_test_on_analytic_field():
    """harness_eikonal_cuda_tests.cpp:19-52: the stencil on f = |x| - R
    sampled at 0.35 mm (R 20) is within 0.001 of 1 for >= 99.9% of the
    samples, RMS <= 0.001 (pins the evaluator)."""
    R, h = 20.0, 0.35
    b = int(math.ceil((R + 6.0 * h + h) / h))
    mn = (-b, -b, -b)
    dims = [2 * b + 1] * 3
    ax = np.arange(-b, b + 1, dtype=np.float64) * h
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    grid = (np.sqrt(X * X + Y * Y + Z * Z) - R).astype(np.float32)
    p = _shell_samples(42, 55000, R, h, mode=4)
    err = _gradient_error(grid, mn, dims, h, p)
    frac = float((err <= 0.001).mean())
    rms = float(np.sqrt((err ** 2).mean()))
    assert frac >= 0.999 and rms <= 0.001, (frac, rms)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["learned", "auto"])
def test_reference_test_10_eikonal_property(rhz, geom, torch_gpu, method):
    """harness_eikonal_cuda_tests.cpp:100-126 case (a): the built field of
    the rasterised sphere (R 20, 0.35 mm) has ||grad f| - 1| <= 0.02 on
    >= 85% of the shell samples."""
    R, h = 20.0, 0.35
    v, f = geom.uv_sphere(R, 96, 96)
    mn, dims, field = _build(rhz, v, f, h, _methods(rhz)[method])
    p = _shell_samples(1337, 55000, R, h, mode=0)
    err = _gradient_error(field, mn, dims, h, p)
    frac = float((err <= 0.02).mean())
    print(f"Test 10 (a) {method}: in tolerance {frac:.4f}, max {err.max():.4f}, RMS {np.sqrt((err ** 2).mean()):.5f}")
    assert frac >= 0.85 and np.isfinite(err).all()


def test_obj_writer_round_trip_is_exact(geom, tmp_path):
    from __graft_entry__ import load_obj
    obj = load_obj()
    v, f = geom.torus(25.0, 5.0, center=(0.0, 0.0, 0.0), axis="z", nu=24, nv=12)
    path = str(tmp_path / "t.obj")
    obj.save_obj(path, v, f)
    v2, f2 = obj.load_obj(path)
    assert np.array_equal(np.asarray(v, np.float32), v2) and np.array_equal(np.asarray(f, np.uint32), f2)


@pytest.mark.gpu
def test_reference_test_8_round_trip(rhz, geom, torch_gpu, tmp_path):
    """harness_vessel_tests.cpp:326-381: the vessel built at 0.15 mm,
    extracted at iso 0, written as OBJ and read back; the re-imported
    vertices and edge midpoints (cuda_verification.cu:307-349) lie within
    0.100 mm of the original mesh.  Distances are the exact unsigned ones of
    rhz_query_distances, on every 40th vertex and every 80th triangle's three
    midpoints (the full set is ~2e13 point-triangle pairs)."""
    from __graft_entry__ import load_obj
    obj = load_obj()
    v, f = geom.vessel()
    h = 0.15
    mn, dims, field = _build(rhz, v, f, h, rhz.METHOD_FAST_LEARNED)
    verts, tris = _extract(rhz, mn, dims, field, h, 0.0)
    path = str(tmp_path / "vessel_roundtrip.obj")
    obj.save_obj(path, verts, tris)
    rv, rf = obj.load_obj(path)
    assert len(rv) == len(verts) and len(rf) == len(tris)
    qv = rv[::40].astype(np.float32)
    t = rf[::80].astype(np.int64)
    a, b, c = rv[t[:, 0]], rv[t[:, 1]], rv[t[:, 2]]
    qm = np.concatenate([(a + b) * 0.5, (b + c) * 0.5, (c + a) * 0.5]).astype(np.float32)
    dv = rhz.query_distances(v, f, qv)
    dm = rhz.query_distances(v, f, qm)
    print(f"Test 8 round trip: {len(rv)} verts, sampled {len(qv)} vertices / {len(qm)} midpoints, "
          f"vertex max {dv.max():.4f} mm, midpoint max {dm.max():.4f} mm")
    assert np.isfinite(dv).all() and np.isfinite(dm).all()
    assert dv.max() <= 0.100 and dm.max() <= 0.100


@pytest.mark.gpu
def test_reference_test_15_marching_cubes(rhz, torch_gpu):
    """harness_eikonal_cuda_tests.cpp:360-385: f = |p| - 10 on a 30^3 lattice
    at 1 mm from (-15, -15, -15) extracts a non-empty surface; here also every
    vertex lies within 0.02 mm of the sphere (linear interpolation of the exact
    field along an edge errs by about h^2 / (8 R) = 0.0125 mm)."""
    n, h, R = 30, 1.0, 10.0
    ax = (np.arange(n, dtype=np.float64) - 15.0) * h
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    grid = (np.sqrt(X * X + Y * Y + Z * Z) - R).astype(np.float32)
    verts, tris = rhz.iso_surface(grid.reshape(-1), [n, n, n], h, [-15.0, -15.0, -15.0], 0.0)
    assert len(tris) > 0
    dev = np.abs(np.linalg.norm(verts.astype(np.float64), axis=1) - R)
    assert dev.max() <= 0.02, dev.max()
