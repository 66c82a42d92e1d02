"""Pins the oracle's band splat, scanline parity, winding fixup, probe
positions and patch boxes against brute force and closed forms (CPU)."""
import numpy as np
import pytest


def seg_tri_hits_fp64(tris, rx, ry):
    """z of every crossing of the vertical line (rx, ry) with the mesh (FP64,
    barycentric), skipping near-degenerate projections."""
    a, b, c = tris[:, 0:3].astype(np.float64), tris[:, 3:6].astype(np.float64), tris[:, 6:9].astype(np.float64)

    def edge(p, q):
        return (q[:, 0] - p[:, 0]) * (ry - p[:, 1]) - (q[:, 1] - p[:, 1]) * (rx - p[:, 0])

    d0, d1, d2 = edge(a, b), edge(b, c), edge(c, a)
    inside = ((d0 > 0) & (d1 > 0) & (d2 > 0)) | ((d0 < 0) & (d1 < 0) & (d2 < 0))
    area = d0 + d1 + d2
    ok = inside & (np.abs(area) > 1e-12)
    z = (d1 * a[:, 2] + d2 * b[:, 2] + d0 * c[:, 2]) / np.where(ok, area, 1)
    return np.sort(z[ok])


def test_parity_vs_brute_force(orc, geom):
    v, f = geom.perturbed_icosphere(seed=4, subdiv=3)
    tris = orc.build_triangles(v, f)
    h = 1.5
    mn, mx = (-16, -16, -16), (15, 15, 15)
    dims = orc.dims_of(mn, mx)
    sign, flag = orc.parity_signs(tris, h, mn, dims)
    assert flag == 0
    s = sign.reshape(dims[2], dims[1], dims[0])
    checked = 0
    for cy in range(0, dims[1], 3):
        for cx in range(0, dims[0], 3):
            rx = (np.float32(mn[0] + cx) + np.float32(0.037)) * np.float32(h)
            ry = (np.float32(mn[1] + cy) + np.float32(0.051)) * np.float32(h)
            hits = seg_tri_hits_fp64(tris, float(rx), float(ry))
            assert len(hits) % 2 == 0
            zc = (mn[2] + np.arange(dims[2])) * h
            below = np.searchsorted(hits, zc, side="left")
            # skip voxel centres within 1e-4 of a crossing (FP32 vs FP64 hit height)
            near = np.min(np.abs(zc[:, None] - hits[None, :]), axis=1) < 1e-4 if len(hits) else np.zeros(len(zc), bool)
            exp = (below % 2).astype(np.uint8)
            assert np.array_equal(s[~near, cy, cx], exp[~near])
            checked += 1
    assert checked > 50


def test_parity_equals_winding_away_from_surface(orc, small_sphere):
    """Jordan-Brouwer: on a watertight mesh parity = winding sign wherever the
    jittered ray point and the voxel centre are on the same side (d > 0.0631 h)."""
    tris, h, mn, mx = small_sphere
    dims = orc.dims_of(mn, mx)
    sign, flag = orc.parity_signs(tris, h, mn, dims)
    d = orc.splat_band(tris, h, mn, dims)
    rng = np.random.default_rng(2)
    far = np.nonzero(d > 0.07 * h)[0]
    pick = rng.choice(far, 1500, replace=False)
    w = orc.winding_at_voxels(tris, h, mn, dims, pick, mode=2)
    assert np.array_equal(sign[pick] == 1, np.abs(w) >= 0.5)


def test_open_mesh_columns_deferred(orc, geom):
    v, f = geom.open_box()
    tris = orc.build_triangles(v, f)
    mn, mx = (-14, -14, -14), (13, 13, 13)
    dims = orc.dims_of(mn, mx)
    sign, flag = orc.parity_signs(tris, 1.0, mn, dims)
    assert flag == 1
    s = sign.reshape(dims[2], dims[1], dims[0])
    # a column through the open top (x, y inside the box) is all 2
    assert np.all(s[:, 14, 14] == 2)
    # a column outside the box footprint is all 0
    assert np.all(s[:, 0, 0] == 0)
    k = orc.lib().orc_winding_fixup(orc._f32(tris).reshape(-1), len(tris), 1.0, orc._i32(mn), orc._i32(dims), sign, 3)
    assert k > 0 and not np.any(sign == 2)


def test_splat_vs_brute_force(orc, geom):
    v, f = geom.perturbed_icosphere(seed=5, subdiv=2)
    tris = orc.build_triangles(v, f)
    h = 2.0
    mn, mx = (-14, -14, -14), (13, 13, 13)
    dims = orc.dims_of(mn, mx)
    bw = 8 * h
    d = orc.splat_band(tris, h, mn, dims)
    L = orc.lib()
    rng = np.random.default_rng(3)
    pick = rng.choice(len(d), 400, replace=False)
    for vi in pick:
        x, y, z = vi % dims[0], (vi // dims[0]) % dims[1], vi // (dims[0] * dims[1])
        p = np.array([(mn[0] + x) * h, (mn[1] + y) * h, (mn[2] + z) * h], np.float32)
        d2 = min(L.orc_point_triangle_distance_sq(p, t[0:3].copy(), t[3:6].copy(), t[6:9].copy()) for t in tris)
        exact = np.sqrt(np.float32(d2))
        if d2 <= bw * bw:
            assert d[vi] == exact
        else:
            assert d[vi] == np.float32(1e30)


def test_verify_indices_closed_form(orc):
    for total in (4096, 4097, 10 ** 6 + 3, 670752407):
        for gen, off in ((0, 0.5), (1, 1 / 6), (2, 5 / 6)):
            got = np.zeros(4096, np.int64)
            orc.lib().orc_verify_indices(total, 4096, gen, got)
            exp = np.minimum(((np.arange(4096) + off) * total / 4096).astype(np.int64), total - 1)
            assert np.array_equal(got, exp)


def test_mark_patch_closed_form(orc):
    dims = np.array([30, 20, 40], np.int32)
    sites = np.array([0, 5 + 30 * 7 + 600 * 33], np.int64)
    s = np.zeros(int(np.prod(dims)), np.uint8)
    orc.lib().orc_mark_patch(s, sites, 2, 4, orc._i32(dims))
    exp = np.zeros((40, 20, 30), np.uint8)
    exp[0:5, 0:5, 0:5] = 2
    exp[29:38, 3:12, 1:10] = 2
    assert np.array_equal(s.reshape(40, 20, 30), exp)


def test_fast_sdf_signs_match_exact_winding(orc, small_sphere):
    """Fast path (reference policy) on a near-sphere: every sign away from the
    jitter band equals the exact winding sign, |d| within the band equals the
    brute-force exact distance."""
    tris, h, mn, mx = small_sphere
    rc, out, sign, st = orc.fast_sdf(tris, h, mn, mx, want_sign=True)
    assert rc == 0 and st.probes >= 1
    dims = orc.dims_of(mn, mx)
    d = orc.splat_band(tris, h, mn, dims)
    rng = np.random.default_rng(4)
    pick = rng.choice(np.nonzero(d > 0.07 * h)[0], 1000, replace=False)
    w = orc.winding_at_voxels(tris, h, mn, dims, pick, mode=2)
    assert np.array_equal(out[pick] < 0, np.abs(w) >= 0.5)
    band = d <= 8 * h
    # the sweeps only ever lower a value (the reference writes whenever the
    # Godunov update is smaller, band voxels included), and the voxels
    # closest to the surface keep their exact splat distance
    assert np.all(np.abs(out[band]) <= d[band])
    near = d <= 0.5 * h
    assert np.array_equal(np.abs(out[near]), d[near])


def test_csr_decline(orc, small_sphere):
    tris, h, mn, mx = small_sphere
    rc, _, _, st = orc.fast_sdf(tris, h, mn, mx, orc.default_params(csr_max_entries=10))
    assert rc == 1 and st.decline_reason == 2


def test_band_f64_vs_binary32_band_and_brute_force(orc, geom):
    """The binary64 band (precision-study yardstick) agrees with the binary32
    splat to binary32 rounding, covers the same voxels, and equals a direct
    binary64 brute-force minimum at sampled voxels."""
    v, f = geom.icosphere(2)
    v = v * 5.0
    tris = orc.build_triangles(v, f)
    h = 0.5
    mn, mx = (-14, -14, -14), (13, 13, 13)
    dims = orc.dims_of(mn, mx)
    d32 = orc.splat_band(tris, h, mn, dims, band_voxels=3)
    d64 = orc.band_f64(tris, h, mn, dims, band_voxels=3)
    inb = d32 < 1e29
    assert np.array_equal(inb, d64 < 1e29) or np.sum(inb != (d64 < 1e29)) <= 2
    both = inb & (d64 < 1e29)
    assert np.max(np.abs(d32[both] - d64[both])) <= 1e-5 * h
    rng = np.random.default_rng(0)
    for i in rng.choice(np.flatnonzero(both), 50, replace=False):
        x = i % dims[0]
        y = (i // dims[0]) % dims[1]
        z = i // (dims[0] * dims[1])
        p = np.array([np.float32(mn[0] + x) * np.float32(h), np.float32(mn[1] + y) * np.float32(h),
                      np.float32(mn[2] + z) * np.float32(h)], np.float64)
        t = tris.reshape(-1, 12).astype(np.float64)
        L = orc.lib()
        best = min(L.orc_point_triangle_distance_sq_f64(p, r[0:3].copy(), r[3:6].copy(), r[6:9].copy())
                   for r in t)
        assert abs(np.sqrt(best) - d64[i]) <= 1e-12
