"""Pins the oracle's sweeps: exact linear fields from planar seeds, idempotence
on a converged field, the double-buffered schedule, and the round policy."""
import numpy as np
import pytest


def planar_seed(dims, h, axis, band=8):
    """Band of a plane through index 0 of `axis`: d = |i| h for |i| <= band."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = (x, y, z)[axis].astype(np.float32)
    d = np.where(idx <= band, idx * np.float32(h), np.float32(1e30)).astype(np.float32)
    return d.reshape(-1), idx.reshape(-1)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_planar_seed_gives_linear_field(orc, axis):
    dims = np.array([30, 20, 25], np.int32)
    h = 0.5
    d, idx = planar_seed(dims, h, axis)
    rounds = orc.fsm_converge(d, dims, h)
    assert rounds <= 3
    np.testing.assert_allclose(d, idx * h, rtol=1e-6, atol=1e-6)


def test_converged_field_is_idempotent(orc, small_sphere):
    tris, h, mn, mx = small_sphere
    dims = orc.dims_of(mn, mx)
    d = orc.splat_band(tris, h, mn, dims)
    orc.fsm_converge(d, dims, h)
    before = d.copy()
    for ax in range(3):
        for dr in (1, -1):
            assert orc.lib().orc_fsm_sweep(d, orc._i32(dims), h, ax, dr) == 0
    assert np.array_equal(before, d)


def test_double_buffered_sweep_reads_only_src(orc, small_sphere):
    """The pinned schedule: dst depends only on src (cross-line reads come
    from src); running it twice from the same src gives the same dst, and a
    line-after-line in-place sweep is a different interleaving."""
    tris, h, mn, mx = small_sphere
    dims = orc.dims_of(mn, mx)
    src = orc.splat_band(tris, h, mn, dims)
    a = np.empty_like(src)
    b = np.empty_like(src)
    orc.lib().orc_fsm_sweep_db(src, a, orc._i32(dims), h, 0, 1)
    orc.lib().orc_fsm_sweep_db(src, b, orc._i32(dims), h, 0, 1)
    assert np.array_equal(a, b)
    c = src.copy()
    orc.lib().orc_fsm_sweep_inplace(c, orc._i32(dims), h, 1, 1)
    d = np.empty_like(src)
    orc.lib().orc_fsm_sweep_db(src, d, orc._i32(dims), h, 1, 1)
    assert not np.array_equal(c, d)


def test_sweep_only_decreases(orc, small_sphere):
    tris, h, mn, mx = small_sphere
    dims = orc.dims_of(mn, mx)
    src = orc.splat_band(tris, h, mn, dims)
    dst = np.empty_like(src)
    for ax in range(3):
        orc.lib().orc_fsm_sweep_db(src, dst, orc._i32(dims), h, ax, -1)
        assert np.all(dst <= src)
        src = dst.copy()


def test_far_field_close_to_euclidean(orc, small_sphere):
    """Converged far field of a near-sphere stays within a few voxels of the
    exact distance (the Godunov scheme is O(h); the reference's halved 3-D
    discriminant biases diagonals upward)."""
    tris, h, mn, mx = small_sphere
    dims = orc.dims_of(mn, mx)
    d = orc.splat_band(tris, h, mn, dims)
    orc.fsm_converge(d, dims, h)
    band = d <= 8 * h
    # brute force exact distance on a sample of far voxels
    rng = np.random.default_rng(1)
    far = np.nonzero(~band)[0]
    pick = rng.choice(far, 200, replace=False)
    exact = np.sqrt([min_dist2(tris, h, mn, dims, v) for v in pick])
    err = np.abs(d[pick] - exact) / h
    assert err.max() < 3.0
    assert np.median(err) < 1.0


def min_dist2(tris, h, mn, dims, v):
    x, y, z = v % dims[0], (v // dims[0]) % dims[1], v // (dims[0] * dims[1])
    p = np.array([(mn[0] + x) * h, (mn[1] + y) * h, (mn[2] + z) * h], np.float64)
    a, b, c = tris[:, 0:3].astype(np.float64), tris[:, 3:6].astype(np.float64), tris[:, 6:9].astype(np.float64)
    # lower bound by vertex distance is enough to select candidates; exact via oracle
    cand = np.argsort(((a - p) ** 2).sum(1))[:64]
    import oracle
    L = oracle.lib()
    return min(L.orc_point_triangle_distance_sq_f64(p, a[i].copy(), b[i].copy(), c[i].copy()) for i in cand)


def test_round_policy_reference_cap(orc, small_sphere):
    """One round of x/y/z sweeps reaches every voxel (a staircase path x, then
    y, then z connects any voxel to the band), so sentinels survive only when
    no round runs: then the audit-gated extension (or, with no extension
    rounds, the safe clamp) takes over."""
    tris, h, mn, mx = small_sphere
    rc, out, _, st = orc.fast_sdf(tris, h, mn, mx, orc.default_params(max_rounds=1))
    assert rc == 0 and st.extra_rounds == 0 and st.eikonal_rounds == 1 and st.far_residual == 0
    rc, out, _, st = orc.fast_sdf(tris, h, mn, mx, orc.default_params(max_rounds=0))
    assert rc == 0
    assert st.extra_rounds == 1 and st.far_residual == 0
    rc, out, _, st = orc.fast_sdf(tris, h, mn, mx, orc.default_params(max_rounds=0, max_extra_rounds=0))
    assert st.safe_clamped > 0 and st.far_residual == 0
    band = 8 * h
    assert set(np.unique(np.abs(out[np.abs(out) == band]))) <= {np.float32(band)}
