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


def _converge_octants(orc, d, dims, h, max_rounds=512):
    dm = orc._i32(dims)
    for r in range(max_rounds):
        ch = 0
        for sz in (1, -1):
            for sy in (1, -1):
                for sx in (1, -1):
                    ch |= orc.lib().orc_fsm_sweep_octant(d, dm, h, sx, sy, sz)
        if not ch:
            return r + 1
    return max_rounds


def test_reference_update_fixed_point_depends_on_sweep_order(orc, small_sphere):
    """The reference's Godunov update halves the squared differences of its
    3-D discriminant (cuda_rasterizer.cu:416-418), which makes it
    non-monotone in the largest neighbour: the 6-axis sweeps and the classical
    8-ordering sweeps (Zhao 2005) converge to different fixed points.  This is
    why the sweeps are pinned to one schedule bit for bit rather than to a
    converged solution."""
    tris, h, mn, mx = small_sphere
    dims = orc.dims_of(mn, mx)
    band = orc.splat_band(tris, h, mn, dims)
    a = band.copy()
    b = band.copy()
    orc.fsm_converge(a, dims, h)
    _converge_octants(orc, b, dims, h)
    assert np.all(a < 1e29) and np.all(b < 1e29)
    diff = np.abs(a.astype(np.float64) - b.astype(np.float64))
    assert diff.max() > 0.1 * h  # measured: 0.31 voxel, in about a fifth of the cells


def _godunov_textbook(a, b, c, h):
    s0, s1, s2 = sorted((a, b, c))
    u = s0 + h
    if u > s1:
        u = (s0 + s1 + np.sqrt(max(2 * h * h - (s1 - s0) ** 2, 0.0))) / 2
        if u > s2:
            disc = 3 * h * h - ((s1 - s0) ** 2 + (s2 - s1) ** 2 + (s2 - s0) ** 2)
            u = (s0 + s1 + s2 + np.sqrt(max(disc, 0.0))) / 3
    return u


def _textbook_sweep(g, h, order):
    """In-place Gauss-Seidel over the cells in `order` (float64, textbook
    Godunov update)."""
    nz, ny, nx = g.shape
    big = 1e30
    changed = False
    for z, y, x in order:
        ax = min(g[z, y, x - 1] if x > 0 else big, g[z, y, x + 1] if x < nx - 1 else big)
        ay = min(g[z, y - 1, x] if y > 0 else big, g[z, y + 1, x] if y < ny - 1 else big)
        az = min(g[z - 1, y, x] if z > 0 else big, g[z + 1, y, x] if z < nz - 1 else big)
        if min(ax, ay, az) >= big:
            continue
        u = _godunov_textbook(ax, ay, az, h)
        if u < g[z, y, x]:
            g[z, y, x] = u
            changed = True
    return changed


def test_textbook_update_fixed_point_is_order_independent():
    """With the textbook discriminant 3h^2 - sum of squared differences the
    update is monotone, and the axis-line sweeps and the 8-ordering sweeps
    reach the same fixed point (the scheme's unique discrete solution)."""
    n, h = 12, 1.0
    rng = np.random.default_rng(4)
    seed = np.full((n, n, n), 1e30)
    for _ in range(3):
        z, y, x = rng.integers(0, n, 3)
        seed[z, y, x] = rng.uniform(0.0, 1.0)
    idx = [(z, y, x) for z in range(n) for y in range(n) for x in range(n)]

    def axis_orders():
        for axis in range(3):
            for d in (1, -1):
                # lines along `axis`, each walked in direction d (cells of one
                # line in order; lines one after another)
                key = [lambda c: (c[0], c[1], c[2] * d), lambda c: (c[0], c[2], c[1] * d),
                       lambda c: (c[1], c[2], c[0] * d)][axis]
                yield sorted(idx, key=key)

    def octant_orders():
        for sz in (1, -1):
            for sy in (1, -1):
                for sx in (1, -1):
                    yield sorted(idx, key=lambda c: (c[0] * sz, c[1] * sy, c[2] * sx))

    a = seed.copy()
    for _ in range(50):
        if not any([_textbook_sweep(a, h, o) for o in axis_orders()]):
            break
    b = seed.copy()
    for _ in range(50):
        if not any([_textbook_sweep(b, h, o) for o in octant_orders()]):
            break
    assert np.all(a < 1e29)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)


ORDERINGS = [(sx, sy, sz) for sz in (1, -1) for sy in (1, -1) for sx in (1, -1)]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_octant_planar_seed_gives_linear_field(orc, axis):
    """The 8 classical orderings from a planar seed reach the exact linear
    field (h a power of two: every sum is exact) in one round; a second
    round changes nothing."""
    dims = np.array([30, 20, 25], np.int32)
    h = 0.5
    d, idx = planar_seed(dims, h, axis)
    assert orc.octant_converge(d, dims, h) == 2
    assert np.array_equal(d, (idx * np.float32(h)).astype(np.float32))


def _wavefront_sweep(orc, g, dims, h, sx, sy, sz, wrong=False):
    """The ordering as diagonal levels kx + ky + kz = L (each axis counted in
    its direction), cell by cell with the oracle's Godunov update.  A cell of
    level L reads level L - 1 (updated) and L + 1 (not yet): the same values
    as the lexicographic in-place order.  wrong=True reverses the level order
    (the downwind side goes first), a planted mistake."""
    nx, ny, nz = (int(v) for v in dims)
    a = g.reshape(nz, ny, nx)
    big = np.float32(1e30)
    levels = range(nx + ny + nz - 2)
    for L in (reversed(levels) if wrong else levels):
        for kx in range(nx):
            for ky in range(ny):
                kz = L - kx - ky
                if not 0 <= kz < nz:
                    continue
                x = kx if sx > 0 else nx - 1 - kx
                y = ky if sy > 0 else ny - 1 - ky
                z = kz if sz > 0 else nz - 1 - kz
                ax = min(a[z, y, x - 1] if x > 0 else big, a[z, y, x + 1] if x < nx - 1 else big)
                ay = min(a[z, y - 1, x] if y > 0 else big, a[z, y + 1, x] if y < ny - 1 else big)
                az = min(a[z - 1, y, x] if z > 0 else big, a[z + 1, y, x] if z < nz - 1 else big)
                u = np.float32(orc.lib().orc_godunov(ax, ay, az, h))
                if u < a[z, y, x]:
                    a[z, y, x] = u


def _seeded_field(dims, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    d = np.full(n, 1e30, np.float32)
    pick = rng.choice(n, 5, replace=False)
    d[pick] = rng.uniform(0.0, 1.5, 5).astype(np.float32)
    return d


@pytest.mark.parametrize("seed", [1, 2])
def test_octant_sweep_equals_wavefront_restatement(orc, seed):
    """orc_fsm_sweep_octant against an independent traversal (diagonal
    levels) for every ordering, bit for bit; the planted level reversal must
    differ."""
    dims = np.array([9, 7, 6], np.int32)
    h = 0.7
    d = _seeded_field(dims, seed)
    ref = d.copy()
    for o in ORDERINGS:
        before = d.copy()
        orc.lib().orc_fsm_sweep_octant(d, orc._i32(dims), h, *o)
        w = before.copy()
        _wavefront_sweep(orc, w, dims, h, *o)
        assert np.array_equal(d, w), o
    bad = ref.copy()
    for o in ORDERINGS:
        _wavefront_sweep(orc, bad, dims, h, *o, wrong=True)
    assert not np.array_equal(bad, d)


def test_octant_policy_in_the_oracle_pipeline(orc, small_sphere):
    """fsm_converge = 2 runs octant rounds until nothing changes: the
    pipeline's field equals octant_converge on its band, with the signs."""
    tris, h, mn, mx = small_sphere
    dims = orc.dims_of(mn, mx)
    rc, out, sg, st = orc.fast_sdf(tris, h, mn, mx, orc.default_params(fsm_converge=2), want_sign=True)
    assert rc == 0 and st.far_residual == 0
    band = orc.splat_band(tris, h, mn, dims)
    rounds = orc.octant_converge(band, dims, h)
    assert st.eikonal_rounds == rounds
    assert np.array_equal(np.abs(out), band)
