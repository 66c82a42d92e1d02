"""Pipeline-level parity: the HIP fast path, plain-exact path and legacy
kernel against the CPU oracle of the same policy, through the host C-ABI
(rhz_sdf_compute), plus the standalone entry points.

Tolerances: signs bit-exact; distances bit-exact where both sides evaluate
identical binary32 expressions (band, sweeps, apply) and within 1e-3 voxel
otherwise."""
import numpy as np
import pytest

from conftest import dims_of

pytestmark = pytest.mark.gpu

STAT_KEYS = ["unknown_flag", "fixup_voxels", "eikonal_rounds", "extra_rounds", "far_residual", "safe_clamped",
             "probes", "sites", "patched_voxels", "patch_rounds", "last_probe_m", "csr_entries"]


def _compare_fast(rhz, orc, tris, h, mn, mx, gpu_kw, orc_kw):
    p = rhz.default_params(method=rhz.METHOD_FAST, **gpu_kw)
    rc, field, st = rhz.compute_sdf(tris, h, mn, mx, p)
    orc_rc, ref, ref_sign, ost = orc.fast_sdf(tris, h, mn, mx, orc.default_params(**orc_kw), want_sign=True)
    assert rc == orc_rc, (st.as_dict(), ost.as_dict())
    if rc != rhz.RHZ_OK:
        assert st.decline_reason == ost.decline_reason
        return st, ost, None, None
    got = field.reshape(-1)
    for k in STAT_KEYS:
        assert getattr(st, k) == getattr(ost, k), (k, getattr(st, k), getattr(ost, k))
    assert np.array_equal(np.signbit(got), np.signbit(ref))
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
        f"max |diff| {np.abs(got - ref).max() / h:.3e} voxel"
    return st, ost, got, ref


def test_fast_path_small_reference_policy(rhz, orc, small_sphere):
    tris, h, mn, mx = small_sphere
    st, ost, got, ref = _compare_fast(rhz, orc, tris, h, mn, mx, {}, {})
    assert st.method_used == rhz.METHOD_FAST
    assert st.eikonal_rounds >= 1


def test_fast_path_small_converged(rhz, orc, small_sphere):
    tris, h, mn, mx = small_sphere
    st, ost, got, ref = _compare_fast(rhz, orc, tris, h, mn, mx, {"fsm_policy": rhz.FSM_CONVERGE},
                                      {"fsm_converge": 1})
    # converged: another round changes nothing
    dims = dims_of(mn, mx)
    d = np.abs(ref).astype(np.float32)
    assert orc.fsm_converge(d.copy(), dims, h, max_rounds=1) == 1


def test_fast_path_ico5_128(rhz, orc, ico5_128):
    tris, h, mn, mx = ico5_128
    st, ost, got, ref = _compare_fast(rhz, orc, tris, h, mn, mx, {}, {})
    print(st.log_line())


def test_safe_clamp_path(rhz, orc, small_sphere):
    """No rounds and no extension rounds leave sentinels: the safe clamp runs
    and ships -band inside / +band outside."""
    tris, h, mn, mx = small_sphere
    for intended in (1, 0):
        st, ost, got, ref = _compare_fast(rhz, orc, tris, h, mn, mx,
                                          {"max_rounds": 0, "max_extra_rounds": 0, "clamp_intended_sign": intended},
                                          {"max_rounds": 0, "max_extra_rounds": 0, "clamp_intended_sign": intended})
        assert st.safe_clamped > 0 and st.far_residual == 0


def test_extension_rounds(rhz, orc, small_sphere):
    tris, h, mn, mx = small_sphere
    st, ost, got, ref = _compare_fast(rhz, orc, tris, h, mn, mx, {"max_rounds": 0}, {"max_rounds": 0})
    assert st.extra_rounds == 1 and st.far_residual == 0


def test_extension_rounds_after_rounds_on_the_working_copy(rhz, orc, small_sphere):
    """A lattice above the sphere, farther than the band: its columns cross the
    mesh (so the path does not decline) but nothing is splatted, the rounds
    leave every voxel at the sentinel, and the audit that the write-back of the
    y-fastest working copy runs (41 x 36 x 35, rows padded to 64) finds them:
    the extension round and the safe clamp then run after the rounds, as in
    the oracle."""
    tris, h, _, _ = small_sphere
    mn, mx = (-18, -17, 40), (22, 18, 74)
    st, ost, got, ref = _compare_fast(rhz, orc, tris, h, mn, mx, {"max_rounds": 2, "max_extra_rounds": 1},
                                      {"max_rounds": 2, "max_extra_rounds": 1})
    assert got is not None, st.as_dict()
    assert st.eikonal_rounds == ost.eikonal_rounds and st.extra_rounds == 1
    assert st.safe_clamped == got.size and st.far_residual == 0


def test_decline_csr(rhz, orc, small_sphere):
    tris, h, mn, mx = small_sphere
    st, ost, _, _ = _compare_fast(rhz, orc, tris, h, mn, mx, {"csr_max_entries": 10}, {"csr_max_entries": 10})
    assert st.decline_reason == 2


def test_auto_falls_back_to_legacy_on_decline(rhz, small_sphere):
    tris, h, mn, mx = small_sphere
    p = rhz.default_params(method=rhz.METHOD_AUTO, fast_min_voxels=1, csr_max_entries=10)
    rc, field, st = rhz.compute_sdf(tris, h, mn, mx, p)
    assert rc == rhz.RHZ_OK and st.method_used == rhz.METHOD_LEGACY


def test_overlap_mesh_verify_patches(rhz, orc, geom):
    """Two overlapping cubes (reference Test 5 fixture): winding 2 in the
    overlap, which parity classifies as outside -> the probes find mismatches;
    both sides run the same patch/decline policy."""
    v, f = geom.two_cubes()
    tris = orc.build_triangles(v, f)
    h = 0.5
    mn, mx = (-30, -30, -30), (45, 29, 29)
    _compare_fast(rhz, orc, tris, h, mn, mx, {}, {})


def test_legacy_matches_oracle(rhz, orc, small_sphere):
    tris, h, mn, mx = small_sphere
    p = rhz.default_params(method=rhz.METHOD_LEGACY)
    rc, field, st = rhz.compute_sdf(tris, h, mn, mx, p)
    assert rc == rhz.RHZ_OK
    ref = orc.legacy_sdf(tris, h, mn, mx)
    got = field.reshape(-1)
    assert np.array_equal(np.abs(got).view(np.uint32), np.abs(ref).view(np.uint32))
    # FP32 winding sums differ in the last bits (device vs host atan2f); a sign
    # may differ only where |w| is within 1e-5 of 0.5
    diff = np.nonzero(np.signbit(got) != np.signbit(ref))[0]
    if len(diff):
        dims = dims_of(mn, mx)
        w = orc.winding_at_voxels(tris, h, mn, dims, diff, mode=2)
        assert np.all(np.abs(np.abs(w) - 0.5) < 1e-5)


def test_legacy_sphere_closed_form(rhz, geom, orc):
    """Reference Test 12 (harness_eikonal_cuda_tests.cpp:287-313): sphere
    R=15 at 1 mm on a 37^3 grid, max |err| vs |p|-15 below 1.5 mm."""
    v, f = geom.uv_sphere(15.0, 32, 32)
    tris = rhz.build_triangles(v, f)
    mn, mx = (-18, -18, -18), (18, 18, 18)
    rc, field, st = rhz.compute_sdf(tris, 1.0, mn, mx)
    assert rc == rhz.RHZ_OK and st.method_used == rhz.METHOD_LEGACY
    ax = np.arange(-18, 19, dtype=np.float64)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    exact = np.sqrt(x * x + y * y + z * z) - 15.0
    assert np.abs(field - exact).max() < 1.5


def test_plain_exact_matches_oracle(rhz, orc, small_sphere):
    tris, h, mn, mx = small_sphere
    p = rhz.default_params(method=rhz.METHOD_PLAIN_EXACT)
    rc, field, st = rhz.compute_sdf(tris, h, mn, mx, p)
    assert rc == rhz.RHZ_OK
    ref, ref_sign, rounds = orc.plain_exact_sdf(tris, h, mn, mx, winding_mode=3, want_sign=True)
    got = field.reshape(-1)
    assert np.array_equal(np.signbit(got), np.signbit(ref))
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert st.eikonal_rounds == rounds


def test_build_triangles_matches_oracle(rhz, orc, geom):
    v, f = geom.perturbed_icosphere(seed=1, subdiv=3)
    f = np.concatenate([f, np.array([[0, 0, 1], [5, 6, 999999]], dtype=np.uint32)])  # degenerate + out of range
    got = rhz.build_triangles(v, f)
    ref = orc.build_triangles(v, f)
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_bbox_and_framing(rhz, geom):
    v, f = geom.vessel(step=3.0)
    lo, hi = rhz.mesh_bbox(v)
    assert np.array_equal(lo, v.min(axis=0)) and np.array_equal(hi, v.max(axis=0))
    mn, mx = rhz.frame_grid(v, 0.1)
    assert (mn, mx) == geom.frame_grid(v, 0.1)


def test_winding_numbers_test5_fixtures(rhz, orc, geom):
    """Reference Test 5 (harness_vessel_tests.cpp:54-121): 3 dirty meshes x
    10k hashed points, >= 99% classified correctly."""
    cases = [(geom.open_box(), 0, 42), (geom.uv_sphere(15.0, 48, 48), 1, 1337), (geom.two_cubes(), 2, 9001)]
    for (v, f), mode, seed in cases:
        pts = orc.gwn_test_points(mode, seed)
        w = rhz.winding_numbers(v, f, pts)
        ref = orc.winding_points(v, f, pts)
        assert np.abs(w - ref).max() < 1e-4
        in_t = 0.50 if mode == 1 else 0.45
        out_t = 0.35 if mode == 0 else 0.25
        ok = np.concatenate([w[:5000] >= in_t, w[5000:] < out_t])
        assert ok.mean() >= 0.99, (mode, ok.mean())


def test_eikonal_redblack_bit_exact(rhz, orc):
    rng = np.random.default_rng(5)
    g = rng.uniform(-3, 3, size=(20, 30, 40)).astype(np.float32)
    g[5:15, 5:25, 5:35] = 1e6
    got = rhz.eikonal_redblack(g, 0.5, 8)
    ref = g.copy().reshape(-1)
    orc.lib().orc_eikonal_redblack(ref, np.array([40, 30, 20], dtype=np.int32), 0.5, 8)
    assert np.array_equal(got.reshape(-1).view(np.uint32), ref.view(np.uint32))
