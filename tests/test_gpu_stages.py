"""Stage-level parity of the HIP kernels against the CPU oracle, through the
C-ABI stage entry points (device buffers from torch).

Integer/byte work (signs, indices, counts) is compared bit-exactly.  The band
distances and the FSM sweeps are bit-exact too: the kernels and the oracle
evaluate the same IEEE binary32 expressions (no contraction) with the same
sweep schedule."""
import numpy as np
import pytest

from conftest import dims_of

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _host(t):
    return t.detach().cpu().numpy()


def test_fill_and_count_far(rhz, torch_gpu):
    torch = torch_gpu
    for n in (1, 3, 17, 1 << 20, (1 << 20) + 5):
        d = torch.empty(n, dtype=torch.float32, device="cuda")
        rhz.Stages.fill(d, n, 1e30)
        torch.cuda.synchronize()
        assert bool((d == 1e30).all())
        assert rhz.Stages.count_far(d, n) == n
        d[: n // 2] = 3.0
        assert rhz.Stages.count_far(d, n) == n - n // 2


def test_splat_bit_exact(rhz, orc, torch_gpu, small_sphere):
    torch = torch_gpu
    tris, h, mn, mx = small_sphere
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    ref = orc.splat_band(tris, h, mn, dims)
    dt = _dev(torch, tris)
    d = torch.empty(n, dtype=torch.float32, device="cuda")
    rhz.Stages.fill(d, n, 1e30)
    rhz.Stages.splat(dt, len(tris), h, np.float32(8) * np.float32(h), mn, dims, d)
    torch.cuda.synchronize()
    got = _host(d)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert (got < 1e29).sum() > 0


def test_splat_grid_clipping(rhz, orc, torch_gpu, small_sphere):
    """Grid that cuts through the mesh (triangles partly or wholly outside)."""
    torch = torch_gpu
    tris, h, _, _ = small_sphere
    mn, mx = (-5, -30, 3), (40, 2, 30)
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    ref = orc.splat_band(tris, h, mn, dims)
    d = torch.empty(n, dtype=torch.float32, device="cuda")
    rhz.Stages.fill(d, n, 1e30)
    rhz.Stages.splat(_dev(torch, tris), len(tris), h, np.float32(8) * np.float32(h), mn, dims, d)
    torch.cuda.synchronize()
    assert np.array_equal(_host(d).view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("grid", ["small", "clipped"])
def test_parity_bit_exact(rhz, orc, torch_gpu, small_sphere, grid):
    torch = torch_gpu
    tris, h, mn, mx = small_sphere
    if grid == "clipped":
        mn, mx = (-5, -30, 3), (40, 2, 30)
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    ref, flag = orc.parity_signs(tris, h, mn, dims)
    s = torch.zeros(n, dtype=torch.uint8, device="cuda")
    rc, gflag, entries = rhz.Stages.parity(_dev(torch, tris), len(tris), h, mn, dims, s)
    assert rc == rhz.RHZ_OK
    assert gflag == flag
    assert entries == orc.lib().orc_csr_entries(orc._f32(tris).reshape(-1), len(tris), h, 4, orc._i32(mn),
                                                 orc._i32(dims))
    assert np.array_equal(_host(s), ref)


def test_parity_open_mesh_flags_columns(rhz, orc, torch_gpu, geom):
    """The reference's open box (no top face): every column through the
    opening has an odd hit count and is deferred (sign 2)."""
    torch = torch_gpu
    v, f = geom.open_box()
    tris = orc.build_triangles(v, f)
    h = 1.0
    mn, mx = (-14, -14, -14), (13, 13, 13)
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    ref, flag = orc.parity_signs(tris, h, mn, dims)
    assert flag == 1
    s = torch.zeros(n, dtype=torch.uint8, device="cuda")
    rc, gflag, _ = rhz.Stages.parity(_dev(torch, tris), len(tris), h, mn, dims, s)
    assert gflag == 1
    got = _host(s)
    assert np.array_equal(got, ref)
    assert (got == 2).sum() > 0


def test_parity_overflow_is_bad_column(rhz, orc, torch_gpu, small_sphere):
    torch = torch_gpu
    tris, h, mn, mx = small_sphere
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    ref, flag = orc.parity_signs(tris, h, mn, dims, max_hits=1)
    s = torch.zeros(n, dtype=torch.uint8, device="cuda")
    _, gflag, _ = rhz.Stages.parity(_dev(torch, tris), len(tris), h, mn, dims, s, max_hits=1)
    assert flag == gflag == 1
    assert np.array_equal(_host(s), ref)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("direction", [1, -1])
def test_sweep_bit_exact(rhz, orc, torch_gpu, small_sphere, axis, direction):
    torch = torch_gpu
    tris, h, mn, mx = small_sphere
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    src = orc.splat_band(tris, h, mn, dims)
    # a second sweep on an already-propagated field exercises the 2-D/3-D branches
    tmp = np.empty_like(src)
    orc.lib().orc_fsm_sweep_db(src, tmp, orc._i32(dims), h, 2, 1)
    src = tmp.copy()
    ref = np.empty_like(src)
    ch_ref = orc.lib().orc_fsm_sweep_db(src, ref, orc._i32(dims), h, axis, direction)
    ds = _dev(torch, src)
    dd = torch.empty_like(ds)
    ch = torch.zeros(1, dtype=torch.int32, device="cuda")
    rhz.Stages.sweep(ds, dd, dims, h, axis, direction, ch)
    torch.cuda.synchronize()
    got = _host(dd)
    assert int(ch.item()) == ch_ref
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
        f"{(got != ref).sum()} voxels differ, max {np.abs(got - ref).max()}"


def test_sweep_sqrt_is_ieee_over_its_whole_range(rhz):
    """The lane sweeps take their square roots as a hardware reciprocal
    square root plus one Newton step; on this device that equals IEEE sqrt
    (binary32, and binary64 rounded) for every binary32 argument in
    [2^-96, inf), which is every argument the sweeps pass it."""
    assert rhz.selftest_sweep_sqrt() == 0


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 7, 5), (70, 3, 2), (65, 66, 3), (5, 130, 67)])
def test_sweep_odd_shapes(rhz, orc, torch_gpu, shape):
    """Ragged shapes: partial waves, partial x-chunks, single-voxel axes."""
    torch = torch_gpu
    dx, dy, dz = shape
    rng = np.random.default_rng(7)
    src = np.full(dx * dy * dz, 1e30, dtype=np.float32)
    seeds = rng.choice(src.size, size=max(1, src.size // 50), replace=False)
    src[seeds] = rng.uniform(0.0, 2.0, size=len(seeds)).astype(np.float32)
    dims = np.array(shape, dtype=np.int32)
    for axis in range(3):
        for direction in (1, -1):
            ref = np.empty_like(src)
            orc.lib().orc_fsm_sweep_db(src, ref, orc._i32(dims), 0.25, axis, direction)
            ds = _dev(torch, src)
            dd = torch.empty_like(ds)
            ch = torch.zeros(1, dtype=torch.int32, device="cuda")
            rhz.Stages.sweep(ds, dd, dims, 0.25, axis, direction, ch)
            torch.cuda.synchronize()
            assert np.array_equal(_host(dd).view(np.uint32), ref.view(np.uint32)), (shape, axis, direction)


@pytest.mark.parametrize("shape,step", [(21, 2), (21, 4), (21, 8), (22, 2), (22, 8), (12, 2), (12, 4), (14, 2),
                                        (14, 8), (1, 4), (1, 8)])
def test_sweep_tunings_bit_exact(rhz, orc, torch_gpu, shape, step):
    """Every sweep launch shape equals the oracle bit for bit, on a lattice
    with interior tiles (no lattice face) as well as ragged face tiles, and at
    a voxel size below the short square root's range (generic kernels)."""
    torch = torch_gpu
    keys = ("sweep_shape_x", "sweep_shape_y", "sweep_shape_z", "sweep_step")
    saved = {k: rhz.tuning_get(k) for k in keys}
    for k in keys[:3]:
        rhz.tuning_set(k, shape)
    rhz.tuning_set("sweep_step", step)
    try:
        rng = np.random.default_rng(11)
        for shape, h in (((300, 71, 41), 0.2), ((130, 9, 7), 1e-10)):
            dx, dy, dz = shape
            src = np.full(dx * dy * dz, 1e30, dtype=np.float32)
            seeds = rng.choice(src.size, size=max(1, src.size // 60), replace=False)
            src[seeds] = (rng.uniform(0.0, 4.0, size=len(seeds)) * h / 0.2).astype(np.float32)
            dims = np.array(shape, dtype=np.int32)
            # one propagated round first, so 2-D and 3-D updates occur
            tmp = np.empty_like(src)
            orc.lib().orc_fsm_sweep_db(src, tmp, orc._i32(dims), h, 2, 1)
            src = tmp
            for axis in range(3):
                for direction in (1, -1):
                    ref = np.empty_like(src)
                    ch_ref = orc.lib().orc_fsm_sweep_db(src, ref, orc._i32(dims), h, axis, direction)
                    ds = _dev(torch, src)
                    dd = torch.empty_like(ds)
                    ch = torch.zeros(1, dtype=torch.int32, device="cuda")
                    rhz.Stages.sweep(ds, dd, dims, h, axis, direction, ch)
                    torch.cuda.synchronize()
                    got = _host(dd)
                    assert int(ch.item()) == ch_ref, (shape, axis, direction)
                    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
                        (shape, axis, direction, int((got != ref).sum()))
            # a full round: the x-sweeps run on the transposed copy
            ref = src.copy()
            ch_ref = 0
            for axis in range(3):
                for direction in (1, -1):
                    ch_ref |= orc.lib().orc_fsm_sweep(ref, orc._i32(dims), h, axis, direction)
            a = _dev(torch, src)
            b = torch.empty_like(a)
            ch = torch.zeros(1, dtype=torch.int32, device="cuda")
            rhz.Stages.sweep_round(a, b, dims, h, ch)
            torch.cuda.synchronize()
            assert int(ch.item()) == ch_ref
            assert np.array_equal(_host(a).view(np.uint32), ref.view(np.uint32)), (shape, "round")
    finally:
        for k, v in saved.items():
            rhz.tuning_set(k, v)


@pytest.mark.parametrize("shape", [(2, 2, 2), (65, 66, 3), (130, 67, 9), (883, 10, 4)])
def test_sweep_round_bit_exact(rhz, orc, torch_gpu, shape):
    """A full round as the pipeline runs it (x-sweeps on the transposed copy)
    equals the oracle's six sweeps bit for bit, over three rounds."""
    torch = torch_gpu
    dx, dy, dz = shape
    rng = np.random.default_rng(9)
    src = np.full(dx * dy * dz, 1e30, dtype=np.float32)
    seeds = rng.choice(src.size, size=max(1, src.size // 40), replace=False)
    src[seeds] = rng.uniform(0.0, 3.0, size=len(seeds)).astype(np.float32)
    dims = np.array(shape, dtype=np.int32)
    a = _dev(torch, src)
    b = torch.empty_like(a)
    ref = src.copy()
    for r in range(3):
        ch_ref = 0
        for axis in range(3):
            for direction in (1, -1):
                ch_ref |= orc.lib().orc_fsm_sweep(ref, orc._i32(dims), 0.3, axis, direction)
        ch = torch.zeros(1, dtype=torch.int32, device="cuda")
        rhz.Stages.sweep_round(a, b, dims, 0.3, ch)
        torch.cuda.synchronize()
        assert int(ch.item()) == ch_ref
        assert np.array_equal(_host(a).view(np.uint32), ref.view(np.uint32)), (shape, r)


def test_clamp_and_apply(rhz, orc, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(3)
    n = 100003
    d = rng.uniform(0, 5, n).astype(np.float32)
    d[rng.choice(n, 1000, replace=False)] = 1e30
    s = rng.integers(0, 2, n).astype(np.uint8)
    for intended in (0, 1):
        ref = d.copy()
        orc.lib().orc_clamp_far(ref, s, n, 4.0, intended)
        orc.lib().orc_apply_signs(ref, s, n)
        dd = _dev(torch, d)
        ds = _dev(torch, s)
        rhz.Stages.clamp_far(dd, ds, n, 4.0, intended)
        rhz.Stages.apply(dd, ds, n)
        torch.cuda.synchronize()
        assert np.array_equal(_host(dd).view(np.uint32), ref.view(np.uint32))


def test_mark_patch(rhz, orc, torch_gpu):
    torch = torch_gpu
    dims = np.array([61, 50, 77], dtype=np.int32)
    n = int(np.prod(dims))
    sites = np.array([0, n - 1, 12345, 200000, 12346], dtype=np.int64)
    ref = np.zeros(n, dtype=np.uint8)
    orc.lib().orc_mark_patch(ref, sites, len(sites), 24, orc._i32(dims))
    s = torch.zeros(n, dtype=torch.uint8, device="cuda")
    rhz.Stages.mark_patch(s, _dev(torch, sites), len(sites), 24, dims)
    torch.cuda.synchronize()
    assert np.array_equal(_host(s), ref)


def test_winding_voxels_decisions(rhz, orc, torch_gpu, small_sphere):
    torch = torch_gpu
    tris, h, mn, mx = small_sphere
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    rng = np.random.default_rng(11)
    vox = np.sort(rng.choice(n, 3000, replace=False)).astype(np.int64)
    ref = orc.winding_at_voxels(tris, h, mn, dims, vox, mode=2)
    w = torch.zeros(len(vox), dtype=torch.float64, device="cuda")
    rhz.Stages.winding_voxels(_dev(torch, tris), len(tris), h, mn, dims, _dev(torch, vox), len(vox), w)
    torch.cuda.synchronize()
    got = _host(w)
    assert np.array_equal(np.abs(got) >= 0.5, np.abs(ref) >= 0.5)
    assert np.abs(got - ref).max() < 1e-4


def test_fixup_matches_oracle(rhz, orc, torch_gpu, geom):
    torch = torch_gpu
    v, f = geom.open_box()
    tris = orc.build_triangles(v, f)
    h = 1.0
    mn, mx = (-14, -14, -14), (13, 13, 13)
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    ref, _ = orc.parity_signs(tris, h, mn, dims)
    k_ref = orc.lib().orc_winding_fixup(orc._f32(tris).reshape(-1), len(tris), h, orc._i32(mn), orc._i32(dims),
                                        ref, 3)
    s = torch.zeros(n, dtype=torch.uint8, device="cuda")
    rhz.Stages.parity(_dev(torch, tris), len(tris), h, mn, dims, s)
    k = rhz.Stages.fixup(_dev(torch, tris), len(tris), h, mn, dims, s)
    assert k == k_ref > 0
    got = _host(s)
    # The box is aligned with the lattice, so two kinds of voxel have no
    # well-defined sign: voxels on a face (distance 0 up to the rounding of
    # the face coordinates, where w jumps and the coplanar triangles
    # contribute +-2 pi by the sign of a zero determinant) and voxels on the
    # plane of the missing face (w = 1/2 exactly by symmetry).  Any other
    # difference is a failure.
    diff = np.nonzero(got != ref)[0]
    if len(diff):
        d = orc.splat_band(tris, h, mn, dims)[diff]
        w = orc.winding_at_voxels(tris, h, mn, dims, diff, mode=2)
        tie = np.abs(np.abs(w) - 0.5) < 1e-12
        assert np.all((d <= 1e-5 * h) | tie), (d, w)


@pytest.mark.parametrize("gen", [0, 1, 2])
def test_verify_probe(rhz, orc, torch_gpu, small_sphere, gen):
    torch = torch_gpu
    tris, h, mn, mx = small_sphere
    dims = dims_of(mn, mx)
    n = int(np.prod(dims))
    sign, _ = orc.parity_signs(tris, h, mn, dims)
    # plant wrong signs at some probe positions: the probe must report exactly them
    idx = np.zeros(4096, dtype=np.int64)
    orc.lib().orc_verify_indices(n, 4096, gen, idx)
    exact = np.abs(orc.winding_at_voxels(tris, h, mn, dims, idx, mode=3)) >= 0.5
    planted = sign.copy()
    flip = idx[::97]
    planted[flip] = 1 - planted[flip]
    expected = set(int(v) for v, e in zip(idx, exact) if (planted[v] == 1) != e)
    s = _dev(torch, planted)
    sites = torch.zeros(4096, dtype=torch.int64, device="cuda")
    m = rhz.Stages.verify(_dev(torch, tris), len(tris), h, mn, dims, 4096, gen, s, sites)
    assert m == len(expected)
    assert set(_host(sites)[:m].tolist()) == expected
