"""Batched bake (rhz_sdf_bake): every item's field and counters equal the
single-mesh entry point's, whatever the number of concurrent lanes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 0], ids=["batched_rounds", "lanes"])
def bake_mode(rhz, request):
    """Both bake modes: the eikonal rounds of all lattices as batched
    launches (default), or every item's whole build on a lane."""
    saved = rhz.tuning_get("bake_batched")
    rhz.tuning_set("bake_batched", request.param)
    yield request.param
    rhz.tuning_set("bake_batched", saved)


def test_bake_matches_single_builds(rhz, geom, torch_gpu, bake_mode):
    torch = torch_gpu
    meshes = geom.bake_batch(n=8, seed=77, tri_range=(2_000, 20_000), side_range=(32, 96))
    items, singles = [], []
    for v, f, h in meshes:
        tris = rhz.build_triangles(v, f)
        mn, mx = geom.frame_grid(v, h)
        V = int(np.prod([mx[i] - mn[i] + 1 for i in range(3)]))
        d_t = torch.from_numpy(tris).cuda()
        out = torch.full((V,), np.nan, dtype=torch.float32, device="cuda")
        ref = torch.empty(V, dtype=torch.float32, device="cuda")
        items.append((d_t, len(tris), h, mn, mx, out))
        singles.append(ref)
    p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED)
    ref_stats = []
    for (d_t, nt, h, mn, mx, _), ref in zip(items, singles):
        rc, st = rhz.compute_sdf_device(d_t, nt, h, mn, mx, ref, p, stream=torch.cuda.current_stream())
        assert rc == rhz.RHZ_OK
        ref_stats.append(st)
    torch.cuda.synchronize()
    for lanes in (1, 3):
        for it in items:
            it[5].fill_(float("nan"))
        torch.cuda.synchronize()
        rc, stats, ms = rhz.bake(items, p, n_lanes=lanes)
        assert rc == rhz.RHZ_OK
        for it, ref, st, rs, t in zip(items, singles, stats, ref_stats, ms):
            assert torch.equal(it[5].view(torch.int32), ref.view(torch.int32))
            assert st.status == rhz.RHZ_OK and t > 0
            for k in ("eikonal_rounds", "candidates", "risky_candidates", "reprobed_voxels", "mismatches_found"):
                assert getattr(st, k) == getattr(rs, k), k


def test_bake_after_growth_and_release(rhz, geom, torch_gpu, bake_mode):
    """Lanes that outgrow their workspace retire the old buffers and free them
    after the call; a batch of larger lattices after a small one, and a bake
    after the workspaces were released, still equal the single builds."""
    torch = torch_gpu
    p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED)

    def batch(seed, side_range):
        items, refs = [], []
        for v, f, h in geom.bake_batch(n=6, seed=seed, tri_range=(2_000, 12_000), side_range=side_range):
            tris = rhz.build_triangles(v, f)
            mn, mx = geom.frame_grid(v, h)
            V = int(np.prod([mx[i] - mn[i] + 1 for i in range(3)]))
            d_t = torch.from_numpy(tris).cuda()
            ref = torch.empty(V, dtype=torch.float32, device="cuda")
            rc, _ = rhz.compute_sdf_device(d_t, len(tris), h, mn, mx, ref, p, stream=torch.cuda.current_stream())
            assert rc == rhz.RHZ_OK
            items.append((d_t, len(tris), h, mn, mx, torch.full((V,), np.nan, dtype=torch.float32, device="cuda")))
            refs.append(ref)
        torch.cuda.synchronize()
        return items, refs

    def check(items, refs):
        rc, stats, _ = rhz.bake(items, p, n_lanes=3)
        assert rc == rhz.RHZ_OK
        for it, ref in zip(items, refs):
            assert torch.equal(it[5].view(torch.int32), ref.view(torch.int32))

    rhz.release_workspace()
    small = batch(91, (24, 40))
    check(*small)
    large = batch(92, (64, 112))
    check(*large)
    rhz.release_workspace()
    for it in large[0]:
        it[5].fill_(float("nan"))
    torch.cuda.synchronize()
    check(*large)
    check(*small)


def test_bake_rejects_bad_lane_count(rhz, geom, torch_gpu):
    with pytest.raises(rhz.RhzError):
        rhz.bake([], rhz.default_params(), n_lanes=0)


@pytest.mark.parametrize("case", ["small_sphere", "ico5_128"])
def test_fsm_batched_rounds_and_graph_match_round_loop(rhz, torch_gpu, request, case):
    """All FSM rounds enqueued at once, and the same rounds captured and
    replayed as a hipGraph, give the round-by-round loop's field and round
    count (rounds after a round without change change nothing)."""
    torch = torch_gpu
    tris, h, mn, mx = request.getfixturevalue(case)
    V = int(np.prod([mx[i] - mn[i] + 1 for i in range(3)]))
    d_t = torch.from_numpy(tris).cuda()
    stream = torch.cuda.Stream()
    saved = {k: rhz.tuning_get(k) for k in ("fsm_batch_rounds_max_voxels", "fsm_graph")}
    outs, rounds = [], []
    try:
        for mv, g in ((0, 0), (1 << 40, 0), (1 << 40, 1), (1 << 40, 1)):
            rhz.tuning_set("fsm_batch_rounds_max_voxels", mv)
            rhz.tuning_set("fsm_graph", g)
            out = torch.empty(V, dtype=torch.float32, device="cuda")
            with torch.cuda.stream(stream):
                rc, st = rhz.compute_sdf_device(d_t, len(tris), h, mn, mx, out,
                                                rhz.default_params(method=rhz.METHOD_FAST_LEARNED), stream=stream)
            stream.synchronize()
            assert rc == rhz.RHZ_OK
            outs.append(out)
            rounds.append(st.eikonal_rounds)
    finally:
        for k, v in saved.items():
            rhz.tuning_set(k, v)
    for o in outs[1:]:
        assert torch.equal(o.view(torch.int32), outs[0].view(torch.int32))
    assert len(set(rounds)) == 1, rounds


@pytest.mark.timeout(600)
def test_full_batch_matches_single_builds_and_winding(rhz, geom, torch_gpu, bake_mode):
    """The full batched-bake workload: 64 seeded vessels (2k-500k triangles,
    longest side 64-384 at 0.25-0.5 mm; lattices of >= 16M voxels run the
    overlapped sign stage inside their lane).  Every field equals its single
    build bit for bit, and at every candidate voxel of every item (the only
    voxels whose scanline sign can be wrong) the sign equals the brute-force
    exact winding sign (on-surface voxels within 1e-5 voxel exempt)."""
    torch = torch_gpu
    p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED)
    items, refs, total = [], [], 0
    for v, f, h in geom.bake_batch(n=64):
        tris = rhz.build_triangles(v, f)
        assert 2_000 <= len(f) <= 500_000
        mn, mx = geom.frame_grid(v, h)
        V = int(np.prod([mx[i] - mn[i] + 1 for i in range(3)]))
        total += V
        d_t = torch.from_numpy(tris).cuda()
        ref = torch.empty(V, dtype=torch.float32, device="cuda")
        rc, _ = rhz.compute_sdf_device(d_t, len(tris), h, mn, mx, ref, p)
        assert rc == rhz.RHZ_OK
        items.append((d_t, len(tris), h, mn, mx, torch.full((V,), np.nan, dtype=torch.float32, device="cuda")))
        refs.append(ref)
    torch.cuda.synchronize()
    assert max(it[5].numel() for it in items) >= 16 << 20
    rc, stats, ms = rhz.bake(items, p, n_lanes=8)
    assert rc == rhz.RHZ_OK
    print(f"bake: {len(items)} items, {total} voxels, slowest item {max(ms):.1f} ms", flush=True)
    cap = 1 << 22
    d_vox = torch.empty(cap, dtype=torch.int64, device="cuda")
    checked = 0
    for k, ((d_t, nt, h, mn, mx, out), ref, st) in enumerate(zip(items, refs, stats)):
        assert st.status == rhz.RHZ_OK
        assert torch.equal(out.view(torch.int32), ref.view(torch.int32)), k
        dims = [mx[i] - mn[i] + 1 for i in range(3)]
        n = rhz.Proposer.candidates(d_t, nt, h, mn, mx, d_vox, None, cap, params=p)
        assert 0 < n <= cap
        w = torch.empty(n, dtype=torch.float64, device="cuda")
        rhz.Stages.winding_voxels(d_t, nt, h, mn, dims, d_vox, n, w)
        fv = out[d_vox[:n]]
        wrong = (torch.signbit(fv) != (w.abs() >= 0.5)) & (fv.abs() > 1e-5 * h)
        assert int(wrong.sum()) == 0, (k, int(wrong.sum()))
        checked += n
    print(f"candidate voxels checked against brute-force winding: {checked}", flush=True)
