"""rhz_stage_winding_voxels_noalloc against rhz_stage_winding_voxels: the
same winding numbers bit for bit, on two calls in a row (the second one
smaller, so the fixed scratch holds the first call's counter, list and
partials when it starts), in both precisions; a call beyond the fixed
scratch is rejected on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_entry_point(rhz, ico5_128, torch_gpu, precision):
    torch = torch_gpu
    tris, h, mn, mx = ico5_128
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    V = int(np.prod(dims))
    d_t = torch.from_numpy(tris).cuda()
    rng = np.random.default_rng(31 + precision)
    # voxels near the surface (where binary64 recomputes happen) and anywhere
    p = np.stack(np.meshgrid(*[np.arange(d) for d in dims[::-1]], indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    r = np.linalg.norm((p + np.array(mn)) * h, axis=1)
    near = np.flatnonzero(np.abs(r - 20.0) < 2.5)
    for n in (200_000, 5_000):
        ids = np.concatenate([rng.choice(near, n // 2, replace=False), rng.choice(V, n - n // 2, replace=False)])
        d_ids = torch.from_numpy(ids.astype(np.int64)).cuda()
        ref = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
        got = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
        rhz.Stages.winding_voxels(d_t, len(tris), h, mn, dims, d_ids, n, ref, precision=precision)
        rhz.Stages.winding_voxels_noalloc(d_t, len(tris), h, mn, dims, d_ids, n, got, precision=precision)
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.int64), ref.view(torch.int64)), n
        assert not torch.isnan(got).any()


def test_noalloc_rejects_more_than_its_scratch(rhz, ico5_128, torch_gpu):
    torch = torch_gpu
    tris, h, mn, mx = ico5_128
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    d_t = torch.from_numpy(tris).cuda()
    n = (1 << 20) + 1
    d_ids = torch.zeros(n, dtype=torch.int64, device="cuda")
    w = torch.empty(n, dtype=torch.float64, device="cuda")
    with pytest.raises(rhz.RhzError, match="2\\^20"):
        rhz.Stages.winding_voxels_noalloc(d_t, len(tris), h, mn, dims, d_ids, n, w)
