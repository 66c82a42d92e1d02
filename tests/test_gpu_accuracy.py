"""The accuracy report (rhz_sdf_accuracy) against a closed form: for the cube
[-10, 10]^3 the distance from a point to the surface is max(|p| - 10, 0)
component-wise under the Euclidean norm outside, and 10 - max |p_i| inside.
The sample positions are recomputed here from the documented generator, so
every figure of the report is checked against an independent evaluation."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def sample_ids(V, n, seed):
    """Jittered stratified ids: sample i in [i V / n, (i + 1) V / n) at a
    splitmix64 position of (seed, i)."""
    out = np.empty(n, dtype=np.int64)
    for i in range(n):
        z = (seed + 0x9E3779B97F4A7C15 * (i + 1)) & M64
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        z ^= z >> 31
        u = float(z >> 11) * (1.0 / 9007199254740992.0)
        out[i] = min(max(int((float(i) + u) * float(V) / float(n)), 0), V - 1)
    return out


def cube_distance(p):
    q = np.abs(p) - 10.0
    outside = np.sqrt((np.maximum(q, 0.0) ** 2).sum(1))
    inside = -q.max(1)
    return np.where((q > 0).any(1), outside, inside)


def test_accuracy_report_matches_closed_form(rhz, geom, torch_gpu):
    torch = torch_gpu
    v, f = geom.two_cubes()
    v, f = v[:8], f[:12]  # the first cube alone
    tris = rhz.build_triangles(v, f)
    h = 0.5
    mn, mx = (-32, -32, -32), (31, 31, 31)
    dims = [64, 64, 64]
    V = 64 ** 3
    d_t = torch.from_numpy(tris).cuda()
    d_out = torch.empty(V, dtype=torch.float32, device="cuda")
    rc, _ = rhz.compute_sdf_device(d_t, len(tris), h, mn, mx, d_out,
                                   rhz.default_params(method=rhz.METHOD_FAST_LEARNED))
    assert rc == rhz.RHZ_OK
    torch.cuda.synchronize()
    n, seed = 20_000, 7
    acc = rhz.accuracy(d_t, len(tris), h, mn, mx, d_out, n_samples=n, band_voxels=8, seed=seed)

    ids = sample_ids(V, n, seed)
    x, y, z = ids % dims[0], (ids // dims[0]) % dims[1], ids // (dims[0] * dims[1])
    h32 = np.float32(h)
    p = np.stack([(np.int32(mn[0]) + x).astype(np.float32) * h32, (np.int32(mn[1]) + y).astype(np.float32) * h32,
                  (np.int32(mn[2]) + z).astype(np.float32) * h32], 1).astype(np.float64)
    truth = cube_distance(p)
    fv = d_out.cpu().numpy()[ids].astype(np.float64)
    inside_true = (np.abs(p) < 10.0).all(1)
    surface = truth <= 1e-5 * h
    e = (np.abs(fv) - truth) / h
    band = truth <= 8 * h
    assert acc.samples == n
    assert acc.on_surface == int(surface.sum())
    assert acc.sign_mismatches == int(((fv < 0) != inside_true)[~surface].sum()) == 0
    assert acc.unreached == 0
    assert acc.band_samples == int(band.sum()) and acc.far_samples == int((~band).sum())
    for got, want in ((acc.band_max_err, np.abs(e[band]).max()), (acc.band_rms_err, np.sqrt((e[band] ** 2).mean())),
                      (acc.band_mean_err, e[band].mean()), (acc.far_max_err, np.abs(e[~band]).max()),
                      (acc.far_rms_err, np.sqrt((e[~band] ** 2).mean())), (acc.far_mean_err, e[~band].mean())):
        assert abs(got - want) <= 1e-9, (got, want)
    assert acc.band_max_err < 1.0
