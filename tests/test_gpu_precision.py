"""Precision study: the ico^5 icosphere (20,480 triangles) with 1% of its triangles
split into slivers of aspect ~1e4 (still watertight), 128^3 at 0.5 mm.

* Band distances: the HIP splat (binary32 point-triangle distances) against
  the binary64 band of the oracle from the same binary32 lattice points and
  vertices.  Bar: max error <= 1e-3 voxel.
* Winding numbers: compensated binary32 with binary64 recompute where
  ||w| - 1/2| < 1e-3 (the default) against binary64 sums, at every voxel
  within 2 voxels of the surface plus 20,000 random voxels.  Bar: no inside /
  outside decision differs (voxels on the surface to within 1e-5 voxel
  excepted); the plain binary32 sum is reported beside it.
* The learned path on this mesh: 0 sign mismatches against the plain exact
  path.
The printed report is the study's result."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sliver_ico5_128(geom, orc):
    v, f = geom.perturbed_icosphere(seed=1)
    v, f = geom.with_slivers(v, f, frac=0.01, eps=1e-4, seed=3)
    tris = orc.build_triangles(v, f)
    return tris, 0.5, (-64, -64, -64), (63, 63, 63)


def test_precision_study(rhz, orc, torch_gpu, sliver_ico5_128):
    torch = torch_gpu
    tris, h, mn, mx = sliver_ico5_128
    dims = orc.dims_of(mn, mx)
    V = int(np.prod(dims.astype(np.int64)))
    d_t = torch.from_numpy(tris).cuda()
    d = torch.empty(V, dtype=torch.float32, device="cuda")
    rhz.Stages.fill(d, V, 1e30)
    rhz.Stages.splat(d_t, len(tris), h, 8 * np.float32(h), mn, dims, d)
    torch.cuda.synchronize()
    g32 = d.cpu().numpy()
    d64 = orc.band_f64(tris, h, mn, dims, band_voxels=8)
    both = (g32 < 1e29) & (d64 < 1e29)
    assert np.sum((g32 < 1e29) != (d64 < 1e29)) <= 0.001 * np.sum(both)
    err = np.abs(g32[both].astype(np.float64) - d64[both]) / h
    report = {"band_voxels": int(both.sum()), "band_max_err_voxel": float(err.max()),
              "band_rms_err_voxel": float(np.sqrt(np.mean(err ** 2)))}
    assert report["band_max_err_voxel"] <= 1e-3

    # winding numbers near the surface and at random voxels
    rng = np.random.default_rng(7)
    near = np.flatnonzero(d64 <= 2 * h)
    pick = np.unique(np.concatenate([near, rng.choice(V, 20000, replace=False)])).astype(np.int64)
    ref = orc.winding_at_voxels(tris, h, mn, dims, pick, mode=2)
    d_v = torch.from_numpy(pick).cuda()
    d_w = torch.empty(len(pick), dtype=torch.float64, device="cuda")
    on_surface = d64[pick] <= 1e-5 * h
    for prec, name in ((rhz.WINDING_FP32_KAHAN, "kahan_fp64_near_half"), (rhz.WINDING_FP32, "plain_fp32")):
        rhz.Stages.winding_voxels(d_t, len(tris), h, mn, dims, d_v, len(pick), d_w, precision=prec)
        torch.cuda.synchronize()
        w = d_w.cpu().numpy()
        dw = np.abs(w - ref)
        wrong = ((np.abs(w) >= 0.5) != (np.abs(ref) >= 0.5)) & ~on_surface
        report[name] = {"voxels": int(len(pick)), "max_abs_dw": float(dw.max()),
                        "rms_dw": float(np.sqrt(np.mean(dw ** 2))), "sign_mismatches": int(wrong.sum())}
    assert report["kahan_fp64_near_half"]["sign_mismatches"] == 0

    # the learned path end to end
    rc, got, st = rhz.compute_sdf(tris, h, mn, mx, rhz.default_params(method=rhz.METHOD_FAST_LEARNED))
    rc2, pe, _ = rhz.compute_sdf(tris, h, mn, mx, rhz.default_params(method=rhz.METHOD_PLAIN_EXACT))
    assert rc == rc2 == rhz.RHZ_OK
    got, pe = got.reshape(-1), pe.reshape(-1)
    wrong = (np.signbit(got) != np.signbit(pe)) & (np.abs(pe) > 1e-5 * h)
    report["learned_sign_mismatches_vs_plain_exact"] = int(wrong.sum())
    report["learned_exact_probes"] = int(st.reprobed_voxels)
    print("PRECISION", json.dumps(report))
    assert report["learned_sign_mismatches_vs_plain_exact"] == 0
