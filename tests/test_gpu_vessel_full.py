"""The vessel at 0.10 mm at full size: 883 x 823 x 923 = 670,752,407 voxels,
199,730 triangles, the default learned build on the device.  Every part of
the result is checked against an independent evaluation:

* signs: the brute-force exact winding kernel (every triangle, compensated
  FP32 with a binary64 recompute where ||w| - 1/2| < 1e-3; no far-field
  certificate) at every candidate voxel (band distance within the scanline
  jitter's reach, the only voxels whose scanline sign can be wrong) and at
  1M jittered stratified voxels: 0 mismatches, voxels on the surface to
  within 1e-5 voxel exempt (w = 1/2 there to rounding);
* band: the device splat equals the oracle's splat bit for bit on two slabs
  of 4 full z-layers (the handle's mid-plane and the cup floor);
* eikonal: |field| equals, bit for bit over all 670.75M voxels, the oracle's
  rounds of the pinned sweep schedule run on the device splat, with the
  round count, the sentinel audit and the clamp of the reference policy.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 0.10


def _log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def vessel(rhz, geom, torch_gpu):
    torch = torch_gpu
    v, f = geom.vessel()
    tris = rhz.build_triangles(v, f)
    mn, mx = geom.frame_grid(v, H)
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    V = int(np.prod(np.array(dims, dtype=np.int64)))
    assert dims == [883, 823, 923] and len(tris) == 199_730
    d_t = torch.from_numpy(tris).cuda()
    d_out = torch.empty(V, dtype=torch.float32, device="cuda")
    p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED)
    t0 = time.perf_counter()
    rc, st = rhz.compute_sdf_device(d_t, len(tris), H, mn, mx, d_out, p)
    torch.cuda.synchronize()
    assert rc == rhz.RHZ_OK, st.as_dict()
    _log(f"vessel 0.10 mm learned build: {time.perf_counter() - t0:.2f} s (first call), {st.log_line()}")
    # the exact band as the pipeline splats it (same band width arithmetic)
    d_band = torch.full((V,), 1e30, dtype=torch.float32, device="cuda")
    bw = float(np.float32(8) * np.float32(H))
    rhz.Stages.splat(d_t, len(tris), H, bw, mn, dims, d_band)
    torch.cuda.synchronize()
    band = d_band.cpu().numpy()
    del d_band
    return dict(tris=tris, d_t=d_t, mn=mn, mx=mx, dims=dims, V=V, d_out=d_out, st=st, band=band)


def test_vessel_signs_match_brute_force_winding(rhz, vessel, torch_gpu):
    torch = torch_gpu
    tris, d_t, mn, mx, dims, V = (vessel[k] for k in ("tris", "d_t", "mn", "mx", "dims", "V"))
    cap = 4_000_000
    d_vox = torch.empty(cap, dtype=torch.int64, device="cuda")
    n = rhz.Proposer.candidates(d_t, len(tris), H, mn, mx, d_vox, None, cap,
                                params=rhz.default_params(method=rhz.METHOD_FAST_LEARNED))
    assert 500_000 < n <= cap
    m = 1_000_000
    rng = np.random.default_rng(2024)
    strat = ((np.arange(m) + rng.random(m)) * (V / m)).astype(np.int64).clip(0, V - 1)
    vox = torch.cat([d_vox[:n], torch.from_numpy(strat).cuda()]).contiguous()
    w = torch.empty(len(vox), dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    rhz.Stages.winding_voxels(d_t, len(tris), H, mn, dims, vox, len(vox), w)
    torch.cuda.synchronize()
    _log(f"brute-force winding: {n} candidates + {m} stratified voxels x {len(tris)} triangles in "
         f"{time.perf_counter() - t0:.2f} s")
    f = vessel["d_out"][vox]
    inside = w.abs() >= 0.5
    surface = f.abs() <= 1e-5 * H
    wrong = (torch.signbit(f) != inside) & ~surface
    _log(f"sign mismatches {int(wrong.sum())}, on-surface exempt {int(surface.sum())}, "
         f"inside {int(inside.sum())}")
    assert int(wrong.sum()) == 0
    assert int(surface.sum()) < 1000


@pytest.mark.parametrize("z0", [461, 12])
def test_vessel_band_matches_oracle(orc, vessel, z0):
    """Four full z-layers of the device splat against the oracle's splat of a
    lattice that holds only those layers (the splat of a voxel depends only on
    its lattice point and the triangles)."""
    dims, mn = vessel["dims"], vessel["mn"]
    layers = 4
    smn = (mn[0], mn[1], mn[2] + z0)
    sdims = np.array([dims[0], dims[1], layers], dtype=np.int32)
    ref = orc.splat_band(vessel["tris"], H, smn, sdims)
    got = vessel["band"].reshape(dims[2], dims[1], dims[0])[z0:z0 + layers].reshape(-1)
    in_band = int((ref < 1e29).sum())
    _log(f"slab z {z0}..{z0 + layers - 1}: {in_band} band voxels")
    assert in_band > 10_000
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))


@pytest.mark.timeout(900)
def test_vessel_eikonal_matches_oracle_rounds(orc, vessel):
    """The reference round policy (rounds until none lowers a voxel, at most
    8; sentinel audit; extension rounds; clamp) in the oracle, from the device
    splat, against the magnitudes of the learned build."""
    st, dims = vessel["st"], vessel["dims"]
    d = vessel["band"].copy()
    t0 = time.perf_counter()
    rounds = 0
    for r in range(8):
        rounds = r + 1
        if not orc.fsm_round(d, dims, H):
            break
    far = int((d >= 1e29).sum())
    _log(f"oracle rounds: {rounds} in {time.perf_counter() - t0:.1f} s "
         f"({orc.lib().orc_num_threads()} threads), {far} voxels left at the sentinel")
    assert rounds == st.eikonal_rounds
    assert far == 0 and st.far_residual == 0 and st.extra_rounds == 0 and st.safe_clamped == 0
    got = vessel["d_out"].abs().cpu().numpy()
    diff = got.view(np.int32) != d.view(np.int32)
    _log(f"voxels differing from the oracle: {int(diff.sum())}")
    assert not diff.any()
