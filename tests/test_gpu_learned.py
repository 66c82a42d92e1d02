"""The learned sign gate (FAST_LEARNED) and the proposer's HIP kernels.

* MLP forward / gradient / Adam kernels against the CPU oracle (oracle/mlp.py,
  which emulates the kernels' BF16 operand rounding; sums are FP32 on the GPU
  and FP64 in the oracle, so logits agree to within accumulation order and
  rare one-ulp BF16 flips of an activation: max |dz| <= 2e-2 (1 + |z|),
  median <= 1e-5 (1 + |z|); gradients per parameter block within 2e-2
  relative L2, cosine >= 0.9995).
* The learned method against the plain exact-winding method on the same
  lattice: signs identical (0 mismatches; only voxels lying on the surface
  to within 1e-5 voxel may differ, where the exact winding number is 1/2 to
  rounding), distance magnitudes bit-identical (same band and sweeps).
"""
import numpy as np
import pytest

from conftest import dims_of
from oracle import mlp

pytestmark = pytest.mark.gpu


def _blob(seed):
    rng = np.random.default_rng(seed)
    b = mlp.init_blob(seed, mean=rng.normal(size=12) * 0.1, istd=1.0 + rng.random(12))
    b[mlp.OFF_B1:mlp.OFF_W2] = rng.normal(size=64) * 0.1
    b[mlp.OFF_B2:mlp.OFF_W3] = rng.normal(size=64) * 0.1
    b[mlp.OFF_B3:mlp.OFF_W4] = rng.normal(size=64) * 0.1
    b[mlp.OFF_B4] = -0.3
    return b


@pytest.mark.parametrize("n", [1, 16, 1003, 70001])
def test_mlp_forward_matches_oracle(rhz, torch_gpu, n):
    torch = torch_gpu
    blob = _blob(11)
    X = np.random.default_rng(n).normal(size=(n, 12)).astype(np.float32)
    d_b = torch.from_numpy(blob).cuda()
    d_x = torch.from_numpy(X).cuda()
    d_z = torch.full((n,), np.nan, dtype=torch.float32, device="cuda")
    rhz.Proposer.forward(d_b, d_x, n, d_z)
    torch.cuda.synchronize()
    z = d_z.cpu().numpy().astype(np.float64)
    ref, _ = mlp.forward(blob, X, bf16=True)
    err = np.abs(z - ref) / (1 + np.abs(ref))
    assert np.all(np.isfinite(z))
    assert err.max() <= 2e-2, err.max()
    assert np.median(err) <= 1e-5, np.median(err)


@pytest.mark.parametrize("n", [37, 4099])
def test_mlp_grad_matches_oracle(rhz, torch_gpu, n):
    torch = torch_gpu
    blob = _blob(5)
    rng = np.random.default_rng(n)
    X = rng.normal(size=(n, 12)).astype(np.float32)
    Y = (rng.random(n) < 0.2).astype(np.float32)
    d_b = torch.from_numpy(blob).cuda()
    d_x = torch.from_numpy(X).cuda()
    d_y = torch.from_numpy(Y).cuda()
    d_g = torch.zeros(mlp.N_PARAM, dtype=torch.float32, device="cuda")
    d_l = torch.zeros(1, dtype=torch.float32, device="cuda")
    rhz.Proposer.grad(d_b, d_x, d_y, n, 4.0, d_g, d_l)
    torch.cuda.synchronize()
    g = d_g.cpu().numpy().astype(np.float64)
    loss = float(d_l.cpu().numpy()[0])
    ref_loss, ref_g = mlp.loss_and_grad(blob, X, Y, 4.0, bf16=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    blocks = [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_B2, mlp.OFF_W3),
              (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.OFF_W4), (mlp.OFF_W4, mlp.OFF_B4),
              (mlp.OFF_B4, mlp.N_PARAM)]
    for a, b in blocks:
        ga, ra = g[a:b], ref_g[a:b]
        rel = np.linalg.norm(ga - ra) / max(np.linalg.norm(ra), 1e-12)
        assert rel <= 2e-2, (a, b, rel)
    cos = np.dot(g, ref_g) / (np.linalg.norm(g) * np.linalg.norm(ref_g))
    assert cos >= 0.9995, cos


def test_mlp_grad_accumulates(rhz, torch_gpu):
    """Two calls add: grad(A) + grad(B) == grad(A u B) up to summation order."""
    torch = torch_gpu
    blob = _blob(7)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(512, 12)).astype(np.float32)
    Y = (rng.random(512) < 0.5).astype(np.float32)
    d_b = torch.from_numpy(blob).cuda()
    d_x = torch.from_numpy(X).cuda()
    d_y = torch.from_numpy(Y).cuda()
    g1 = torch.zeros(mlp.N_PARAM, dtype=torch.float32, device="cuda")
    g2 = torch.zeros_like(g1)
    l1 = torch.zeros(1, dtype=torch.float32, device="cuda")
    l2 = torch.zeros_like(l1)
    rhz.Proposer.grad(d_b, d_x, d_y, 512, 1.0, g1, l1)
    rhz.Proposer.grad(d_b, d_x[:200], d_y[:200], 200, 1.0, g2, l2)
    rhz.Proposer.grad(d_b, d_x[200:], d_y[200:], 312, 1.0, g2, l2)
    torch.cuda.synchronize()
    a, b = g1.cpu().numpy(), g2.cpu().numpy()
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(a)


def test_adam_matches_oracle(rhz, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(2)
    blob = _blob(3)
    m = np.zeros(mlp.N_PARAM, np.float32)
    v = np.zeros(mlp.N_PARAM, np.float32)
    d_b = torch.from_numpy(blob.copy()).cuda()
    d_m = torch.zeros(mlp.N_PARAM, dtype=torch.float32, device="cuda")
    d_v = torch.zeros_like(d_m)
    ref = blob.copy()
    for step in range(1, 4):
        g = rng.normal(size=mlp.N_PARAM).astype(np.float32)
        d_g = torch.from_numpy(g).cuda()
        rhz.Proposer.adam(d_b, d_g, d_m, d_v, mlp.N_PARAM, lr=1e-3, step=step, grad_scale=0.5, weight_decay=1e-4)
        mlp.adam(ref, g, m, v, mlp.N_PARAM, lr=1e-3, step=step, grad_scale=0.5, weight_decay=1e-4)
    torch.cuda.synchronize()
    got = d_b.cpu().numpy()
    assert np.allclose(got[:mlp.N_PARAM], ref[:mlp.N_PARAM], rtol=1e-5, atol=1e-7)
    # the normalisation statistics behind the parameters are not touched
    assert np.array_equal(got[mlp.N_PARAM:], blob[mlp.N_PARAM:])


def _compare_learned(rhz, tris, h, mn, mx, **kw):
    pl = rhz.default_params(method=rhz.METHOD_FAST_LEARNED, **kw)
    rc, got, st = rhz.compute_sdf(tris, h, mn, mx, pl)
    assert rc == rhz.RHZ_OK, st.as_dict()
    assert st.method_used == rhz.METHOD_FAST_LEARNED
    pe = rhz.default_params(method=rhz.METHOD_PLAIN_EXACT)
    rc, ref, est = rhz.compute_sdf(tris, h, mn, mx, pe)
    assert rc == rhz.RHZ_OK
    got = got.reshape(-1)
    ref = ref.reshape(-1)
    assert np.array_equal(np.abs(got).view(np.uint32), np.abs(ref).view(np.uint32))
    wrong = np.signbit(got) != np.signbit(ref)
    on_surface = np.abs(ref) <= 1e-5 * h
    assert not np.any(wrong & ~on_surface), f"{int(np.sum(wrong & ~on_surface))} sign mismatches; {st.as_dict()}"
    return st, est, int(wrong.sum())


@pytest.mark.parametrize("scorer", [0, 1])
def test_learned_small_sphere(rhz, small_sphere, scorer):
    tris, h, mn, mx = small_sphere
    st, est, _ = _compare_learned(rhz, tris, h, mn, mx, scorer=scorer)
    assert st.candidates > 0
    assert st.reprobed_voxels < est.reprobed_voxels


def test_learned_ico5_128(rhz, ico5_128):
    tris, h, mn, mx = ico5_128
    st, est, _ = _compare_learned(rhz, tris, h, mn, mx)
    V = int(np.prod(dims_of(mn, mx).astype(np.int64)))
    # the exact kernel probes a small fraction of what the plain path probes
    assert st.reprobed_voxels <= 0.02 * V
    print(st.log_line(), "reprobed", st.reprobed_voxels, "mismatches", st.mismatches_found)


def test_learned_overlapping_solids(rhz, geom, orc):
    """Two overlapping closed cubes: winding number 2 in the overlap; the
    signed crossing count keeps it inside where parity would flip it."""
    v, f = geom.two_cubes()
    tris = orc.build_triangles(v, f)
    h = 0.25
    mn, mx = geom.frame_grid(v, h)
    _compare_learned(rhz, tris, h, mn, mx)


def test_learned_vessel_coarse(rhz, geom):
    v, f = geom.vessel()
    tris = rhz.build_triangles(v, f)
    h = 0.5
    mn, mx = geom.frame_grid(v, h)
    st, est, _ = _compare_learned(rhz, tris, h, mn, mx)
    print(st.log_line(), "reprobed", st.reprobed_voxels, "mismatches", st.mismatches_found)


def _plain_signs(rhz, tris, h, mn, mx):
    rc, ref, _ = rhz.compute_sdf(tris, h, mn, mx, rhz.default_params(method=rhz.METHOD_PLAIN_EXACT))
    assert rc == rhz.RHZ_OK
    return ref.reshape(-1)


@pytest.mark.parametrize("case", ["small_sphere", "ico5_128", "vessel_1mm"])
def test_certificate_covers_every_scanline_error(rhz, geom, torch_gpu, request, case):
    """Every voxel whose scanline sign differs from the exact winding sign is
    a candidate (band distance within the jitter reach) and is flagged by the
    segment certificate; the flag is a small superset of the errors."""
    torch = torch_gpu
    if case == "vessel_1mm":
        v, f = geom.vessel()
        tris = rhz.build_triangles(v, f)
        h = 1.0
        mn, mx = geom.frame_grid(v, h)
    else:
        tris, h, mn, mx = request.getfixturevalue(case)
    V = int(np.prod(dims_of(mn, mx).astype(np.int64)))
    d_t = torch.from_numpy(tris).cuda()
    n = rhz.Proposer.candidates(d_t, len(tris), h, mn, mx, None, None, 0)
    assert n > 0
    d_v = torch.zeros(n, dtype=torch.int64, device="cuda")
    d_f = torch.zeros((n, 12), dtype=torch.float32, device="cuda")
    d_r = torch.zeros(n, dtype=torch.uint8, device="cuda")
    d_s = torch.zeros(V, dtype=torch.uint8, device="cuda")
    n2 = rhz.Proposer.candidates(d_t, len(tris), h, mn, mx, d_v, d_f, n, d_risk=d_r, d_scan=d_s)
    torch.cuda.synchronize()
    assert n2 == n
    vox = d_v.cpu().numpy()
    feats = d_f.cpu().numpy()
    risk = d_r.cpu().numpy().astype(bool)
    scan = d_s.cpu().numpy()
    assert vox.min() >= 0 and vox.max() < V and len(np.unique(vox)) == n
    assert np.all(np.isfinite(feats)) and np.all(feats[:, 0] >= 0)
    assert set(np.unique(feats[:, 1])) <= {0.0, 1.0}
    assert set(np.unique(scan)) <= {0, 1}
    ref = _plain_signs(rhz, tris, h, mn, mx)
    exact = (~np.signbit(ref)).astype(np.uint8) ^ 1  # 1 inside
    on_surface = np.abs(ref) <= 1e-5 * h
    wrong = (scan != exact) & ~on_surface
    is_cand = np.zeros(V, bool)
    is_cand[vox] = True
    flagged = np.zeros(V, bool)
    flagged[vox[risk]] = True
    assert not np.any(wrong & ~is_cand), "scanline error outside the candidate set"
    assert not np.any(wrong & ~flagged), f"{int(np.sum(wrong & ~flagged))} errors not flagged"
    # the flag is selective: at most 2x the actual errors plus a few
    print(case, "candidates", n, "flagged", int(risk.sum()), "errors", int(wrong.sum()))
    assert risk.sum() <= 2 * wrong.sum() + 16


def test_ranked_probes_alone_can_miss(rhz, small_sphere):
    """Without the certificate a blind (hashed subset) probe budget leaves
    scanline errors behind; with it the result is exact again."""
    tris, h, mn, mx = small_sphere
    ref = _plain_signs(rhz, tris, h, mn, mx)
    on_surface = np.abs(ref) <= 1e-5 * h
    res = {}
    for scorer in (1, 2):
        p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED, scorer=scorer, learned_certify=0,
                               reprobe_fraction=5e-4)
        rc, got, st = rhz.compute_sdf(tris, h, mn, mx, p)
        assert rc == rhz.RHZ_OK
        res[scorer] = int(np.sum((np.signbit(got.reshape(-1)) != np.signbit(ref)) & ~on_surface))
        assert st.fallback_probes == 0
    # the planar-margin ranking leaves fewer errors than blind probes
    assert res[2] > 0
    assert res[1] < res[2], res


def test_network_gain_over_blind_probes(rhz, ico5_128):
    """The trained proposer against the method without it.  Config: the
    perturbed icosphere on 128^3 (a shape family the proposer never saw in
    training), top 0.1% of voxels re-probed exactly, certificate off: the
    network leaves far fewer scanline errors behind than blind
    probes of the same budget.  With the certificate on (default), every
    scorer ends at 0 errors and the network spends fewer ranked probes than
    the planar-margin rule."""
    if rhz.Proposer.builtin() is None:
        pytest.skip("library built without trained proposer weights")
    tris, h, mn, mx = ico5_128
    ref = _plain_signs(rhz, tris, h, mn, mx)
    on_surface = np.abs(ref) <= 1e-5 * h
    resid, stats = {}, {}
    for scorer in (0, 1, 2):
        p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED, scorer=scorer, learned_certify=0,
                               reprobe_fraction=1e-3)
        rc, got, st = rhz.compute_sdf(tris, h, mn, mx, p)
        assert rc == rhz.RHZ_OK
        resid[scorer] = int(np.sum((np.signbit(got.reshape(-1)) != np.signbit(ref)) & ~on_surface))
        stats[scorer] = st.ranked_probes
    print("residual errors at a 0.1% budget: network", resid[0], "margin", resid[1], "blind", resid[2],
          "probes", stats)
    assert resid[2] > 0
    assert resid[0] * 4 <= resid[2]
    ranked = {}
    for scorer in (0, 1):
        p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED, scorer=scorer)
        rc, got, st = rhz.compute_sdf(tris, h, mn, mx, p)
        assert rc == rhz.RHZ_OK
        assert not np.any((np.signbit(got.reshape(-1)) != np.signbit(ref)) & ~on_surface)
        ranked[scorer] = (st.ranked_probes, st.fallback_probes, st.reprobed_voxels)
    print("classifier mode (ranked, fallback, total): network", ranked[0], "margin", ranked[1])
    assert ranked[0][2] <= ranked[1][2]


@pytest.mark.parametrize("case", ["ico5_128", "two_cubes", "vessel_coarse"])
def test_far_field_certificate_keeps_every_decision(rhz, geom, orc, request, case):
    """The far-field certificate (clusters beyond beta x radius as bounded
    dipoles, undecided voxels to the brute-force kernel) changes no sign:
    fields are bit-identical for beta = 0 (brute force), 2, 6 and 20."""
    if case == "ico5_128":
        tris, h, mn, mx = request.getfixturevalue("ico5_128")
    elif case == "two_cubes":
        v, f = geom.two_cubes()
        tris = orc.build_triangles(v, f)
        h = 0.25
        mn, mx = geom.frame_grid(v, h)
    else:
        v, f = geom.vessel()
        tris = rhz.build_triangles(v, f)
        h = 0.5
        mn, mx = geom.frame_grid(v, h)
    fields, stats = [], []
    for beta in (0.0, 2.0, 6.0, 20.0):
        rc, out, st = rhz.compute_sdf(tris, h, mn, mx, rhz.default_params(method=rhz.METHOD_FAST_LEARNED,
                                                                           far_field_beta=beta))
        assert rc == rhz.RHZ_OK
        fields.append(out.reshape(-1))
        stats.append(st)
    for fo in fields[1:]:
        assert np.array_equal(fo.view(np.uint32), fields[0].view(np.uint32))
    st6 = stats[2]
    print(case, "clusters", st6.n_clusters, "parents", st6.n_parent_clusters, "decided", st6.far_decided,
          "second pass", st6.far_second_pass, "undecided", st6.far_undecided)
    assert stats[0].far_decided == 0 and st6.far_decided + st6.far_undecided == st6.reprobed_voxels
    assert st6.far_decided >= 0.9 * st6.reprobed_voxels


def test_candidate_features_against_fp64(rhz, orc, torch_gpu, ico5_128):
    """Features 0 (nearest distance / h), 2-4 (nearest normal) and 8
    (log1p of the triangles within h) of a sample of candidates equal a
    binary64 evaluation over EVERY triangle of the mesh: the kernel's column
    lists and its z-extent reject drop no triangle these features need."""
    torch = torch_gpu
    tris, h, mn, mx = ico5_128
    d_tris = torch.from_numpy(tris).to("cuda")
    cap = 1 << 20
    d_vox = torch.empty(cap, dtype=torch.int64, device="cuda")
    d_feat = torch.empty(cap * 12, dtype=torch.float32, device="cuda")
    n = rhz.Proposer.candidates(d_tris, len(tris), h, mn, mx, d_vox, d_feat, cap)
    assert n > 1000
    vox = d_vox[:n].cpu().numpy()
    feat = d_feat[:12 * n].cpu().numpy().reshape(n, 12)
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    L = orc.lib()
    rng = np.random.default_rng(12)
    V = tris.reshape(-1, 4, 3).astype(np.float64)
    cen = V[:, :3].mean(axis=1)
    reach = 2 * h + np.max(np.linalg.norm(V[:, :3] - cen[:, None], axis=2))  # every triangle within 2 h
    checked = 0
    # the list is appended to by many waves: sample it in voxel order, so that
    # the same voxels are checked on every run
    for k in np.argsort(vox)[rng.choice(n, 150, replace=False)]:
        v = int(vox[k])
        x, y, z = v % dims[0], (v // dims[0]) % dims[1], v // (dims[0] * dims[1])
        p = np.array([(mn[0] + x) * h, (mn[1] + y) * h, (mn[2] + z) * h], np.float64)
        cand = np.nonzero(np.linalg.norm(cen - p, axis=1) <= reach)[0]
        d2 = np.full(len(V), np.inf)
        d2[cand] = [L.orc_point_triangle_distance_sq_f64(p, *(np.ascontiguousarray(V[t, j]) for j in range(3)))
                    for t in cand]
        if np.any(np.abs(np.sqrt(d2) - h) < 1e-4 * h):
            continue  # a triangle at the 1 h threshold: its count depends on rounding
        best = int(np.argmin(d2))
        assert feat[k, 0] == pytest.approx(np.sqrt(d2[best]) / h, rel=1e-4, abs=1e-5)
        # the nearest point may be a shared vertex or edge: any triangle tied at
        # the minimum, to within binary32 rounding of the distance (1e-4 h)
        tied = np.nonzero(np.sqrt(d2) <= np.sqrt(d2[best]) * (1 + 1e-5) + 1e-4 * h)[0]
        normals = tris.reshape(-1, 4, 3)[tied, 3]
        assert np.any(np.all(np.abs(normals - feat[k, 2:5]) <= 1e-6, axis=1))
        assert feat[k, 8] == pytest.approx(np.log1p(np.sum(d2 <= h * h)), rel=1e-6)
        checked += 1
    assert checked > 100


def _axis_parity_brute_force(tris, h, mn, vox, dims, axis, jx, jy):
    """Features 5 (axis 0) / 7 (axis 1) of the voxels `vox` over EVERY
    triangle, in float32 with the kernel's operation order: signed crossings
    of the jittered ray along the axis, counted below the voxel centre."""
    f32 = np.float32
    T = tris.reshape(-1, 4, 3)[:, :3].astype(np.float32)
    perm = [1, 2, 0] if axis == 0 else [2, 0, 1]
    R = T[:, :, perm]  # (u, v, w) of every vertex
    ax, ay, az = R[:, 0, 0], R[:, 0, 1], R[:, 0, 2]
    bx, by, bz = R[:, 1, 0], R[:, 1, 1], R[:, 1, 2]
    cx, cy, cz = R[:, 2, 0], R[:, 2, 1], R[:, 2, 2]
    mnr = [mn[k] for k in perm]
    out = []
    for v in vox:
        x, y, z = v % dims[0], (v // dims[0]) % dims[1], v // (dims[0] * dims[1])
        idx = [(x, y, z)[k] for k in perm]
        ru = (f32(mnr[0] + idx[0]) + f32(jx)) * f32(h)
        rv = (f32(mnr[1] + idx[1]) + f32(jy)) * f32(h)
        pw = f32(mnr[2] + idx[2]) * f32(h)
        d0 = (bx - ax) * (rv - ay) - (by - ay) * (ru - ax)
        d1 = (cx - bx) * (rv - by) - (cy - by) * (ru - bx)
        d2 = (ax - cx) * (rv - cy) - (ay - cy) * (ru - cx)
        pos = (d0 > 0) & (d1 > 0) & (d2 > 0)
        neg = (d0 < 0) & (d1 < 0) & (d2 < 0)
        area2 = d0 + d1 + d2
        hit = (pos | neg) & (np.abs(area2) >= f32(1e-12))
        with np.errstate(divide="ignore", invalid="ignore"):
            wh = (d1 * az + d2 * bz + d0 * cz) / area2
        sgn = np.where(pos, 1, -1)
        total = int(np.sum(sgn[hit]))
        below = int(np.sum(sgn[hit & (wh < pw)]))
        out.append(0.5 if total != 0 else (1.0 if below != 0 else 0.0))
    return np.array(out, np.float32)


@pytest.mark.parametrize("case", ["ico5_128", "vessel_1mm"])
def test_axis_parity_features_against_brute_force(rhz, geom, torch_gpu, request, case):
    """Features 5 and 7 (inside by the signed crossings of +x and +y rays,
    the other two of the three axis parity counts) equal a float32
    evaluation over EVERY triangle with the kernel's arithmetic: the rotated
    column lists drop no crossing.  On a closed mesh every line closes (no
    0.5), and away from the jitter reach the x and y answers agree with the
    z scan sign."""
    torch = torch_gpu
    if case == "ico5_128":
        tris, h, mn, mx = request.getfixturevalue("ico5_128")
    else:
        v, f = geom.vessel()
        tris = rhz.build_triangles(v, f)
        h = 1.0
        mn, mx = geom.frame_grid(v, h)
    d_tris = torch.from_numpy(tris).to("cuda")
    cap = 1 << 21
    d_vox = torch.empty(cap, dtype=torch.int64, device="cuda")
    d_feat = torch.empty(cap * 12, dtype=torch.float32, device="cuda")
    n = rhz.Proposer.candidates(d_tris, len(tris), h, mn, mx, d_vox, d_feat, cap)
    assert 1000 < n <= cap
    vox = d_vox[:n].cpu().numpy()
    feat = d_feat[:12 * n].cpu().numpy().reshape(n, 12)
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    p = rhz.default_params()
    for slot in (5, 7):
        assert set(np.unique(feat[:, slot])) <= {0.0, 1.0}, slot
    rng = np.random.default_rng(21)
    pick = rng.choice(n, 300, replace=False)
    for axis, slot in ((0, 5), (1, 7)):
        ref = _axis_parity_brute_force(tris, h, mn, vox[pick], dims, axis, p.jitter_x, p.jitter_y)
        assert np.array_equal(feat[pick, slot], ref), (axis, int(np.sum(feat[pick, slot] != ref)))
    # candidates whose segment certificate rules out a scanline error have the
    # exact sign in feature 1; the x and y rays mostly agree with it
    clear = feat[:, 10] == 0
    agree = np.mean((feat[clear, 5] == feat[clear, 1]) & (feat[clear, 7] == feat[clear, 1]))
    print(case, "candidates", n, "x/y agree with a certified scan sign:", round(float(agree), 4))
    assert agree > 0.5


def test_sign_stage_overlap_matches_sequential(rhz, geom, torch_gpu):
    """On a lattice of >= 16M voxels the sign stage (features, ranking, exact
    probes, commits) runs on a side stream while the eikonal rounds run; the
    field and every count equal the sequential order bit for bit."""
    torch = torch_gpu
    v, f = geom.vessel()
    tris = rhz.build_triangles(v, f)
    h = 0.25
    mn, mx = geom.frame_grid(v, h)
    V = int(np.prod(dims_of(mn, mx).astype(np.int64)))
    assert V >= 1 << 24
    d_t = torch.from_numpy(tris).cuda()
    saved = rhz.tuning_get("verify_overlap")
    outs, stats = [], []
    try:
        for ov in (0, 1, 1, 0):
            rhz.tuning_set("verify_overlap", ov)
            out = torch.empty(V, dtype=torch.float32, device="cuda")
            rc, st = rhz.compute_sdf_device(d_t, len(tris), h, mn, mx, out,
                                            rhz.default_params(method=rhz.METHOD_FAST_LEARNED))
            torch.cuda.synchronize()
            assert rc == rhz.RHZ_OK
            outs.append(out)
            stats.append(st)
    finally:
        rhz.tuning_set("verify_overlap", saved)
    keys = ("candidates", "risky_candidates", "ranked_probes", "ranked_mismatches", "fallback_probes",
            "fallback_mismatches", "mismatches_found", "reprobed_voxels", "eikonal_rounds", "sweeps",
            "far_residual", "fixup_voxels")
    for o, st in zip(outs[1:], stats[1:]):
        assert torch.equal(o.view(torch.int32), outs[0].view(torch.int32))
        for k in keys:
            assert getattr(st, k) == getattr(stats[0], k), k
    print(stats[1].log_line(), "| overlapped: eikonal", round(stats[1].ms_eikonal, 2), "ms, sign stage",
          round(stats[1].ms_proposer + stats[1].ms_verify, 2), "ms, total", round(stats[1].ms_total, 2),
          "| sequential total", round(stats[0].ms_total, 2))


def test_sign_stage_overlap_decline(rhz, geom, torch_gpu):
    """A decline raised on the side stream (candidate lists over their cap)
    comes back as RHZ_DECLINED with its reason, and AUTO falls back."""
    torch = torch_gpu
    v, f = geom.vessel()
    tris = rhz.build_triangles(v, f)
    h = 0.25
    mn, mx = geom.frame_grid(v, h)
    V = int(np.prod(dims_of(mn, mx).astype(np.int64)))
    d_t = torch.from_numpy(tris).cuda()
    out = torch.empty(V, dtype=torch.float32, device="cuda")
    saved = rhz.tuning_get("verify_overlap")
    try:
        rhz.tuning_set("verify_overlap", 1)
        rc, st = rhz.compute_sdf_device(d_t, len(tris), h, mn, mx, out,
                                        rhz.default_params(method=rhz.METHOD_FAST_LEARNED, csr_max_entries=10))
        assert rc == rhz.RHZ_DECLINED and st.decline_reason == 2
    finally:
        rhz.tuning_set("verify_overlap", saved)


def test_sign_stage_overlap_with_safe_clamp(rhz, geom, torch_gpu):
    """No rounds and no extension rounds leave far voxels at the sentinel, so
    the safe clamp runs after the join (it reads the signs the side stream
    wrote): the overlapped build equals the sequential one bit for bit."""
    torch = torch_gpu
    v, f = geom.vessel()
    tris = rhz.build_triangles(v, f)
    h = 0.25
    mn, mx = geom.frame_grid(v, h)
    V = int(np.prod(dims_of(mn, mx).astype(np.int64)))
    d_t = torch.from_numpy(tris).cuda()
    saved = rhz.tuning_get("verify_overlap")
    outs, stats = [], []
    try:
        for ov in (0, 1):
            rhz.tuning_set("verify_overlap", ov)
            out = torch.empty(V, dtype=torch.float32, device="cuda")
            p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED, max_rounds=0, max_extra_rounds=0)
            rc, st = rhz.compute_sdf_device(d_t, len(tris), h, mn, mx, out, p)
            torch.cuda.synchronize()
            assert rc == rhz.RHZ_OK
            outs.append(out)
            stats.append(st)
    finally:
        rhz.tuning_set("verify_overlap", saved)
    assert stats[1].safe_clamped > 0 and stats[1].safe_clamped == stats[0].safe_clamped
    assert stats[1].far_residual == stats[0].far_residual
    assert torch.equal(outs[1].view(torch.int32), outs[0].view(torch.int32))
