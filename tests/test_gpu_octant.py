"""The classical 8-ordering fast sweeping on the GPU (rhz_stage_sweep_octant,
RHZ_FSM_OCTANT) against the oracle's lexicographic sweep (orc_fsm_sweep_octant),
bit for bit: every ordering on odd and degenerate lattices, rounds to
convergence on the plain small configuration's lattice (icosphere, 128^3 at
0.5 mm), and the pipeline policy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ORDERINGS = [(sx, sy, sz) for sz in (1, -1) for sy in (1, -1) for sx in (1, -1)]


def _seeded_field(dims, seed, n_seeds=7):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    d = np.full(n, 1e30, np.float32)
    pick = rng.choice(n, min(n_seeds, n), replace=False)
    d[pick] = rng.uniform(0.0, 2.0, len(pick)).astype(np.float32)
    return d


@pytest.mark.parametrize("dims", [(37, 29, 23), (1, 5, 7), (64, 1, 3), (2, 2, 2)])
def test_each_ordering_matches_oracle(rhz, orc, torch_gpu, dims):
    torch = torch_gpu
    dims = np.array(dims, np.int32)
    h = 0.37
    d = _seeded_field(dims, int(dims.sum()))
    g = torch.from_numpy(d.copy()).cuda()
    flag = torch.zeros(1, dtype=torch.int32, device="cuda")
    for o in ORDERINGS:
        flag.zero_()
        rhz.Stages.sweep_octant(g, dims, h, *o, flag)
        ch = orc.lib().orc_fsm_sweep_octant(d, orc._i32(dims), h, *o)
        torch.cuda.synchronize()
        assert np.array_equal(g.cpu().numpy().view(np.uint32), d.view(np.uint32)), o
        assert int(flag.item()) == int(ch), o


def test_rounds_to_convergence_on_the_small_lattice(rhz, orc, torch_gpu, ico5_128):
    torch = torch_gpu
    tris, h, mn, mx = ico5_128
    dims = orc.dims_of(mn, mx)
    band = orc.splat_band(tris, h, mn, dims)
    ref = band.copy()
    rounds = orc.octant_converge(ref, dims, h)
    g = torch.from_numpy(band.copy()).cuda()
    flag = torch.zeros(1, dtype=torch.int32, device="cuda")
    got_rounds = 0
    for r in range(64):
        flag.zero_()
        for o in ORDERINGS:
            rhz.Stages.sweep_octant(g, dims, h, *o, flag)
        torch.cuda.synchronize()
        got_rounds = r + 1
        if int(flag.item()) == 0:
            break
    print("octant rounds to convergence:", rounds)
    assert got_rounds == rounds
    assert np.array_equal(g.cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("method", ["plain", "learned"])
def test_octant_policy_in_the_pipeline(rhz, orc, torch_gpu, ico5_128, method):
    """fsm_policy = RHZ_FSM_OCTANT: magnitudes equal the oracle's octant
    convergence on the band, signs equal the default policy's, rounds and
    sweep count agree."""
    tris, h, mn, mx = ico5_128
    dims = orc.dims_of(mn, mx)
    m = rhz.METHOD_PLAIN_EXACT if method == "plain" else rhz.METHOD_FAST_LEARNED
    rc, out, st = rhz.compute_sdf(tris, h, mn, mx, rhz.default_params(method=m, fsm_policy=rhz.FSM_OCTANT))
    assert rc == rhz.RHZ_OK
    rc, base, _ = rhz.compute_sdf(tris, h, mn, mx, rhz.default_params(method=m))
    assert rc == rhz.RHZ_OK
    ref = orc.splat_band(tris, h, mn, dims)
    rounds = orc.octant_converge(ref, dims, h)
    out = out.reshape(-1)
    assert np.array_equal(np.abs(out).view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(np.signbit(out), np.signbit(base.reshape(-1)))
    assert st.eikonal_rounds == rounds and st.sweeps == 8 * rounds and st.far_residual == 0
    print(method, "octant rounds", rounds, "eikonal ms", round(st.ms_eikonal, 2), "total ms", round(st.ms_total, 2))


def test_octant_policy_rejects_bad_values(rhz):
    with pytest.raises(rhz.RhzError):
        rhz.compute_sdf(np.zeros((1, 12), np.float32), 1.0, (0, 0, 0), (3, 3, 3), rhz.default_params(fsm_policy=3))
