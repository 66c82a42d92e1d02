"""Every sweep on the y-fastest copy (tuning "fsm_layout" 1: the y-sweep
walking along the rows in register chunks, z-sweeps with lanes over y, no
transposes between rounds) gives the fields of the other layout bit for bit,
for every y-sweep kernel (sweep_run_kernel with 16- and 32-cell chunks, the
prefetching kernel),
on lattices whose y extent is and is not a multiple of the chunk and whose x
extent is and is not a multiple of 64, in the plain and learned methods (the
learned one on a lattice large enough for the overlapped sign stage), and for
rounds to convergence."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture()
def layout(rhz):
    saved = (rhz.tuning_get("fsm_layout"), rhz.tuning_get("sweep_run_chunk"), rhz.tuning_get("sweep_run_block"))
    yield
    rhz.tuning_set("fsm_layout", saved[0])
    rhz.tuning_set("sweep_run_chunk", saved[1])
    rhz.tuning_set("sweep_run_block", saved[2])


# (41: the prefetching kernel, 0: sweep_run_kernel; cells per chunk of sweep_run_kernel)
RUN_VARIANTS = [(0, 32), (0, 16), (41, 32), (42, 32)]


# lattices cut from the small sphere's frame: x past one 64-row block, y
# shorter than one chunk or ragged, z shorter than a block of planes
CLIPS = {"clip_70x5x9": ((-35, -2, -4), (34, 2, 4)), "clip_65x37x3": ((-30, -18, -1), (34, 18, 1)),
         "clip_2x2x2": ((0, 0, 0), (1, 1, 1))}


def _case(name, request, geom, rhz):
    if name in ("small_sphere", "ico5_128"):
        return request.getfixturevalue(name)
    if name in CLIPS:
        tris, h, _, _ = request.getfixturevalue("small_sphere")
        return (tris, h) + CLIPS[name]
    h = {"vessel_1mm": 1.0, "vessel_0p25mm": 0.25}[name]
    v, f = geom.vessel()
    tris = rhz.build_triangles(v, f)
    mn, mx = geom.frame_grid(v, h)
    return tris, h, mn, mx


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name,method,policy", [
    ("small_sphere", "plain", "reference"), ("ico5_128", "learned", "reference"),
    ("clip_70x5x9", "plain", "converge"), ("clip_65x37x3", "plain", "converge"), ("clip_2x2x2", "plain", "reference"),
    ("vessel_1mm", "learned", "reference"), ("vessel_1mm", "plain", "converge"),
    ("vessel_0p25mm", "learned", "reference")])
def test_y_fastest_rounds_match_default(rhz, geom, torch_gpu, request, layout, name, method, policy):
    tris, h, mn, mx = _case(name, request, geom, rhz)
    m = rhz.METHOD_PLAIN_EXACT if method == "plain" else rhz.METHOD_FAST_LEARNED
    fsm = rhz.FSM_REFERENCE if policy == "reference" else rhz.FSM_CONVERGE
    p = rhz.default_params(method=m, fsm_policy=fsm)
    rhz.tuning_set("fsm_layout", 0)
    rc, ref, st0 = rhz.compute_sdf(tris, h, mn, mx, p)
    assert rc == rhz.RHZ_OK
    for block, chunk in RUN_VARIANTS:
        rhz.tuning_set("fsm_layout", 1)
        rhz.tuning_set("sweep_run_block", block)
        rhz.tuning_set("sweep_run_chunk", chunk)
        rc, got, st1 = rhz.compute_sdf(tris, h, mn, mx, p)
        assert rc == rhz.RHZ_OK
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (block, chunk)
        assert st1.eikonal_rounds == st0.eikonal_rounds and st1.sweeps == st0.sweeps
        print(name, method, policy, "block", block, "chunk", chunk, "eikonal ms", round(st0.ms_eikonal, 2), "->",
              round(st1.ms_eikonal, 2), "transposes ms", round(st0.ms_transposes, 2), "->",
              round(st1.ms_transposes, 2), flush=True)
