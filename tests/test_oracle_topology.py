"""The oracle's mesh orientation (orc_orient_mesh, orc_signed_volume) and
red-black eikonal (orc_eikonal_redblack) pinned by closed forms.

Orientation: faces flipped on purpose come back exactly; an inside-out
sphere (the reference's Test 5 fixture, harness_vessel_tests.cpp:79-82) is
inverted; a Moebius band is non-orientable; a nested inside-out shell keeps
its winding (only the total volume decides a global inversion); volumes of a
unit cube and a tetrahedron are exact.  A link direction taken the wrong way
round, or flips applied relative to the wrong triangle, fails these.

Red-black: from a planar seed every interior cell within reach of the seed
ends at its exact distance |x - c| h with the seed side's sign, and the reach
after k iterations is 2k - ((c + 1 + y + z) & 1) because the colour
(x + y + z) & 1 = 0 runs first (the other order shifts every reach by one);
face cells are never written.
"""
import numpy as np
import pytest

CUBE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                  np.float32)
CUBE_F = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [1, 2, 6], [1, 6, 5],
                   [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.uint32)


def test_signed_volume_closed_forms(orc):
    assert orc.signed_volume(CUBE_V, CUBE_F) == pytest.approx(1.0, abs=1e-12)
    tet_v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    tet_f = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.uint32)
    assert orc.signed_volume(tet_v, tet_f) == pytest.approx(1.0 / 6.0, abs=1e-12)
    assert orc.signed_volume(CUBE_V, CUBE_F[:, ::-1]) == pytest.approx(-1.0, abs=1e-12)


def _sphere(geom):
    v, f = geom.icosphere(4)
    return (v * 20.0).astype(np.float32), f


def _flip(f, rows):
    g = f.copy()
    g[rows] = g[rows][:, [0, 2, 1]]
    return g


def test_seeded_flips_come_back(orc, geom):
    v, f = _sphere(geom)
    rng = np.random.default_rng(4)
    rows = rng.choice(np.arange(1, len(f)), size=len(f) // 10, replace=False)
    out, orientable, inconsistent, inverted, vol = orc.orient_mesh(v, _flip(f, rows))
    assert orientable and inconsistent and not inverted
    assert np.array_equal(out, f)
    assert vol == pytest.approx(orc.signed_volume(v, f), rel=1e-12)
    assert vol == pytest.approx(4.0 / 3.0 * np.pi * 20.0 ** 3, rel=0.01)


def test_component_follows_its_first_triangle_then_inverts(orc, geom):
    """Triangle 0 flipped: the component is unified to triangle 0 (inward),
    its volume is negative, so every triangle is reversed: the original."""
    v, f = _sphere(geom)
    rng = np.random.default_rng(5)
    rows = np.concatenate([[0], rng.choice(np.arange(1, len(f)), size=200, replace=False)])
    out, orientable, inconsistent, inverted, vol = orc.orient_mesh(v, _flip(f, rows))
    assert orientable and inconsistent and inverted and vol > 0
    assert np.array_equal(out, f)


def test_inside_out_sphere_fixture_is_inverted(orc, geom):
    v, f = geom.uv_sphere(15.0, 48, 48)  # the reference's Test 5 inverted-normals mesh, flipped
    out, orientable, inconsistent, inverted, vol = orc.orient_mesh(v, f[:, ::-1].copy())
    assert orientable and not inconsistent and inverted and vol > 0
    assert np.array_equal(out, _flip(f[:, ::-1].copy(), np.arange(len(f))))


def test_nested_inside_out_shell_keeps_its_winding(orc, geom):
    vo, fo = _sphere(geom)
    vi = (vo * 0.5).astype(np.float32)
    fi = fo[:, ::-1] + len(vo)
    v = np.concatenate([vo, vi])
    f = np.concatenate([fo, fi]).astype(np.uint32)
    out, orientable, inconsistent, inverted, vol = orc.orient_mesh(v, f)
    assert orientable and not inconsistent and not inverted
    assert np.array_equal(out, f)
    assert vol == pytest.approx(orc.signed_volume(vo, fo) * (1 - 0.125), rel=1e-9)


def test_moebius_band_is_non_orientable(orc, geom):
    v, f = geom.moebius()
    _, orientable, inconsistent, _, _ = orc.orient_mesh(v, f)
    assert not orientable and inconsistent
    # the same strip cut open (last quad removed) is orientable and consistent
    _, orientable, inconsistent, _, _ = orc.orient_mesh(v, f[:-2])
    assert orientable and not inconsistent


def test_unlinked_triangles_are_left_alone(orc):
    """Non-manifold edges (three triangles) and boundary edges link nothing."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1]], np.float32)
    f = np.array([[0, 1, 2], [0, 1, 3], [0, 1, 4]], np.uint32)  # three faces on edge (0, 1)
    out, orientable, inconsistent, inverted, vol = orc.orient_mesh(v, f)
    assert orientable and not inconsistent
    assert np.array_equal(out, f) or np.array_equal(out, f[:, [0, 2, 1]])


@pytest.mark.parametrize("c,iters", [(10, 3), (11, 4)])
def test_redblack_reach_from_planar_seed(orc, c, iters):
    dims = np.array([32, 9, 7], np.int32)
    h = np.float32(0.5)
    big = np.float32(1e30)
    g = np.full((dims[2], dims[1], dims[0]), big, np.float32)
    g[:, :, :c] = -big
    g[:, :, c] = 0.0
    ref = g.copy()
    flat = g.reshape(-1).copy()
    orc.lib().orc_eikonal_redblack(flat, dims, h, iters)
    out = flat.reshape(g.shape)
    for z in range(dims[2]):
        for y in range(dims[1]):
            for x in range(dims[0]):
                face = x in (0, dims[0] - 1) or y in (0, dims[1] - 1) or z in (0, dims[2] - 1)
                if face:
                    assert out[z, y, x] == ref[z, y, x]
                    continue
                d = abs(x - c)
                reach = 2 * iters - ((c + 1 + y + z) & 1)
                if d <= reach:
                    assert out[z, y, x] == np.float32(d) * h * (-1 if x < c else 1), (x, y, z)
                else:
                    assert abs(out[z, y, x]) == big and np.signbit(out[z, y, x]) == (x < c)
