"""The C++ drop-in (include/rhizome/cuda_rasterizer.hpp) driven like the
reference's callers: build_from_mesh and compute_winding_numbers_gpu through a
host program (tests/cxx/dropin_main, built by the package build) must give the
same bits as the same calls made through the Python binding."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cxx", "dropin_main")


def _write_mesh(path, v, f):
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(f, np.uint32).reshape(-1, 3)
    with open(path, "wb") as fh:
        fh.write(np.array([len(v), len(f)], np.int64).tobytes())
        fh.write(v.tobytes())
        fh.write(f.tobytes())


def _run(*args):
    assert os.path.exists(BIN), "tests/cxx/dropin_main missing: run __graft_entry__.build()"
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return r.stdout


def test_dropin_reports_device():
    out = _run("info").split()
    assert out[0] == "available" and out[1] == "1"
    assert out[2] == "memory" and out[3] == "1" and int(out[5]) > 100 * 2 ** 30


def test_dropin_build_from_mesh_matches_binding(rhz, geom, tmp_path):
    v, f = geom.perturbed_icosphere(seed=1, subdiv=4)
    mesh = tmp_path / "mesh.bin"
    out = tmp_path / "out.bin"
    _write_mesh(mesh, v, f)
    print(_run("sdf", str(mesh), "1.0", str(out)))
    raw = np.fromfile(out, dtype=np.uint8)
    mn = tuple(raw[:12].view(np.int32))
    mx = tuple(raw[12:24].view(np.int32))
    field = raw[24:].view(np.float32)
    rmn, rmx, ref, st = rhz.build_from_mesh(v, f, 1.0)
    assert mn == tuple(rmn) and mx == tuple(rmx)
    assert np.array_equal(field.view(np.uint32), ref.reshape(-1).view(np.uint32))
    # negative inside: the centre of the sphere
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    c = [-mn[i] for i in range(3)]
    assert field[(c[2] * dims[1] + c[1]) * dims[0] + c[0]] < 0


def test_dropin_winding_matches_binding(rhz, geom, tmp_path):
    v, f = geom.perturbed_icosphere(seed=1, subdiv=3)
    mesh = tmp_path / "mesh.bin"
    pts = tmp_path / "pts.bin"
    out = tmp_path / "w.bin"
    _write_mesh(mesh, v, f)
    p = np.random.default_rng(5).uniform(-30, 30, size=(4096, 3)).astype(np.float32)
    with open(pts, "wb") as fh:
        fh.write(np.array([len(p)], np.int64).tobytes())
        fh.write(p.tobytes())
    _run("wind", str(mesh), str(pts), str(out))
    w = np.fromfile(out, dtype=np.float32)
    ref = rhz.winding_numbers(v, f, p)
    assert np.array_equal(w.view(np.uint32), np.asarray(ref, np.float32).view(np.uint32))
    r = np.linalg.norm(p, axis=1)
    assert np.all(np.abs(w[r < 15]) > 0.99) and np.all(np.abs(w[r > 25]) < 0.01)


def test_dropin_query_distances_matches_binding(rhz, orc, geom, tmp_path):
    """query_distance_batch_gpu through the C++ drop-in equals the binding and
    the oracle bit for bit."""
    v, f = geom.perturbed_icosphere(seed=2, subdiv=3)
    mesh = tmp_path / "mesh.bin"
    pts = tmp_path / "pts.bin"
    out = tmp_path / "d.bin"
    _write_mesh(mesh, v, f)
    p = np.random.default_rng(6).uniform(-30, 30, size=(3000, 3)).astype(np.float32)
    with open(pts, "wb") as fh:
        fh.write(np.array([len(p)], np.int64).tobytes())
        fh.write(p.tobytes())
    _run("dist", str(mesh), str(pts), str(out))
    d = np.fromfile(out, dtype=np.float32)
    ref = rhz.query_distances(v, f, p)
    assert np.array_equal(d.view(np.uint32), np.asarray(ref, np.float32).view(np.uint32))
    assert np.array_equal(d.view(np.uint32), orc.query_distances(v, f, p).view(np.uint32))


def test_query_distances_empty_inputs(rhz, geom):
    """No query points: nothing to do; no triangles: every distance is 0
    (cuda_mesh_ops.cu:714-718)."""
    v, f = geom.perturbed_icosphere(seed=2, subdiv=1)
    assert rhz.query_distances(v, f, np.zeros((0, 3), np.float32)).shape == (0,)
    d = rhz.query_distances(v, f[:0], np.ones((5, 3), np.float32))
    assert np.array_equal(d, np.zeros(5, np.float32))


def test_dropin_fast_path_logs_reference_line(rhz, geom, tmp_path):
    """Above the 32Mi-voxel gate compute_mesh_sdf takes the fast path and
    logs the reference's summary line on channel CUDA (stderr by default)."""
    v, f = geom.perturbed_icosphere(seed=1, subdiv=4)
    mesh = tmp_path / "mesh.bin"
    out = tmp_path / "out.bin"
    _write_mesh(mesh, v, f)
    r = subprocess.run([BIN, "sdf", str(mesh), "0.125", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    raw = np.fromfile(out, dtype=np.uint8)
    mn, mx = tuple(raw[:12].view(np.int32)), tuple(raw[12:24].view(np.int32))
    dims = [int(mx[i] - mn[i] + 1) for i in range(3)]
    assert np.prod(dims) >= 32 * 2 ** 20
    line = [l for l in r.stderr.splitlines() if l.startswith("[CUDA] Fast SDF path OK:")]
    assert len(line) == 1, r.stderr
    assert f"| {dims[0]}x{dims[1]}x{dims[2]} voxels, {len(f)} tris, band 8 vox," in line[0]
    print(line[0])


def test_dropin_winding_prepass_matches_binding(rhz, geom, tmp_path):
    """MeshWindingPass::orient_mesh (the OBJ loader's pre-pass) through the
    C++ drop-in: the binding's indices and flags, the reference's message."""
    v, f = geom.icosphere(3)
    v = (v * 20.0).astype(np.float32)
    f = np.array(f, dtype=np.uint32)
    f[5:40] = f[5:40][:, [0, 2, 1]]
    mesh = tmp_path / "mesh.bin"
    out = tmp_path / "out.bin"
    _write_mesh(mesh, v, f)
    lines = _run("orient", str(mesh), str(out)).splitlines()
    ref = rhz.orient_mesh(v, f)
    assert np.array_equal(np.fromfile(out, dtype=np.uint32).reshape(-1, 3), ref[0])
    assert lines[0].startswith("orientable 1 inconsistent 1 inverted 0 volume")
    assert lines[1] == "Inconsistent face winding detected and unified."
