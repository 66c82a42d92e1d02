"""Host entry points called from several host threads at once.

Every host-buffer entry point stages its inputs in the device's shared
workspace; the library lock covers staging, compute and the copy back, so a
call made while another thread's call is in flight still returns its own
result (ctypes releases the GIL for the duration of each call)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_compute_sdf_and_interior_count_in_parallel(rhz, geom, torch_gpu):
    v, f = geom.perturbed_icosphere(seed=5, subdiv=4)
    tris = rhz.build_triangles(v, f)
    h, mn, mx = 0.75, (-36, -36, -36), (35, 35, 35)
    p = rhz.default_params(method=rhz.METHOD_FAST_LEARNED)
    rc, ref_field, _ = rhz.compute_sdf(tris, h, mn, mx, p)
    assert rc == rhz.RHZ_OK
    rng = np.random.default_rng(11)
    # a field of another size: the staged output slot must grow under the
    # other thread if the lock did not cover its whole call
    grid = rng.standard_normal(3_000_017).astype(np.float32)
    ref_count = rhz.interior_count(grid)
    assert ref_count == int((grid <= 0).sum())

    errors, fields, counts = [], [], []

    def builder():
        try:
            for _ in range(4):
                rc, fld, _ = rhz.compute_sdf(tris, h, mn, mx, p)
                assert rc == rhz.RHZ_OK
                fields.append(fld)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    def counter():
        try:
            for k in range(12):
                g = grid[: len(grid) - 977 * k]
                counts.append((k, rhz.interior_count(g)))
        except Exception as e:
            errors.append(e)

    th = [threading.Thread(target=builder), threading.Thread(target=counter)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    assert len(fields) == 4 and len(counts) == 12
    for fld in fields:
        assert np.array_equal(fld.view(np.int32), ref_field.view(np.int32))
    for k, c in counts:
        assert c == int((grid[: len(grid) - 977 * k] <= 0).sum())
