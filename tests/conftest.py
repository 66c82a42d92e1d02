"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the
CPU (the oracle, the geometry generators, the host logic, the C-ABI exports)."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

import __graft_entry__ as entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def rhz():
    return entry.load_package()


@pytest.fixture(scope="session")
def geom():
    return entry.load_geometry()


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def torch_gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


@pytest.fixture(scope="session")
def small_sphere(geom, orc):
    """icosphere ico^4 (5,120 tris), R 20 +- 2 mm, on a 48^3 grid at 1 mm."""
    v, f = geom.perturbed_icosphere(seed=1, subdiv=4)
    tris = orc.build_triangles(v, f)
    return tris, 1.0, (-24, -24, -24), (23, 23, 23)


@pytest.fixture(scope="session")
def ico5_128(geom, orc):
    """The plain-method small workload: ico^5 (20,480 tris) on 128^3 at 0.5 mm."""
    v, f = geom.perturbed_icosphere(seed=1)
    tris = orc.build_triangles(v, f)
    return tris, 0.5, (-64, -64, -64), (63, 63, 63)


def dims_of(mn, mx):
    return np.array([mx[i] - mn[i] + 1 for i in range(3)], dtype=np.int32)
