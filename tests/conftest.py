import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def vortex_meshes(tmp_path_factory):
    """The isentropic-vortex meshes: 20x20 (800 triangles) and 40x40 quads."""
    from gocfd_amd import meshgen
    d = tmp_path_factory.mktemp("meshes")
    return {n: meshgen.vortex_mesh(str(d / f"vortex{n}.su2"), n) for n in (4, 20, 40)}
