import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _ensure_built():
    """Load what is built; build only when nothing is."""
    from oracle import build as obuild
    from gudni_amd import build as pbuild
    if not os.path.exists(obuild.LIB):
        obuild.build()
    if not os.path.exists(pbuild.LIB):
        pbuild.build()


_ensure_built()


@pytest.fixture(scope="session")
def rasterizer():
    import gudni_amd as G
    r = G.Rasterizer()
    yield r
    r.close()
