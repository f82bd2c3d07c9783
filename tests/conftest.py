import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _ensure_built():
    """Rebuild a library that is missing or older than its sources (a
    fingerprint of the sources and flags is stamped beside each build), so a
    test never runs a stale binary."""
    from oracle import build as obuild
    from gudni_amd import build as pbuild
    if obuild.stale():
        obuild.build()
    if pbuild.stale():
        pbuild.build()


_ensure_built()


@pytest.fixture(scope="session")
def rasterizer():
    import gudni_amd as G
    r = G.Rasterizer()
    yield r
    r.close()
