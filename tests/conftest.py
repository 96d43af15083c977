import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libzkg.so")


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the whole session (GPU tests only)."""
    import ziros_amd

    c = ziros_amd.Context(0)
    for field, width in ((ziros_amd.GOLDILOCKS, 16), (ziros_amd.GOLDILOCKS, 8), (ziros_amd.BABYBEAR, 16)):
        c.load_constants(field, width, 42)
    yield c
    c.close()
