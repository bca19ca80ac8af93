import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X through the C ABI")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the GPU tests."""
    from isosurface_amd import Context

    c = Context(0)
    yield c
    c.close()
