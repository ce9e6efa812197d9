import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on the GPU box)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded; a GPU test fails (never skips) if it cannot run."""
    from openem_amd import _lib
    lib = _lib.load()
    assert _lib.device_count() >= 1, "no HIP device visible"
    return lib
