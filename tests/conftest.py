import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")


@pytest.fixture(scope="session")
def ctx():
    """One device context shared by every GPU test (one process, one GPU)."""
    import viprs_amd as v

    c = v.Context(0)
    yield c
    c.synchronize()
    c.close()
