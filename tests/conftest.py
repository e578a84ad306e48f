import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: large CPU work")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_py as O
    O.build()
    return O


@pytest.fixture(scope="session")
def gpu_ctx():
    """A library context on device 0.  No skip: on the GPU box a missing
    library or device is a failure."""
    import voxelizer_amd as va
    ctx = va.Context(0)
    yield ctx
    ctx.close()
