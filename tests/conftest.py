import os
import sys

import pytest

try:
    # the differentiable harness hands torch tensors to librtx: torch must be
    # loaded first so that both bind the same HIP runtime
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the device)")
    config.addinivalue_line("markers", "slow: takes more than ~10 s on the CPU")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def rtx():
    import rtxplay_amd
    rtxplay_amd.lib()
    return rtxplay_amd
