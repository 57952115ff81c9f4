import os
import sys

import pytest

# One HIP runtime per process: torch first, so that the native library binds
# to torch's runtime and the torch-based tests share the device with it.
try:
    import torch  # noqa: F401
except ImportError:  # torch is optional for everything but the PPR op
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build_library()
    return oracle
