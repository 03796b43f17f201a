import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device and the built liblubatch.so")


@pytest.fixture(scope="session")
def gpu():
    """The device layer; fails loudly when the native library is missing."""
    import matrixinversion_amd as mia

    n = mia.device_count()
    if n == 0:
        pytest.skip("no HIP device visible")
    info = mia.device_info(0)
    print(f"[gpu] {info.name} {info.arch} CUs={info.cu_count}", flush=True)
    return mia
