import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def lace():
    import lace_amd
    lace_amd.lib()  # raises if the library is not built
    return lace_amd


@pytest.fixture(scope="session")
def gpu(lace):
    if not lace.gpu_available():
        pytest.fail("GPU test selected but no gfx950 device could be initialised: "
                    + lace.lib().lace_last_error().decode())
    return lace


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
