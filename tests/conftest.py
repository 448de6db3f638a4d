import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "emt6-ro_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


def golden(*parts):
    return os.path.join(GOLDEN, *parts)


@pytest.fixture(scope="session")
def default_params_path():
    return golden("default-parameters.json")


def gpu_available():
    """Asks the library's own HIP runtime.  (PyTorch is not used here: its
    bundled HIP runtime and the library's must not both initialise in one
    process.)"""
    try:
        from emt6ro._native import lib
        return lib.emt6ro_device_count() > 0
    except Exception:  # noqa: BLE001
        return False
