import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")
    config.addinivalue_line("markers", "slow: longer CPU runs")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with at least one visible device; fails loudly when the
    library is missing (there is no CPU fallback)."""
    from nonnewlbm_amd import _abi, lbm355
    _abi.load()
    n = lbm355.device_count()
    if n < 1:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return _abi.load()
