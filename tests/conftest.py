import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP path through the C-ABI")
    config.addinivalue_line("markers", "slow: larger instances")


@pytest.fixture(scope="session")
def capi():
    from pramaan_amd import _capi
    _capi.lib()
    return _capi


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding
    binding.lib()
    return binding
