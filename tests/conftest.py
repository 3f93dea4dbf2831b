import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built native library")
    config.addinivalue_line("markers", "slow: long CPU-only scenario")


@pytest.fixture(scope="session")
def lib():
    import coffee_sim_amd as cs

    return cs.load_library()


@pytest.fixture(scope="session")
def orc():
    import orc as _orc

    _orc.lib()
    return _orc


@pytest.fixture(scope="session")
def ctx(lib):
    import coffee_sim_amd as cs

    if cs.device_count() < 1:
        pytest.fail("a gpu test ran without a visible HIP device")
    c = cs.Context(0)
    yield c
    c.close()
