import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "polymult-challenge_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpolymult.so on the device)")
    config.addinivalue_line("markers", "slow: full-size CPU oracle work (tens of seconds)")


@pytest.fixture(scope="session")
def O():
    from oracle import oracle

    return oracle


@pytest.fixture(scope="session")
def pm():
    import pmlib

    pmlib.load()
    return pmlib
