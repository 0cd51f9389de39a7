import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libhstgpu.so)")
    config.addinivalue_line("markers", "slow: longer CPU checks")


def deck_path(name: str) -> str:
    return os.path.join(GOLDEN, "decks", name + ".in")


@pytest.fixture(scope="session")
def oracle_built():
    from oracle import oracle

    oracle.build()
    return oracle
