import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP backend)")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture
def progress(capsys):
    """Progress lines of a long test, written to the terminal as they happen
    (not held back by output capture)."""
    def say(msg):
        with capsys.disabled():
            print(msg, flush=True)
    return say


@pytest.fixture(scope="session")
def nemo_mod():
    import nemosim_amd
    return nemosim_amd


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.load()
    return oracle
