import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
RATES = os.path.join(GOLDEN, "atomic_rates_Hui.txt")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def rates_path():
    return RATES


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


def print_case(*a):
    """name the case before a GPU step (flushed, so a dying box still shows it)"""
    print("[case]", *a, flush=True)
