import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: large sizes")


def pytest_collection_modifyitems(config, items):
    # GPU tests fail loudly when no GPU is present: a test marked gpu is meant to be selected
    # with -m gpu on a machine with an MI355X. Nothing here skips them silently.
    pass


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import ora
    ora.build()
    return ora
