import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc  # oracle/oracle.py: test infrastructure only
    orc.lib()
    orc.set_threads(1)
    return orc


@pytest.fixture(scope="session")
def gb():
    import gooseberry_amd
    gooseberry_amd.lib()
    return gooseberry_amd
