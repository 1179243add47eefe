import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpfxhip on the device)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O

    return O.lib()


@pytest.fixture(scope="session")
def native_lib():
    from pfx_amd import _native as N

    return N.lib()


def pytest_collection_modifyitems(config, items):
    # keep the oracle's OpenMP pool modest on shared CPUs
    os.environ.setdefault("OMP_NUM_THREADS", "8")
