import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_collection_modifyitems(config, items):
    # GPU tests fail loudly without a GPU rather than skip: a GPU run that
    # silently skipped would read as green.
    pass


@pytest.fixture(scope="session")
def gpu():
    from remobench_amd import lib as L
    n = L.device_count()
    if n < 1:
        pytest.fail("no GPU visible: -m gpu tests need an MI355X")
    return n
