import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

FIXTURES = ROOT / "tests" / "fixtures"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs HIP kernels through libfvhip)")


@pytest.fixture(scope="session")
def fixtures_dir():
    return FIXTURES


@pytest.fixture(scope="session")
def ctx():
    import foamadapter_amd as fa
    c = fa.Context(0)
    yield c
    c.close()
