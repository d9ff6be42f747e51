import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def pytest_collection_modifyitems(config, items):
    # GPU tests are only skipped when explicitly deselected with -m "not gpu"; on a GPU run a missing
    # device or library is an error, never a silent fallback.
    pass


@pytest.fixture(scope="session")
def device():
    from solarenergy_chier_amd import Device

    dev = Device(0)
    yield dev
    dev.close()
