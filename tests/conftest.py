import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X")


@pytest.fixture(scope="session")
def orc():
    from oracle_lib import load_oracle
    return load_oracle()


@pytest.fixture(scope="session")
def hip():
    import bonsai_amd
    if bonsai_amd.device_count() == 0:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return bonsai_amd
