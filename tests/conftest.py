import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def wedge_path():
    return os.path.join(GOLDEN, "wedge.hmsh")


@pytest.fixture(scope="session")
def wedge_mesh(wedge_path):
    from cfd_optimisations_amd.formats import read_hmsh
    return read_hmsh(wedge_path)
