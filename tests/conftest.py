import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
BASIS = os.path.join(GOLDEN, "basis")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def basis_path(name: str) -> str:
    return os.path.join(BASIS, name + ".txt")


O3 = ([8, 8, 8], [[0, 0, 0], [0, 1, 1], [0, 0, 2]])                      # tests_ints.cpp:115-120
H2O = ([8, 1, 1], [[0.0, 0.0, 0.0], [0.757, 0.586, 0.0], [-0.757, 0.586, 0.0]])  # :68-73
