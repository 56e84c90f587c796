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
C2H4 = ([6, 6, 1, 1, 1, 1], [[3.402, 0.773, -9.252], [4.697, 0.791, -8.909],     # :51-58
                             [2.933, -0.150, -9.521], [2.837, 1.682, -9.258],
                             [5.262, -0.118, -8.904], [5.167, 1.714, -8.641]])
CO2 = ([6, 8, 8], [[0.0, 0.0, 0.32942], [0.0, 1.15362, -0.12353], [0.0, -1.15362, -0.12353]])  # :18-23
