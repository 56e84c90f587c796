"""The lible::ints C++ interface (include/lible/ints/ints.hpp), driven through the C++
known-answer program tests/cpp/ints_selftest.cpp, which mirrors the reference's
test/tests_ints.cpp checks of eri4 / eri4Diagonal / eri3 / eri2 / eri2Diagonal."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

BASIS = os.path.join(ROOT, "tests", "golden", "basis")


def selftest() -> str:
    from lible_amd import build as b
    if not os.path.exists(b.OUT):
        b.build(verbose=False)
    elif not os.path.exists(b.SELFTEST) or \
            os.path.getmtime(b.SELFTEST) < os.path.getmtime(b.SELFTEST_SRC):
        b.build_selftest()
    return b.SELFTEST


def run(mode: str, timeout: int):
    r = subprocess.run([selftest(), BASIS, mode], capture_output=True, text=True,
                       timeout=timeout)
    lines = [l for l in r.stdout.splitlines() if l.startswith(("PASS", "FAIL"))]
    failed = [l for l in lines if l.startswith("FAIL")]
    assert r.returncode == 0 and not failed and lines, \
        f"rc={r.returncode}\n{r.stdout}\n{r.stderr}"
    return lines


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text("#include <lible/ints/ints.hpp>\n#include <lbx.h>\nint main() { return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        f"-I{os.path.join(ROOT, 'include')}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_interface_host_side():
    lines = run("cpu", 60)
    assert len(lines) == 11
    assert any(l.startswith("PASS shell_pair_data_ctor") for l in lines)


@pytest.mark.gpu
def test_cpp_interface_known_answers():
    lines = run("gpu", 300)
    names = {l.split()[1].rstrip(":") for l in lines}
    assert {"eri4", "eri4_symmetry", "coulomb_exchange", "coulomb_exchange_gradient",
            "one_electron_overlap", "one_electron_kinetic", "one_electron_nuclear",
            "one_electron_external_charges",
            "eri4_diagonal", "eri3", "eri2",
            "eri2_diagonal", "ghost_atoms", "deploy_eri4_kernel", "deploy_eri3_kernel",
            "deploy_eri2_kernel", "deploy_eri4d1_kernel", "deploy_eri3d1_kernel",
            "deploy_eri2d1_kernel", "deploy_eri2d2_kernel", "deploy_eri4soc_kernel",
            "deploy_eri3soc_kernel", "eri4_benchmark", "g_shells_def2qzvp"} <= names
