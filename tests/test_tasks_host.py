"""The library's task-table arithmetic (lible_amd/csrc/tasks.hpp: Schwarz cuts, the
lane-block-major enumeration of the one-lane J/K kernels, the shell-ordered lane side and the
bounds check run before every launch) against brute-force enumerations, built with g++ under
the address and undefined-behaviour sanitizers and run on the CPU (tests/cpp/tasks_check.cpp)."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT


def test_task_tables_sanitized(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = os.path.join(ROOT, "tests", "cpp", "tasks_check.cpp")
    exe = str(tmp_path / "tasks_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-static-libasan", "-static-libubsan", src, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "all checks passed" in r.stdout
