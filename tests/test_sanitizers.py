"""The host library code and the CPU oracle under AddressSanitizer and
UndefinedBehaviorSanitizer, and the library's host entry points called from
several threads under ThreadSanitizer.

tests/sanitize_driver.cpp calls the scene generators, the kd-tree build and
export, the camera and bounce ray generators, the Perlin table, the .rsdf
header checks and the oracle's traversal, attribute, terrain, edit and
network routines on small inputs with their edge cases.  It is built here with
g++ (host code only; no HIP source is part of it) from the library's host
sources and the oracle sources, and must finish with no sanitizer report.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HOST = sorted((ROOT / "rayengine_amd" / "csrc" / "host").glob("*.cpp"))
ORACLE = sorted((ROOT / "oracle").glob("*.cpp"))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    out = tmp_path_factory.mktemp("san") / "sanitize_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-ffp-contract=off", "-fopenmp", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "sanitize_driver.cpp"), *map(str, HOST), *map(str, ORACLE), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def test_host_code_and_oracle_are_sanitizer_clean(driver, tmp_path):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=1:strict_string_checks=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run([str(driver), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-6000:]
    assert "0 failed checks" in r.stdout


def test_host_entry_points_are_thread_safe(tmp_path):
    """tests/tsan_driver.cpp under ThreadSanitizer: the context-free host
    entry points called from six threads at once give the one-thread results
    and no race report."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    out = tmp_path / "tsan_driver"
    # the library's host sources only (the oracle needs the OpenMP runtime)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=thread", "-ffp-contract=off",
           f"-I{ROOT / 'include'}", str(ROOT / "tests" / "tsan_driver.cpp"), *map(str, HOST), "-pthread", "-o",
           str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")
    r = subprocess.run([str(out)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-6000:]
    assert "0 failed checks" in r.stdout
