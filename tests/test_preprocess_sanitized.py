"""CPU: the host window-block preprocessing (att_amd/csrc/preprocess.cpp) built
with AddressSanitizer + UndefinedBehaviorSanitizer (g++, host code only) and
driven by tests/sanitize/preprocess_checks.cpp: the reference's toy CSR, random
graphs with hub and empty rows, empty graphs, every value source and output
type, malformed inputs, and the structure's invariants checked on each build.
Any out-of-bounds access, leak or undefined behaviour fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "sanitize", "preprocess_checks.cpp")
SOURCE = os.path.join(ROOT, "att_amd", "csrc", "preprocess.cpp")


def test_host_preprocessing_under_asan_ubsan(tmp_path):
    cxx = shutil.which(os.environ.get("CXX", "g++"))
    if cxx is None:
        pytest.skip("no C++ compiler")
    exe = str(tmp_path / "preprocess_checks")
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-Wall", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fopenmp", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "att_amd", "csrc"), DRIVER, SOURCE, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0 and ("cannot find -lasan" in r.stderr or "cannot find -lubsan" in r.stderr):
        pytest.skip("sanitizer runtimes not installed")
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and r.stdout.startswith("OK"), (r.stdout + r.stderr)[-4000:]
    assert int(r.stdout.split()[1]) > 600
