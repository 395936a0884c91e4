"""The C ABI from C: include/attgnn.h compiles as C99, libattgnn.so loads with
dlopen, and a C caller's window-block build equals the oracle's (CPU only)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from att_amd import _lib
from oracle import block as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def abi_smoke(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path_factory.mktemp("cabi") / "abi_smoke")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_abi", "abi_smoke.c"), "-o", exe, "-ldl"], check=True)
    return exe


def test_c_caller_matches_oracle(abi_smoke):
    out = subprocess.run([abi_smoke, _lib.LIB_PATH], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        name, *vals = line.split()
        got[name] = vals
    row = np.array([0, 3, 4, 7, 8, 10, 10, 12, 13, 15, 16, 18, 19], np.int32)
    col = np.array([0, 4, 9, 1, 7, 2, 5, 3, 0, 11, 6, 8, 7, 2, 10, 9, 1, 11, 0], np.int32)
    d = (1.0 / np.arange(1, 13, dtype=np.float32)).astype(np.float32)
    rp, c, val, win, at = B.block_balance(row, col, d, 8, 4, 2, B.FP32)
    for name, want in (("rp", rp), ("col", c), ("window", win), ("atomic", at)):
        assert np.array_equal(np.array(got[name], np.int64), np.asarray(want, np.int64)), name
    assert np.array_equal(np.array(got["val"], np.float32), np.asarray(val, np.float32))
    assert int(got["err_window"][0]) == -1
    assert int(got["err_format"][0]) == -2
    assert int(got["err_spmm"][0]) == -1


def test_header_compiles_as_c_and_cxx():
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("no C / C++ compiler")
    hdr = os.path.join(ROOT, "include", "attgnn.h")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c", "-fsyntax-only", hdr], check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-x", "c++", "-fsyntax-only", hdr], check=True)
