"""Build libattgnn.so in-tree (host C++ with g++/OpenMP, device code with hipcc for gfx950).

Usage: ``python -m att_amd.build``.  The library lands in ``att_amd/_build/``;
it is git-ignored but travels with the working tree.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(OUT_DIR, "libattgnn.so")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["spmm.hip", "sddmm.hip", "edge.hip", "pack.hip", "preprocess_gpu.hip", "gatv2.hip", "transpose.hip",
               "loss.hip"]
HOST_SOURCES = ["preprocess.cpp"]
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions are disabled in device code: hipcc -O3 emits them
# from plain float code and some of their forms have produced wrong results.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I" + INCLUDE, "-I" + CSRC,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-I" + INCLUDE, "-I" + CSRC, "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build(verbose=False, jobs=None):
    """Compile every source (``jobs`` at a time, default: all) and link
    libattgnn.so (always rebuilds)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    cmds, objs = [], []
    for src in HOST_SOURCES:
        obj = os.path.join(OUT_DIR, src + ".o")
        cmds.append([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(OUT_DIR, src + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    workers = jobs or min(len(cmds), os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        list(pool.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lgomp"])
    shutil.move(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
