"""Writes the golden fixtures for the window-block preprocessing.

Run in a container that has the reference tree (``make -C oracle ref`` first):

    python tests/golden/make_golden.py

Expected outputs come from the reference's own CPU preprocessing
(gcn/Block_gcn/example.cpp, blockProcess_fp16_balance / blockProcess_tf32_balance)
compiled under oracle/_ref.  Inputs:
  * toy.npz   -- the 32-row, 38-entry CSR every reference test uses
                 (e.g. gcn/SpMM/test/mgcn/test_balance.py:20-22), with all-ones and
                 non-trivial degree vectors; full output arrays for every
                 window in {8,16}, wide in {4,8,16,32}, partSize in {1,2,32}.
  * cl2k.npz  -- a seeded Chung-Lu graph (N=2048, mean degree 12) with GCN
                 degrees; SHA-256 digests of every output array for the same
                 parameter grid (the arrays themselves would be megabytes).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import block as B  # noqa: E402

TOY_ROW = [0, 6, 7, 8, 10, 10, 11, 13, 15, 18, 19, 19, 19, 20, 20, 20, 20, 24,
           25, 26, 28, 28, 29, 31, 33, 36, 37, 37, 37, 38, 38, 38, 38]
TOY_COL = [1, 2, 3, 11, 16, 21, 1, 1, 2, 24, 25, 16, 22, 0, 27, 0, 4, 25, 9, 27, 1, 3, 11, 21, 1, 1, 2, 24,
           25, 16, 22, 0, 27, 0, 4, 25, 9, 27]
GRID = [(w, wd, ps) for w in (8, 16) for wd in (4, 8, 16, 32) for ps in (1, 2, 32)]
NAMES = ("rp", "col", "val", "window", "atomic")


def digest(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(str(a.dtype).encode() + str(a.shape).encode() + a.tobytes()).hexdigest()


def run_ref(ref, row, col, deg16, deg32, w, wd, ps):
    import torch
    r = torch.from_numpy(np.asarray(row, np.int32))
    c = torch.from_numpy(np.asarray(col, np.int32))
    o16 = ref.blockProcess_fp16_balance(r, c, torch.from_numpy(deg16), w, wd, ps)
    o32 = ref.blockProcess_tf32_balance(r, c, torch.from_numpy(deg32), w, wd, ps)
    return [t.numpy().copy() for t in o16], [t.numpy().copy() for t in o32]


def main():
    ref = B.load_reference()
    if ref is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    out = {"row": np.array(TOY_ROW, np.int32), "col": np.array(TOY_COL, np.int32)}
    rng = np.random.Generator(np.random.PCG64(7))
    out["deg16_ones"] = np.ones(32, np.float16)
    out["deg16"] = (rng.random(32) * 1.5 + 0.05).astype(np.float16)
    out["deg32"] = (rng.random(32) * 1.5 + 0.05).astype(np.float32)
    for w, wd, ps in GRID:
        for dname in ("deg16_ones", "deg16"):
            o16, o32 = run_ref(ref, out["row"], out["col"], out[dname], out["deg32"], w, wd, ps)
            for n, a in zip(NAMES, o16):
                out[f"fp16_{dname}_{w}_{wd}_{ps}_{n}"] = a
        for n, a in zip(NAMES, o32):
            out[f"tf32_{w}_{wd}_{ps}_{n}"] = a
    np.savez_compressed(os.path.join(HERE, "toy.npz"), **out)

    from att_amd import graphs as G
    g = G.chung_lu(2048, 12, 2.1, seed=0)
    d32 = G.gcn_norm_degree(g)
    d16 = d32.astype(np.float16)
    cl = {"row": g.rowptr32(), "col": g.indices, "deg16": d16, "deg32": d32}
    for w, wd, ps in GRID:
        o16, o32 = run_ref(ref, cl["row"], cl["col"], d16, d32, w, wd, ps)
        for n, a in zip(NAMES, o16):
            cl[f"sha_fp16_{w}_{wd}_{ps}_{n}"] = np.array(digest(a))
        for n, a in zip(NAMES, o32):
            cl[f"sha_tf32_{w}_{wd}_{ps}_{n}"] = np.array(digest(a))
    np.savez_compressed(os.path.join(HERE, "cl2k.npz"), **cl)
    print("wrote", os.path.join(HERE, "toy.npz"), os.path.join(HERE, "cl2k.npz"))


if __name__ == "__main__":
    main()
