"""Manifest of the benchmark workloads: for each configuration the generator
call, and what it produces -- nodes, entries, maximum degree, U (distinct
columns summed over 16-row windows, the gather count of the SpMM byte model),
the condensed columns padded to 32 per window (the compact structure's
column count, whatever the part size) and a SHA-256 of the CSR arrays -- so a
run can show it measured the stated graph.

    python -m att_amd.benchmarks.manifest            # writes manifest.json beside this file
    python -m att_amd.benchmarks.manifest --check    # regenerates and compares

The generators are the ones the benchmarks call (bench.py, gat.py,
precision.py, spmm_small.py, serving.py): configuration 1 the 16,384-node
Chung-Lu graph, 2 the 114.6M-entry GCN graph, 3 the 2.48M-entry attention
graph, 4 the 2^20-node band + random matrix, 5 the serving pool of 1,024
Chung-Lu graphs."""
import argparse
import hashlib
import json
import os
import sys

import numpy as np

from att_amd import graphs as G

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "manifest.json")
WINDOW, WIDE = 16, 32


def csr_digest(g):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(g.indptr, np.int64).tobytes())
    h.update(np.ascontiguousarray(g.indices, np.int32).tobytes())
    if g.values is not None:
        h.update(np.ascontiguousarray(g.values, np.float32).tobytes())
    return h.hexdigest()[:16]


def window_columns(g, window=WINDOW, wide=WIDE):
    """(U, padded): distinct columns summed over `window`-row windows, and the
    same with each window's count rounded up to a multiple of `wide`."""
    win = (g.row_ids().astype(np.int64) // window)
    keys = np.unique(win * np.int64(g.n) + g.indices.astype(np.int64))
    counts = np.bincount(keys // np.int64(g.n))
    return int(keys.size), int(((counts + wide - 1) // wide * wide).sum())


def describe(g):
    u, padded = window_columns(g)
    return {"nodes": int(g.n), "nnz": int(g.nnz), "max_degree": int(g.degrees().max()) if g.n else 0,
            "U16": u, "condensed_cols_w32": padded, "csr_sha256_16": csr_digest(g)}


GENERATORS = {
    "cfg1_chung_lu_16k": ("graphs.chung_lu(16384, 12, 2.1, seed=0)",
                          lambda: G.chung_lu(16384, 12, 2.1, seed=0)),
    "cfg2_gcn_114M": ("graphs.chung_lu(232965, (114600000 - 232965) / 232965, 2.1, seed=0, target_nnz=114600000)",
                      lambda: G.chung_lu(232_965, (114_600_000 - 232_965) / 232_965, 2.1, seed=0,
                                         target_nnz=114_600_000)),
    "cfg3_attention_2.48M": ("graphs.directed_chung_lu_symmetrized(169343, 2480000, seed=0)",
                             lambda: G.directed_chung_lu_symmetrized(169_343, 2_480_000, seed=0)),
    "cfg4_band_random_2^20": ("graphs.band_random(1 << 20, 20000000, 2, seed=0)",
                              lambda: G.band_random(1 << 20, 20_000_000, 2, seed=0)),
}


def serving_pool(size=1024, seed=0):
    """Configuration 5: the serving pool's graph sizes and a digest over all
    its CSRs (features are seeded per key and not part of the digest)."""
    from att_amd.serve import GraphPool
    pool = GraphPool(size, seed=seed)
    h = hashlib.sha256()
    nodes, nnz = [], []
    for k in range(size):
        csr, _ = pool(k)
        pool.memo.pop(k)
        h.update(csr_digest(csr).encode())
        nodes.append(csr.n)
        nnz.append(csr.nnz)
    nodes, nnz = np.array(nodes), np.array(nnz)
    return {"call": f"serve.GraphPool({size}, seed={seed})", "graphs": size,
            "nodes_total": int(nodes.sum()), "nodes_min": int(nodes.min()), "nodes_median": float(np.median(nodes)),
            "nodes_max": int(nodes.max()), "nnz_total": int(nnz.sum()), "nnz_max": int(nnz.max()),
            "pool_sha256_16": h.hexdigest()[:16]}


def build(names=None, pool=True):
    out = {}
    for name, (call, gen) in GENERATORS.items():
        if names and name not in names:
            continue
        out[name] = {"call": call, **describe(gen())}
        print(name, out[name], file=sys.stderr, flush=True)
    if pool and (not names or "cfg5_serving_pool" in names):
        out["cfg5_serving_pool"] = serving_pool()
        print("cfg5_serving_pool", out["cfg5_serving_pool"], file=sys.stderr, flush=True)
    return out


def load():
    with open(PATH) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true", help="regenerate and compare with manifest.json")
    ap.add_argument("--only", default="", help="comma-separated configuration names")
    args = ap.parse_args()
    names = [s for s in args.only.split(",") if s] or None
    got = build(names)
    if args.check:
        want = load()
        bad = [k for k in got if got[k] != want.get(k)]
        print("mismatch: " + ", ".join(bad) if bad else "manifest matches", flush=True)
        return 1 if bad else 0
    with open(PATH, "w") as f:
        json.dump(got, f, indent=1)
        f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
