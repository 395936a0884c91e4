"""Window-block preprocessing: host (C++/OpenMP) against device (HIP) builds.

Graphs: the 2-layer GCN graph (Chung-Lu gamma 2.1, N = 232,965, ~114.6M
entries incl. self-loops, FP16 tiles of the normalised values, compact tiles)
and the attention graph (N = 169,343, ~2.48M entries, edge-id map).  Both
builds use 16-row windows and 32-column blocks; the device result is compared
bit for bit with the host result.  The host build is timed with all host
cores (up to 16); the device build with the CSR already in device memory,
from the first kernel to the last (its two size read-backs included).

    python -m att_amd.benchmarks.preprocess [--scale 1.0] [--reps 3]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from att_amd import graphs as G
from att_amd import ops
from att_amd.layers import _blocks_estimate, default_part_size


def _equal(host, dev, keys):
    for k in keys:
        a = host[k]
        b = dev[k].cpu()
        if a.dtype in (torch.float16, torch.bfloat16):
            a, b = a.view(torch.int16), b.view(torch.int16)
        if not torch.equal(a, b):
            return False
    return host["n_unique"] == dev["n_unique"]


def run(name, csr, src, vals, dtype, part_size, compact, trans, reps, dev, cores):
    rp = csr.rowptr32()
    col = csr.indices
    kw = dict(emit_empty=True, want_trans=trans, compact=compact)
    t = time.perf_counter()
    host = ops.blocks_build(rp, col, src, vals, dtype, 16, 32, part_size, nthreads=cores, **kw)
    t_host = time.perf_counter() - t
    rp_d = torch.from_numpy(rp).to(dev)
    col_d = torch.from_numpy(col).to(dev)
    v_d = None if vals is None else torch.from_numpy(np.asarray(vals, np.float32)).to(dev)
    times = []
    out = None
    for _ in range(reps + 1):
        out = None
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = ops.blocks_build_device(rp_d, col_d, src, v_d, dtype, 16, 32, part_size, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    t_dev = float(np.median(times[1:]))
    keys = ["rp", "col", "val", "window", "atomic"] + (["map_t"] if trans else []) + \
           (["masks", "block_base", "values", "positions"] if compact else [])
    same = _equal(host, out, keys)
    rec = {"graph": name, "nodes": csr.n, "nnz": csr.nnz, "src": src, "tiles": str(dtype).replace("torch.", ""),
           "compact": compact, "transposed_map": trans, "part_size": part_size, "U": host["n_unique"],
           "columns": int(host["col"].numel()), "host_s": round(t_host, 3), "host_threads": cores,
           "device_ms": round(t_dev * 1e3, 2), "speedup": round(t_host / t_dev, 1), "bit_identical": bool(same)}
    print(json.dumps(rec), flush=True)
    return same


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    cores = min(len(os.sched_getaffinity(0)), 16)
    ok = True
    t = time.perf_counter()
    n = int(232_965 * args.scale)
    nnz = int(114_600_000 * args.scale)
    g = G.chung_lu(n, (nnz - n) / n, 2.1, seed=0, target_nnz=nnz).padded(16)
    vals = G.gcn_norm_values(g, G.gcn_norm_degree(g))
    print(f"gcn graph N={g.n} nnz={g.nnz} ({time.perf_counter() - t:.1f}s)", file=sys.stderr, flush=True)
    ps = default_part_size(_blocks_estimate(g, 32))
    ok &= run("gcn-114M", g, "edge_f32", vals, torch.float16, ps, True, False, args.reps, dev, cores)
    del g, vals
    t = time.perf_counter()
    a = G.directed_chung_lu_symmetrized(int(169_343 * args.scale), int(2_480_000 * args.scale), seed=0).padded(16)
    print(f"attention graph N={a.n} nnz={a.nnz} ({time.perf_counter() - t:.1f}s)", file=sys.stderr, flush=True)
    ps = max(default_part_size(_blocks_estimate(a, 32)), 32)
    ok &= run("gat-2.48M", a, "map", None, torch.int32, ps, True, True, args.reps, dev, cores)
    if not ok:
        sys.exit("device build differs from the host build")


if __name__ == "__main__":
    main()
