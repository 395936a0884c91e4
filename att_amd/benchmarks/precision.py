"""SpMM precision sweep: N = 1,048,576, 20M entries (|i-j| <= 2 band plus
uniform off-band entries, seed 0), values N(0,1) FP32, X 1,048,576 x 128
U[-1,1] (seed 1); every mode's time and relative Frobenius error against an
FP64 host product (scipy) on a strided row sample (every 64th row), next to
torch.sparse.mm FP32 (int32 CSR indices) on the same GPU.

    python -m att_amd.benchmarks.precision
"""
import argparse
import json
import sys
import time

import numpy as np
import scipy.sparse as sp
import torch

from att_amd import graphs as G
from att_amd.benchmarks.plain import csr_tensor
from att_amd.precision import BlockMatrix

MODES = ("f32", "bf16x3", "bf16", "fp16", "fp8")


def _timed(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y = fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, y


def run(dev, nodes=1 << 20, nnz=20_000_000, dim=128, reps=10, sample_stride=64, build="device", log=None):
    """Returns {"workload", "sample", "modes": [...], "torch_sparse_fp32": {...}}."""
    t = time.perf_counter()
    g = G.band_random(nodes, nnz, 2, seed=0)
    X = np.random.default_rng(1).uniform(-1, 1, (g.n, dim)).astype(np.float32)
    rows = np.arange(0, g.n, sample_stride)
    counts = np.diff(g.indptr)[rows]
    sub_ip = np.concatenate([[0], np.cumsum(counts)])
    take = np.concatenate([np.arange(g.indptr[r], g.indptr[r + 1]) for r in rows])
    ref = sp.csr_matrix((g.values[take].astype(np.float64), g.indices[take], sub_ip),
                        shape=(rows.size, g.n)) @ X.astype(np.float64)
    nref = np.linalg.norm(ref)
    rows_d = torch.from_numpy(rows).to(dev)
    if log:
        log(f"precision leg: N={g.n} nnz={g.nnz}, FP64 sample of {rows.size} rows ({time.perf_counter() - t:.1f}s)")
    Xt = torch.from_numpy(X).to(dev)
    flops = 2.0 * g.nnz * dim
    out = []
    for mode in MODES:
        A = BlockMatrix(g, g.values, dev, mode, build=build)
        Xq, c = A.quantize_x(Xt)
        ms, y = _timed(lambda: A.matmul(None, Xq=Xq, col_scale=c), reps)
        err = float(np.linalg.norm(y[rows_d].double().cpu().numpy() - ref) / nref)
        out.append({"mode": mode, "ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 3),
                    "rel_fro_vs_fp64": err, "U16": A.n_unique, "parts": A.parts})
        del A, Xq, y
    Acsr = csr_tensor(g, g.values, dev)
    ms, y = _timed(lambda: torch.sparse.mm(Acsr, Xt), reps)
    err = float(np.linalg.norm(y[rows_d].double().cpu().numpy() - ref) / nref)
    return {"workload": f"band (|i-j|<=2) + uniform random, N={g.n}, nnz={g.nnz}, X {g.n}x{dim} U[-1,1]",
            "sample": f"FP64 reference on every {sample_stride}th row ({rows.size} rows)",
            "modes": out,
            "torch_sparse_fp32": {"ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 3), "rel_fro_vs_fp64": err,
                                  "indices": "int32"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nodes", type=int, default=1 << 20)
    ap.add_argument("--nnz", type=int, default=20_000_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    res = run(dev, args.nodes, args.nnz, args.dim, args.reps, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
