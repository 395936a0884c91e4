"""SpMM precision sweep: N = 1,048,576, 20M entries (|i-j| <= 2 band plus
uniform off-band entries, seed 0), values N(0,1) FP32, X 1,048,576 x 128
U[-1,1]; every mode's relative Frobenius error against an FP64 host product
(scipy) and its time, next to torch.sparse.mm FP32 on the same GPU."""
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nodes", type=int, default=1 << 20)
    ap.add_argument("--nnz", type=int, default=20_000_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    t = time.perf_counter()
    g = G.band_random(args.nodes, args.nnz, 2, seed=0)
    X = np.random.default_rng(1).uniform(-1, 1, (g.n, args.dim)).astype(np.float32)
    ref = sp.csr_matrix((g.values.astype(np.float64), g.indices, g.indptr), shape=(g.n, g.n)) @ X.astype(np.float64)
    nref = np.linalg.norm(ref)
    print(f"matrix N={g.n} nnz={g.nnz}, FP64 reference ({time.perf_counter() - t:.1f}s)", file=sys.stderr)
    Xt = torch.from_numpy(X).to(dev)
    flops = 2.0 * g.nnz * args.dim
    rows = []
    for mode in ("f32", "bf16x3", "bf16", "fp16", "fp8"):
        A = BlockMatrix(g, g.values, dev, mode)
        Xq, c = A.quantize_x(Xt)
        y = A.matmul(None, Xq=Xq, col_scale=c)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.reps):
            y = A.matmul(None, Xq=Xq, col_scale=c)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.reps
        err = float(np.linalg.norm(y.cpu().numpy().astype(np.float64) - ref) / nref)
        rows.append({"mode": mode, "ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 2), "rel_fro_vs_fp64": err,
                     "U16": A.n_unique, "parts": A.parts})
        print(json.dumps(rows[-1]), flush=True)
        del A
    Acsr = csr_tensor(g, g.values, dev)
    torch.sparse.mm(Acsr, Xt)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.reps):
        y = torch.sparse.mm(Acsr, Xt)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.reps
    err = float(np.linalg.norm(y.cpu().numpy().astype(np.float64) - ref) / nref)
    print(json.dumps({"mode": "torch.sparse fp32", "ms": round(ms, 4), "tflops": round(flops / ms / 1e9, 2),
                      "rel_fro_vs_fp64": err}), flush=True)


if __name__ == "__main__":
    main()
