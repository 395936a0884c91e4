"""Arithmetic modes of the width-256 SpMM on the GCN benchmark's graph:
Chung-Lu gamma = 2.1, N = 232,965, 114.6M entries after symmetrisation and
self-loops, A_hat = D^-1/2 (A + I) D^-1/2 in FP32, X 232,965 x 256 N(0,1)
(seed 1).  Each mode builds the graph's compact window-block tiles on the
device in its tile type and runs Y = A_hat X with FP32 output:

  fp16 / bf16  16-bit tiles and X, MFMA in that type, FP32 accumulation
  bf16x3       FP32 tiles and X, each split into bf16 hi + lo (3 MFMA terms)
  f32          FP32 tiles and X, FP32 MFMA (16x16x4)

Reported per mode: time per SpMM (HIP events), TFLOP/s (2 nnz D); the
compulsory bytes per second (X read once, the compact structure as read --
per part its rp pair, window and flag, per condensed column a 4-byte id and a
2-byte row mask, per 32-column block an 8-byte offset, per entry its value --
and Y written once) as a fraction of the 8 TB/s HBM peak; beside it the
per-window gather bytes 4(M+1) + nnz (4 + s_A) + U D s_X + M D 4 (U = distinct
columns summed over 16-row windows) per second, which the L2s and the
Infinity Cache mostly serve; and the relative Frobenius error against an
FP64 product of the unrounded FP32 A_hat and X on a sample of 2,000 rows
(scipy)."""
import argparse
import json
import sys
import time

import numpy as np
import scipy.sparse as sp
import torch

from att_amd import graphs as G
from att_amd import ops
from att_amd.layers import GCNGraph

HBM_PEAK_GBS = 8000.0
N_NODES, TARGET_NNZ, DIM = 232_965, 114_600_000, 256
MODES = {  # tile dtype, X dtype, kernel mode
    "fp16": (torch.float16, torch.float16, None),
    "bf16": (torch.bfloat16, torch.bfloat16, None),
    "bf16x3": (torch.float32, torch.float32, "bf16x3"),
    "f32": (torch.float32, torch.float32, "f32"),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--modes", default=",".join(MODES))
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--scale", type=float, default=1.0, help="graph size fraction (testing only)")
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    t = time.perf_counter()
    n, nnz = int(N_NODES * args.scale), int(TARGET_NNZ * args.scale)
    g = G.chung_lu(n, (nnz - n) / n, 2.1, seed=0, target_nnz=nnz)
    vals = G.gcn_norm_values(g, G.gcn_norm_degree(g))
    X = torch.randn(g.n, DIM, generator=torch.Generator().manual_seed(1))
    rows = np.arange(0, g.n, max(g.n // 2000, 1))
    sub_ip = np.concatenate([[0], np.cumsum(np.diff(g.indptr)[rows])])
    sub_ix = np.concatenate([g.indices[g.indptr[r]:g.indptr[r + 1]] for r in rows])
    sub_v = np.concatenate([vals[g.indptr[r]:g.indptr[r + 1]] for r in rows]).astype(np.float64)
    ref = sp.csr_matrix((sub_v, sub_ix, sub_ip), shape=(rows.size, g.n)) @ X.double().numpy()
    nref = np.linalg.norm(ref)
    rows_d = torch.from_numpy(rows).to(dev)
    print(f"graph N={g.n} nnz={g.nnz}, FP64 row-sample reference ({time.perf_counter() - t:.1f}s)",
          file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream(dev)
    for name in args.modes.split(","):
        tile_dt, x_dt, kmode = MODES[name]
        graph = GCNGraph(g, dev, window=16, wide=32, tile_dtype=tile_dt, symmetric=True, values=vals,
                         compact=True, build="device")
        _, col, _, _, _ = graph.fwd["arrays"]
        rp, win, at = graph.fwd["tasks"]
        masks, base, values = graph.fwd["compact"]
        Xq = X.to(x_dt).to(dev)

        def run():
            return ops.spmm_compact(rp, col, masks, base, values, win, at, Xq, graph.num_nodes_ori, 16,
                                    mode=kmode, out_dtype=torch.float32, rows_covered=True)

        for _ in range(3):
            y = run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.reps):
            y = run()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.reps
        s_a = torch.empty(0, dtype=tile_dt).element_size()
        s_x = Xq.element_size()
        P, C = win.numel(), col.numel()
        compulsory = (g.n * DIM * s_x + 4 * (2 * P + 1) + 8 * P + C * (4 + 2) + (C // 32) * 8 + g.nnz * s_a
                      + g.n * DIM * 4)
        gather = 4 * (g.n + 1) + g.nnz * (4 + s_a) + graph.n_unique * DIM * s_x + g.n * DIM * 4
        err = float(np.linalg.norm(y[rows_d].double().cpu().numpy() - ref) / nref)
        print(json.dumps({"mode": name, "nodes": g.n, "nnz": g.nnz, "dim": DIM, "U16": graph.n_unique,
                          "ms": round(ms, 4), "tflops": round(2.0 * g.nnz * DIM / ms / 1e9, 2),
                          "compulsory_gbs": round(compulsory / ms / 1e6, 1),
                          "hbm_frac": round(compulsory / ms / 1e6 / HBM_PEAK_GBS, 4),
                          "gather_gbs": round(gather / ms / 1e6, 1),
                          "rel_fro_vs_fp64": err}), flush=True)
        del graph, masks, base, values, rp, col, win, at, Xq, y
        torch.cuda.empty_cache()


if __name__ == "__main__":
    sys.exit(main())
