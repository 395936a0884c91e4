"""SpMM on the 16,384-node graph: plain FP32 on the host and the GPU against the
window-block kernels.

Graph: Chung-Lu (gamma 2.1, mean degree 12, seed 0), symmetrised with
self-loops, A_hat = D^-1/2 (A+I) D^-1/2 as FP32 CSR; X 16,384 x 64 FP32
U[-1, 1] (seed 1).  Timed: torch.sparse.mm (CSR, FP32) on the host cores and
on the GPU, and the window-block SpMM (16-row windows, 32-column blocks) in
FP32-MFMA, bf16x3 and FP16 (tiles and X rounded to FP16) modes; the 16-row
preprocessing on the host (16 threads) and on the device, whose outputs are
compared bit for bit.  Each GPU result reports its relative Frobenius error
against torch.sparse.mm FP32 on the same inputs.  Host timings: 3 warm-ups,
median of 10.

    python -m att_amd.benchmarks.spmm_small
"""
import json
import os
import sys
import time

import numpy as np
import torch

from att_amd import graphs as G
from att_amd import ops


def _t_gpu(fn, reps=50):
    for _ in range(5):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        out = fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, out


def run(dev):
    """The measurements as one record (see the module docstring)."""
    cores = min(len(os.sched_getaffinity(0)), 16)
    torch.set_num_threads(cores)
    g = G.chung_lu(16384, 12, 2.1, seed=0)
    d = G.gcn_norm_degree(g)
    vals = G.gcn_norm_values(g, d)
    n, nnz, D = g.n, g.nnz, 64
    X = np.random.default_rng(1).uniform(-1, 1, (n, D)).astype(np.float32)
    flops = 2.0 * nnz * D
    # plain FP32 on the host
    A_cpu = torch.sparse_csr_tensor(torch.from_numpy(g.indptr.astype(np.int32)), torch.from_numpy(g.indices),
                                    torch.from_numpy(vals), (n, n))
    Xc = torch.from_numpy(X)
    for _ in range(3):
        torch.sparse.mm(A_cpu, Xc)
    ts = []
    for _ in range(10):
        t = time.perf_counter()
        torch.sparse.mm(A_cpu, Xc)
        ts.append(time.perf_counter() - t)
    t_cpu = float(np.median(ts)) * 1e3
    rec = {"op": "spmm", "graph": "chung-lu 16,384", "nnz": nnz, "dim": D,
           "cpu_torch_sparse_fp32": {"ms": round(t_cpu, 4), "gflops": round(flops / t_cpu / 1e6, 2), "threads": cores}}
    # plain FP32 on the GPU
    A = A_cpu.to(dev)
    Xd = Xc.to(dev)
    t_plain, ref = _t_gpu(lambda: torch.sparse.mm(A, Xd))
    rec["gpu_torch_sparse_fp32"] = {"ms": round(t_plain, 4), "gflops": round(flops / t_plain / 1e6, 2)}
    # preprocessing: host vs device, bit for bit
    t = time.perf_counter()
    hb = ops.blocks_build(g.rowptr32(), g.indices, "edge_f32", vals, torch.float32, 16, 32, 8, emit_empty=True,
                          nthreads=cores)
    t_host = (time.perf_counter() - t) * 1e3
    rp_d = torch.from_numpy(g.rowptr32()).to(dev)
    col_d = torch.from_numpy(g.indices).to(dev)
    v_d = torch.from_numpy(vals).to(dev)
    t_dev, db = _t_gpu(lambda: ops.blocks_build_device(rp_d, col_d, "edge_f32", v_d, torch.float32, 16, 32, 8,
                                                       emit_empty=True), reps=20)
    same = all(torch.equal(hb[k], db[k].cpu()) for k in ("rp", "col", "val", "window", "atomic"))
    rec["preprocess_16row"] = {"host_ms": round(t_host, 3), "device_ms": round(t_dev, 3), "bit_identical": same}
    # window-block SpMM modes
    arrs = [db[k] for k in ("rp", "col", "val", "window", "atomic")]
    modes = {}
    for name, mode in (("f32_mfma", "f32"), ("bf16x3", "bf16x3")):
        ms, y = _t_gpu(lambda: ops.spmm_block(*arrs, Xd, n, 16, 32, mode=mode, rows_covered=True))
        modes[name] = {"ms": round(ms, 4), "gflops": round(flops / ms / 1e6, 2),
                       "rel_err_vs_torch_sparse": float((y - ref).norm() / ref.norm())}
    b16 = ops.blocks_build_device(rp_d, col_d, "edge_f32", v_d, torch.float16, 16, 32, 8, emit_empty=True)
    a16 = [b16[k] for k in ("rp", "col", "val", "window", "atomic")]
    X16 = Xd.half()
    ms, y = _t_gpu(lambda: ops.spmm_block(*a16, X16, n, 16, 32, rows_covered=True, out_dtype=torch.float32))
    modes["fp16"] = {"ms": round(ms, 4), "gflops": round(flops / ms / 1e6, 2),
                     "rel_err_vs_torch_sparse": float((y - ref).norm() / ref.norm())}
    rec["window_block"] = modes
    return rec


def main():
    rec = run(torch.device("cuda", 0))
    print(json.dumps(rec), flush=True)
    if not rec["preprocess_16row"]["bit_identical"]:
        sys.exit("device preprocessing differs from the host build")


if __name__ == "__main__":
    main()
