"""GAT / GATv2 layer forward+backward on the 169,343-node, ~2.48M-entry graph.

Directed Chung-Lu draws symmetrised with self-loops (seed 0); features
169,343 x 128 BF16 ~ N(0,1); heads in {1,2,4,8} of width 64; LeakyReLU 0.2;
attention dropout 0 (timing) -- the window-block path (BF16 storage, FP32
accumulation and softmax) against plain FP32 PyTorch message passing on the
same GPU.  Times are medians of CUDA-event-timed fwd+bwd iterations.
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from att_amd import graphs as G
from att_amd.benchmarks.plain import gat_fp32, gatv2_true_fp32
from att_amd.layers import GATConv_multi, GATGraph, GATv2Conv_multi


def timeit(fn, reps, warm=2):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--heads", default="1,2,4,8")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--nodes", type=int, default=169_343)
    ap.add_argument("--nnz", type=int, default=2_480_000)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 path and the SDDMM variants")
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    t = time.perf_counter()
    g = G.directed_chung_lu_symmetrized(args.nodes, args.nnz, seed=0)
    graph = GATGraph(g, dev, build="device")
    print(f"graph N={g.n} nnz={g.nnz} U16={graph.n_unique} ({time.perf_counter() - t:.1f}s)", file=sys.stderr)
    rows = torch.from_numpy(g.row_ids().astype(np.int64)).to(dev)
    cols = torch.from_numpy(g.indices.astype(np.int64)).to(dev)
    X = torch.randn(g.n, 128, generator=torch.Generator().manual_seed(1)).to(dev)
    Xb = X.bfloat16()
    dim = 64
    # SDDMM alone: out[h, e=(i,j)] = <A[i,h], B[j,h]> in BF16, fp32 out.
    # Algorithmic bytes: 4(M+1) + nnz*(4+4) + M*D*H*2 + U*D*H*2 (U: distinct
    # columns summed over 16-row windows), flops 2*nnz*D*H.
    from att_amd import ops
    rp, col, emap, win, _ = graph.fwd
    for heads in ([] if args.no_plain else [int(h) for h in args.heads.split(",")]):
        A = torch.randn(g.n, heads, dim, device=dev).bfloat16()
        B = torch.randn(g.n, heads, dim, device=dev).bfloat16()
        variants = {"dense_map": lambda: ops.sddmm_block(rp, col, emap, win, A, B, g.nnz, graph.window, graph.wide,
                                                         layout="mhd", out_dtype=torch.float32),
                    "edge_list": lambda: graph.sddmm(A, B)}
        for name, f in variants.items():
            ms = timeit(f, args.reps)
            M = g.n
            nbytes = 4 * (M + 1) + g.nnz * 8 + M * dim * heads * 2 + graph.n_unique * dim * heads * 2
            print(json.dumps({"op": "sddmm", "variant": name, "heads": heads, "dim": dim, "nnz": g.nnz,
                              "ms": round(ms, 4), "gflops": round(2.0 * g.nnz * dim * heads / ms / 1e6, 1),
                              "algorithmic_gbs": round(nbytes / ms / 1e6, 1),
                              "hbm_frac": round(nbytes / ms / 1e6 / 8000.0, 3)}), flush=True)
    # gatv2: the single-head reference file's logits; gatv2_minmax: the
    # multi-head file's per-head min-max logit map; gatv2_true: the paper's GATv2
    for kind in ("gat", "gatv2", "gatv2_minmax", "gatv2_true"):
        for heads in [int(h) for h in args.heads.split(",")]:
            torch.manual_seed(heads)
            if kind == "gat":
                layer = GATConv_multi(128, dim, 0.0, 0.2, heads, dtype=torch.bfloat16).to(dev)
            else:
                layer = GATv2Conv_multi(128, dim, 0.0, 0.2, heads, dtype=torch.bfloat16,
                                        true_gatv2=(kind == "gatv2_true"), minmax=(kind == "gatv2_minmax")).to(dev)
            go = torch.randn(g.n, heads, dim, device=dev)
            go_b = go.bfloat16()            # the upstream gradient in the layer's output type

            def ours():
                layer.zero_grad(set_to_none=True)
                out = layer(Xb, graph)
                out.backward(go_b)

            if kind == "gat":
                W = layer.fc.weight.detach().clone().requires_grad_(True)
                a0 = layer.a0.detach().clone().requires_grad_(True)
                a1 = layer.a1.detach().clone().requires_grad_(True)

                def plain():
                    out = gat_fp32(X, W, a0, a1, rows, cols, g.n, heads, dim)
                    out.backward(go)
            elif kind in ("gatv2", "gatv2_minmax"):
                W0 = layer.fc0.weight.detach().clone().requires_grad_(True)
                W1 = layer.fc1.weight.detach().clone().requires_grad_(True)
                a = layer.a0.detach().clone().requires_grad_(True)
                mm = kind == "gatv2_minmax"

                def plain():
                    out = gat_fp32(X, W0, a, a, rows, cols, g.n, heads, dim, slope=1.0, Wt=W1, v2_alpha=0.2,
                                   minmax=mm)
                    out.backward(go)
            else:
                W0 = layer.fc0.weight.detach().clone().requires_grad_(True)
                W1 = layer.fc1.weight.detach().clone().requires_grad_(True)
                a = layer.a0.detach().clone().view(heads, dim).requires_grad_(True)

                def plain():
                    out = gatv2_true_fp32(X, W0, W1, a, rows, cols, g.n, heads, dim, slope=0.2)
                    out.backward(go)

            t_ours = timeit(ours, args.reps)
            extra = {}
            if kind != "gatv2_true":
                # the same layer with the probabilities rebuilt in the aggregations instead of stored
                layer.fused = True
                extra["ms_fwd_bwd_fused_p"] = round(timeit(ours, args.reps), 3)
                layer.fused = False
            if args.no_plain:
                print(json.dumps({"layer": kind, "heads": heads, "dim": dim, "nodes": g.n, "nnz": g.nnz,
                                  "ms_fwd_bwd": round(t_ours, 3), **extra}), flush=True)
                continue
            t_plain = timeit(plain, max(args.reps // 2, 3))
            # accuracy of the forward output against the FP32 path on the same inputs
            with torch.no_grad():
                out = layer(Xb, graph).float()
                if kind == "gat":
                    ref = gat_fp32(Xb.float(), W.bfloat16().float(), a0, a1, rows, cols, g.n, heads, dim)
                elif kind in ("gatv2", "gatv2_minmax"):
                    ref = gat_fp32(Xb.float(), W0.bfloat16().float(), a, a, rows, cols, g.n, heads, dim, slope=1.0,
                                   Wt=W1.bfloat16().float(), v2_alpha=0.2, minmax=(kind == "gatv2_minmax"))
                else:
                    ref = gatv2_true_fp32(Xb.float(), W0.bfloat16().float(), W1.bfloat16().float(), a, rows, cols,
                                          g.n, heads, dim)
                err = ((out - ref).norm() / ref.norm()).item()
            print(json.dumps({"layer": kind, "heads": heads, "dim": dim, "nodes": g.n, "nnz": g.nnz,
                              "ms_fwd_bwd": round(t_ours, 3), **extra, "plain_fp32_ms": round(t_plain, 3),
                              "speedup": round(t_plain / t_ours, 2), "rel_err_vs_fp32": err}), flush=True)


if __name__ == "__main__":
    main()
