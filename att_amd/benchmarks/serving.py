"""Serving benchmark: Poisson arrivals of 2-layer GCN / 1-layer 4-head GAT
inference requests (50/50) on Chung-Lu graphs with N ~ lognormal(median 2000,
sigma 1.2) clipped to [100, 200000], mean degree U[4, 32], 128 BF16 features.

Compared: the continuous-batching server (on-device structure cache, HBM-budget
admission, block-diagonal packing, HIP-graph replays) and plain FP32 PyTorch
launched per graph, FIFO.  Steps: (1) both servers' saturated throughput (all
requests queued at t = 0); (2) the plain server at 70% of its capacity, whose
p99 latency becomes the bound; (3) the batched server at increasing multiples
of the plain capacity -- the largest multiple whose p99 stays within the bound
is reported.  The graph cache is warmed first (steady state); cold-cache
preprocessing cost is reported separately.
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from att_amd import serve as S


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--pool", type=int, default=1024)
    ap.add_argument("--plain-requests", type=int, default=1024)
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of HIP graphs")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--part-size", type=int, default=S.SERVE_PART_SIZE, help="blocks per part of cached graphs")
    ap.add_argument("--preprocess", choices=["host", "device"], default="device",
                    help="where cold-cache graphs are preprocessed")
    ap.add_argument("--compute-streams", type=int, default=2, help="compute streams the batches alternate over")
    ap.add_argument("--saturated-only", action="store_true",
                    help="warm the cache, then one saturated batched run (for kernel traces)")
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    models = S.Models(dev)
    pool = S.GraphPool(args.pool, seed=args.seed)
    t = time.perf_counter()
    for k in range(args.pool):
        pool(k)
    print(f"pool of {args.pool} graphs generated ({time.perf_counter() - t:.1f}s); "
          f"nodes median {int(np.median(pool.n))} max {int(pool.n.max())}", file=sys.stderr)

    # cold cache: every graph preprocessed on first use, on the host and on the
    # device (the server keeps the latter's cache)
    cold = {}
    for where in ((args.preprocess,) if args.saturated_only else ("host", "device")):
        srv = S.Server(dev, models, pool, use_graph=not args.no_graph, part_size=args.part_size, preprocess=where,
                       compute_streams=args.compute_streams)
        warm = [S.Request(i, 0.0, "gcn" if i % 2 else "gat", i) for i in range(args.pool)]
        torch.cuda.synchronize()
        t = time.perf_counter()
        srv.run(warm)
        torch.cuda.synchronize()
        cold[where] = time.perf_counter() - t
        print(f"cold cache ({where} preprocessing): {args.pool} graphs served in {cold[where]:.2f}s, "
              f"{srv.cache.bytes / 2**30:.2f} GiB resident", file=sys.stderr)
        if where == args.preprocess:
            server = srv
            server.prime()   # every bucket's HIP graph captured before the timed runs
        else:
            del srv
    if args.saturated_only:
        reqs = S.make_requests(args.requests, None, args.pool, seed=2)
        torch.cuda.synchronize()
        server.run(reqs)
        torch.cuda.synchronize()
        print(json.dumps({"phase": "saturated", "batched": S.latency_stats(reqs), "batches": server.batches}),
              flush=True)
        return
    print(json.dumps({"phase": "cold cache", "graphs": args.pool, "host_preprocess_s": round(cold["host"], 3),
                      "device_preprocess_s": round(cold["device"], 3),
                      "speedup": round(cold["host"] / cold["device"], 2)}), flush=True)
    cold_s = cold[args.preprocess]
    plain = S.PlainServer(dev, models, pool)
    for k in range(args.pool):
        plain._get(k)

    def run(srv, n, rate, seed):
        reqs = S.make_requests(n, rate, args.pool, seed=seed)
        torch.cuda.synchronize()
        srv.run(reqs)
        return S.latency_stats(reqs)

    sat_plain = run(plain, args.plain_requests, None, 1)
    sat_ours = run(server, args.requests, None, 2)
    cap_plain = sat_plain["req_per_s"]
    print(json.dumps({"phase": "saturated", "plain_fp32": sat_plain, "batched": sat_ours,
                      "capacity_ratio": sat_ours["req_per_s"] / cap_plain}), flush=True)
    bound = run(plain, args.plain_requests, 0.7 * cap_plain, 3)
    p99_bound = bound["p99_ms"]
    print(json.dumps({"phase": "plain at 70% load", "rate": 0.7 * cap_plain, **bound}), flush=True)
    best = None
    for mult in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64):
        rate = mult * cap_plain
        if rate > 1.2 * sat_ours["req_per_s"]:
            break
        st = run(server, args.requests, rate, 10 + mult)
        ok = st["p99_ms"] <= p99_bound
        print(json.dumps({"phase": "batched", "rate_multiple": mult, "rate": rate, "within_p99_bound": ok, **st}),
              flush=True)
        if ok:
            best = (mult, st)
    print(json.dumps({"summary": True, "plain_capacity_req_s": cap_plain, "p99_bound_ms": p99_bound,
                      "batched_capacity_req_s": sat_ours["req_per_s"],
                      "max_rate_multiple_within_bound": best[0] if best else 0,
                      "req_s_at_bound": best[1]["req_per_s"] if best else None,
                      "batches": server.batches, "cache_hits": server.cache.hits,
                      "cache_misses": server.cache.misses, "cold_warmup_s": cold_s,
                      "preprocess": args.preprocess, "hip_graphs": not args.no_graph,
                      "part_size": args.part_size}), flush=True)


if __name__ == "__main__":
    main()
