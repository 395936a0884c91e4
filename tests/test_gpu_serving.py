"""Serving loop: each request's result is the same alone and in a packed
batch, HIP-graph replays match eager runs across different batches, results
agree with the plain FP32 per-graph path, and admission respects the budget."""
import numpy as np
import pytest
import torch

from att_amd import serve as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(cuda):
    pool = S.GraphPool(24, seed=3, n_max=6000)
    models = S.Models(cuda, seed=1)
    return pool, models


def run_server(cuda, models, pool, reqs, use_graph=True, **kw):
    srv = S.Server(cuda, models, pool, capacities=(1 << 14, 1 << 16), use_graph=use_graph, keep_results=True, **kw)
    srv.run(reqs)
    return srv


def test_alone_equals_mix(cuda, setup):
    pool, models = setup
    mix = S.make_requests(48, None, 24, seed=5)
    run_server(cuda, models, pool, mix)
    alone = {}
    for r in mix[:12]:
        single = S.Request(0, 0.0, r.kind, r.key)
        run_server(cuda, models, pool, [single])
        alone[r.rid] = single.result
    exact = 0
    for r in mix[:12]:
        a, b = alone[r.rid].float(), r.result.float()
        # windows never straddle graphs and the sparse kernels are per window,
        # so packing changes nothing there; the dense GEMMs may pick another
        # kernel for another bucket size, hence a BF16-level tolerance
        err = ((a - b).norm() / b.norm()).item()
        assert err < 1e-2, (r.kind, r.key, err)
        exact += int(torch.equal(a, b))
    print(f"{exact}/12 results bitwise equal alone and in the mix")


@pytest.mark.parametrize("kind", ["gcn", "gat"])
def test_packed_sparse_aggregation_bitwise_equals_alone(cuda, setup, kind):
    """The sparse part of a packed batch is each graph's own result bit for
    bit: rows are padded to the 16-row window so no window straddles two
    graphs, every kernel works per window, and the split windows are combined
    in a fixed order.  Same inputs per graph (random features / edge values)
    through the bucket's packed structure and through the graph's own."""
    pool, models = setup
    srv = S.Server(cuda, models, pool, capacities=(1 << 14, 1 << 16), use_graph=False)
    graphs = [srv._admit(S.Request(0, 0.0, kind, k)) for k in (3, 0, 7, 11, 5, 2)]
    torch.cuda.synchronize()
    bucket = srv._pick_bucket(kind, graphs)
    assert bucket is not None
    rows, edges = bucket.pack(graphs)
    gen = torch.Generator().manual_seed(17)
    if kind == "gcn":
        D = S.GCN_HID
        Z = torch.zeros(bucket.rows_cap, D, dtype=torch.bfloat16, device=cuda)
        ins = []
        for g, r0 in zip(graphs, rows):
            z = torch.randn(g.n_pad, D, generator=gen).bfloat16().to(cuda)
            Z[r0:r0 + g.n_pad] = z
            ins.append(z)
        Yp = S.ops.spmm_block(bucket.rp, bucket.col, bucket.val, bucket.win, bucket.atomic, Z, bucket.rows_cap,
                              S.WINDOW, S.WIDE, out_dtype=torch.float32, rows_covered=True)
        for g, r0, z in zip(graphs, rows, ins):
            s = g.gcn
            Ya = S.ops.spmm_block(s["rp"], s["col"], s["val"], s["window"], s["atomic"], z, g.n_pad, S.WINDOW,
                                  S.WIDE, out_dtype=torch.float32, rows_covered=True)
            assert torch.equal(Yp[r0:r0 + g.n_pad], Ya), g.key
    else:
        H, D = S.GAT_HEADS, S.GAT_DIM
        Xp = torch.zeros(bucket.rows_cap, H, D, dtype=torch.bfloat16, device=cuda)
        P = torch.zeros(H, bucket.nnz_cap, dtype=torch.bfloat16, device=cuda)
        ins = []
        for g, r0, e0 in zip(graphs, rows, edges):
            x = torch.randn(g.n_pad, H, D, generator=gen).bfloat16().to(cuda)
            p = torch.rand(H, g.nnz, generator=gen).bfloat16().to(cuda)
            Xp[r0:r0 + g.n_pad] = x
            P[:, e0:e0 + g.nnz] = p
            ins.append((x, p.contiguous()))
        Yp = S.ops.spmm_edge_map_compact(bucket.rp, bucket.col, bucket.masks, bucket.bbase, bucket.values, P,
                                         bucket.win, bucket.atomic, Xp, "mhd", S.WINDOW, m_ori=bucket.rows_cap,
                                         x_layout="mhd", rows_covered=True)
        for g, r0, (x, p) in zip(graphs, rows, ins):
            s = g.gat
            Ya = S.ops.spmm_edge_map_compact(s["rp"], s["col"], s["masks"], s["bbase"], s["values"], p,
                                             s["window"], s["atomic"], x, "mhd", S.WINDOW, m_ori=g.n_pad,
                                             x_layout="mhd", rows_covered=True)
            assert torch.equal(Yp[r0:r0 + g.n_pad], Ya), g.key


def test_graph_replay_matches_eager(cuda, setup):
    pool, models = setup
    for seed in (7, 8, 9):     # several batches through the same captured graphs
        reqs = S.make_requests(40, None, 24, seed=seed)
        eager = S.make_requests(40, None, 24, seed=seed)
        run_server(cuda, models, pool, reqs, use_graph=True)
        run_server(cuda, models, pool, eager, use_graph=False)
        for r, e in zip(reqs, eager):
            assert torch.equal(r.result, e.result), (seed, r.rid)


def test_matches_plain_fp32(cuda, setup):
    pool, models = setup
    reqs = S.make_requests(32, None, 24, seed=11)
    run_server(cuda, models, pool, reqs)
    plain = S.PlainServer(cuda, models, pool)
    ref = S.make_requests(32, None, 24, seed=11)
    plain.run(ref)
    for r, p in zip(reqs, ref):
        got = r.result.float().reshape(p.result.shape)
        err = ((got - p.result).norm() / p.result.norm()).item()
        assert err < 3e-2, (r.kind, r.key, err)


def test_budget_admission(cuda, setup):
    pool, models = setup
    reqs = S.make_requests(30, None, 24, seed=12)
    need = max(S.estimate_bytes(pool(k)[0]) for k in range(24))
    srv = S.Server(cuda, models, pool, capacities=(1 << 14, 1 << 16), keep_results=True, hbm_budget=1 << 62)
    budget = srv.reserved + 3 * need       # room for about three graphs at a time
    srv.cache.budget = budget
    srv.run(reqs)
    assert all(r.t_done is not None for r in reqs)
    assert srv.cache.bytes + srv.reserved <= budget
    assert srv.cache.evictions > 0


def test_prime_captures_every_bucket_and_keeps_results(cuda, setup):
    """Server.prime captures each bucket's HIP graph ahead of traffic (both
    instances of every capacity and model); the served results afterwards
    equal those of eager launches."""
    pool, models = setup
    srv = S.Server(cuda, models, pool, capacities=(1 << 14, 1 << 16), use_graph=True, keep_results=True)
    srv.run([S.Request(i, 0.0, "gcn" if i % 2 else "gat", i) for i in range(24)])
    srv.prime()
    for kind in ("gcn", "gat"):
        for pair in srv.buckets[kind]:
            assert all(b.graph is not None for b in pair), kind
    reqs = S.make_requests(40, None, 24, seed=13)
    srv.run(reqs)
    eager = S.make_requests(40, None, 24, seed=13)
    run_server(cuda, models, pool, eager, use_graph=False)
    for r, e in zip(reqs, eager):
        assert torch.equal(r.result, e.result), r.rid


def test_prefetched_misses_serve_the_same_results(cuda, setup):
    """Cache misses prepared ahead of their admission (uploads and window
    plans queued early, sizes read back behind an event) serve the same
    results as misses prepared at admission: bit for bit when every graph
    fits (the same batches form), within BF16 tolerance under a budget of
    about three graphs (where admission failures change the batches); the
    budget holds and nothing is left in preparation."""
    pool, models = setup
    need = max(S.estimate_bytes(pool(k)[0]) for k in range(24))
    for tight in (False, True):
        outs = []
        for prefetch in (0, 4):
            srv = S.Server(cuda, models, pool, capacities=(1 << 14, 1 << 16), keep_results=True,
                           hbm_budget=1 << 62, prefetch=prefetch)
            if tight:
                srv.cache.budget = srv.reserved + 3 * need
            reqs = S.make_requests(40, None, 24, seed=14)
            srv.run(reqs)
            assert all(r.t_done is not None for r in reqs)
            assert not srv.preparing
            assert srv.cache.bytes + srv.reserved <= srv.cache.budget
            if tight:
                assert srv.cache.evictions > 0
            outs.append(reqs)
        for a, b in zip(*outs):
            if tight:
                err = ((a.result.float() - b.result.float()).norm() / b.result.float().norm()).item()
                assert err < 1e-2, (a.rid, err)
            else:
                assert torch.equal(a.result, b.result), a.rid


def test_plain_server_budget_lru(cuda, setup):
    """The plain per-graph server under a byte budget (the cold baseline):
    it evicts least recently used graphs, stays within the budget and serves
    the same results as with every graph resident (to 1e-5: the plain GAT
    sums with index_add atomics, whose order varies run to run)."""
    pool, models = setup
    warm = S.PlainServer(cuda, models, pool)
    ref = S.make_requests(30, None, 24, seed=15)
    warm.run(ref)
    total = sum(c["nbytes"] for c in warm.cache.values())
    cold = S.PlainServer(cuda, models, pool, budget_bytes=total // 4)
    reqs = S.make_requests(30, None, 24, seed=15)
    cold.run(reqs)
    assert cold.evictions > 0 and cold.misses > len(warm.cache) // 4
    assert cold.bytes <= cold.budget
    assert sum(c["nbytes"] for c in cold.cache.values()) == cold.bytes
    for r, e in zip(reqs, ref):
        err = ((r.result - e.result).norm() / e.result.norm()).item()
        assert err < 1e-5, (r.rid, err)
