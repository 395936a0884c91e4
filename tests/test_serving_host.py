"""CPU: host-side serving logic -- LRU cache with byte budget and pinning,
request generation, latency statistics, graph pool, and the admission
estimate bounding the real footprint."""
import numpy as np
import pytest
import torch

from att_amd import serve as S


class Fake:
    def __init__(self, key, nbytes):
        self.key, self.nbytes, self.pins = key, nbytes, 0


def test_cache_lru_budget_and_pins():
    c = S.GraphCache(budget_bytes=100)
    for k in range(3):
        assert c.make_room(30, reserved=0)
        c.put(Fake(k, 30))
    assert c.bytes == 90
    assert c.get(0) is not None          # 0 becomes most recent; LRU order 1, 2, 0
    c.entries[1].pins = 1                # in-flight: not evictable
    assert c.make_room(30, reserved=0)   # evicts 2 (LRU among unpinned)
    assert 1 in c.entries and 2 not in c.entries and 0 in c.entries
    assert c.evictions == 1
    assert not c.make_room(90, reserved=20)   # larger than the budget itself
    c.entries[1].pins = 0
    c.entries[0].pins = 1
    assert c.make_room(60, reserved=0)   # must evict 1 but never 0
    assert 0 in c.entries and 1 not in c.entries


def test_requests_and_stats():
    r = S.make_requests(2000, 500.0, 64, seed=3)
    arr = np.array([x.arrival for x in r])
    assert np.all(np.diff(arr) >= 0)
    rate = len(r) / arr[-1]
    assert 400 < rate < 600                               # Poisson at 500/s
    kinds = np.array([x.kind == "gcn" for x in r])
    assert 0.45 < kinds.mean() < 0.55                     # 50/50 GCN / GAT
    assert all(0 <= x.key < 64 for x in r)
    assert all(x.arrival == 0.0 for x in S.make_requests(10, None, 4))
    for i, x in enumerate(r):
        x.t_done = x.arrival + (0.001 if i % 100 else 0.01)
    st = S.latency_stats(r)
    assert abs(st["p50_ms"] - 1.0) < 1e-6 and st["p99_ms"] > 1.0


def test_graph_pool_distribution_and_estimate():
    pool = S.GraphPool(300, seed=1)
    assert pool.n.min() >= 100 and pool.n.max() <= 200_000
    assert 1500 < np.median(pool.n) < 2600                # lognormal median 2000
    assert 4 <= pool.deg.min() and pool.deg.max() <= 32
    for k in range(5):
        csr, feats = pool(k)
        assert feats.shape == (csr.n, S.F_IN)
        host = S.prepare_host(csr, feats)
        real = sum(t.numel() * t.element_size() for t in
                   list(host["gcn"].values()) + list(host["gat"].values()) + [host["X"]])
        assert S.estimate_bytes(csr) >= real              # admission never under-counts
        assert host["n_pad"] % S.WINDOW == 0
        assert torch.equal(host["gat"]["rowptr"][-1], torch.tensor(csr.nnz, dtype=torch.int32))
