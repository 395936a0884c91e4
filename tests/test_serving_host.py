"""CPU: host-side serving logic -- LRU cache with byte budget and pinning,
request generation, latency statistics, graph pool, the admission estimate
bounding the real footprint, and the pack tables: applied to CPU buffers they
produce the window-block structure of the block-diagonal batch graph."""
import collections
import ctypes
import types

import numpy as np
import pytest
import torch

from att_amd import _lib
from att_amd import graphs as G
from att_amd import ops
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


class FakeEvent:
    """Stands for a CUDA event: in flight until synchronised."""
    def __init__(self, log, name):
        self.log, self.name, self.done = log, name, False

    def synchronize(self):
        self.log.append(self.name)
        self.done = True

    def query(self):
        return self.done


def test_pinned_ring_waits_for_every_overlapping_piece():
    """The staging ring hands out pieces round-robin; a new piece must wait
    for every in-flight piece it overlaps, also when pieces of two earlier
    passes over the ring are still in flight (a wrap that left an old piece
    near the end while a newer one sits at the start)."""
    ring = S.PinnedRing(nbytes=100 * 256, pin=False)
    log = []

    def take(name, units):
        start = ring._take(units * 256)
        ring.fences.append((start, start + units * 256, FakeEvent(log, name)))
        return start // 256

    # first pass: ten pieces of 10 units, 0..95 (the last of 5)
    for k in range(9):
        assert take(f"a{k}", 10) == 10 * k
    assert take("z", 5) == 90
    # wrap: a piece of 30 units at 0 waits for a0, a1, a2
    assert take("c", 30) == 0
    assert log == ["a0", "a1", "a2"]
    # a piece of 80 units does not fit after 30: wrap again to 0; it overlaps
    # c (second pass) and a3..a7 (first pass), not z at 90..95
    log.clear()
    assert take("d", 80) == 0
    assert sorted(log) == sorted(["c", "a3", "a4", "a5", "a6", "a7"])
    assert "z" not in log
    # pieces larger than the ring are refused
    assert ring._take(101 * 256) is None


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


def _apply_tables(b):
    """A NumPy stand-in for attgnn_pack_segments over CPU buffers: chunk c
    copies its part of segment chunk_seg[c] (kinds as the C ABI documents)."""
    host = b.table_host.numpy()
    segs = host[: b.map_at].view(S.SEG_DTYPE)
    cmap = host[b.map_at:].view(np.int32)
    arr = lambda addr, ct, n: np.ctypeslib.as_array((ct * int(n)).from_address(int(addr)))
    for c in np.nonzero(cmap >= 0)[0]:
        seg = segs[cmap[c]]
        unit = {S.KIND_BYTES: 1, S.KIND_I64: 8}.get(int(seg["kind"]), 4)
        per = S.CHUNK_BYTES // unit
        i0 = (int(c) - int(seg["first"])) * per
        n = min(per, int(seg["count"]) - i0)
        assert 0 <= i0 and n > 0
        src, dst, off = int(seg["src"]), int(seg["dst"]), int(seg["offset"])
        if seg["kind"] == S.KIND_BYTES:
            ctypes.memmove(dst + i0, src + i0, n)
            continue
        ct = ctypes.c_int64 if seg["kind"] == S.KIND_I64 else ctypes.c_int32
        v = arr(src + i0 * unit, ct, n).copy() if src else np.zeros(n, np.int64 if unit == 8 else np.int32)
        out = v + off
        if seg["kind"] == S.KIND_IDX:
            out = np.where(v < 0, v, out)
        arr(dst + i0 * unit, ct, n)[:] = out


@pytest.mark.parametrize("kind", ["gcn", "gat"])
def test_pack_tables_build_block_diagonal_structure(kind):
    pool = S.GraphPool(24, seed=3, n_max=6000)
    graphs = []
    for key in (0, 3, 5, 7, 11):
        csr, feats = pool(key)
        h = S.prepare_host(csr, feats)
        cg = object.__new__(S.CachedGraph)
        cg.key, cg.n, cg.n_pad, cg.nnz = key, h["n"], h["n_pad"], h["nnz"]
        cg.gcn, cg.gat, cg.X = h["gcn"], h["gat"], h["X"]
        cg.describe()
        graphs.append(cg)
    b = object.__new__(S.Bucket)
    b.kind, b.rows_cap = kind, 1 << 16
    b.allocate("cpu")
    assert b.fits(graphs)
    for batch in (graphs, graphs[1:3]):     # a second, smaller batch rewrites the tails
        offsets = b.build_table(batch)
        _apply_tables(b)
        # the block-diagonal batch graph, built directly
        pads = [g.n_pad for g in batch]
        assert offsets == list(np.cumsum([0] + pads[:-1]))
        ip, idx, vals, node, edge = [0], [], [], 0, 0
        for g in batch:
            csr = pool(g.key)[0]
            p = csr.padded(S.WINDOW)
            d = G.gcn_norm_degree(p)
            d[~np.isfinite(d)] = 0.0
            vals.append(np.asarray(G.gcn_norm_values(p, d), np.float32))
            ip += list(p.indptr[1:] + edge)
            idx.append(p.indices + node)
            node += p.n
            edge += p.nnz
        rowptr, col = np.array(ip, np.int32), np.concatenate(idx).astype(np.int32)
        N, E = node, edge
        if kind == "gcn":
            e = ops.blocks_build(rowptr, col, "edge_f32", np.concatenate(vals), torch.bfloat16, S.WINDOW, S.WIDE,
                                 S.SERVE_PART_SIZE, emit_empty=True)
        else:
            e = ops.blocks_build(rowptr, col, "map", None, torch.int32, S.WINDOW, S.WIDE, S.SERVE_PART_SIZE,
                                 emit_empty=True, compact=True)
        P, C = e["window"].numel(), e["col"].numel()
        assert torch.equal(b.rp[: 2 * P + 1], e["rp"]) and bool((b.rp[2 * P + 1:] == C).all())
        assert torch.equal(b.col[:C], e["col"])
        assert torch.equal(b.win[:P], e["window"]) and bool((b.win[P:] == -1).all())
        assert torch.equal(b.atomic[:P], e["atomic"]) and bool((b.atomic[P:] == 0).all())
        X = torch.cat([g.X for g in batch])
        assert torch.equal(b.X[:N], X)
        if kind == "gcn":
            assert torch.equal(b.val[: C * S.WINDOW].view(torch.int16), e["val"].view(torch.int16))
            continue
        assert torch.equal(b.masks[:C], e["masks"][:C])
        # every block's packed ids (each graph's run of values starts 8-aligned in the batch)
        nb = C // S.WIDE
        ebase = e["block_base"].numpy()
        per = np.diff(ebase)
        got_ids = np.concatenate([b.values.numpy()[o: o + k] for o, k in zip(b.bbase.numpy()[:nb], per)])
        assert np.array_equal(got_ids, e["values"].numpy()[: ebase[nb]])
        assert torch.equal(b.rowptr[: N + 1], torch.from_numpy(rowptr)) and bool((b.rowptr[N + 1:] == E).all())
        assert torch.equal(b.ccol[:E], torch.from_numpy(col))


def test_serve_stage_host_matches_numpy_and_torch():
    """attgnn_serve_stage_host (host work only): the CSR padded to whole
    windows (empty padding rows), d_i = 1/sqrt(row count) in float32 (0 for
    empty rows, as graphs.gcn_norm_degree with the non-finite entries
    zeroed), features rounded to BF16 exactly as torch casts (ties to even;
    a NaN stays a NaN, whose bits torch's casts do not fix) with zero padding
    rows, at the offsets
    attgnn_serve_stage_layout reports; bad arguments are refused."""
    E_ARG = -1   # ATTGNN_E_ARG
    L = _lib.lib()
    g0 = G.chung_lu(37, 6, 2.1, seed=5)
    # row 3 emptied (its degree must come out 0, not inf)
    keep = np.ones(g0.nnz, bool)
    keep[g0.indptr[3]:g0.indptr[4]] = False
    counts = np.diff(g0.indptr)
    counts[3] = 0
    g = G.CSR(np.concatenate([[0], np.cumsum(counts)]).astype(np.int64), g0.indices[keep].astype(np.int32), g0.n)
    indptr = g.indptr
    rng = np.random.default_rng(4)
    f = rng.standard_normal((g.n, 128)).astype(np.float32) * 3.0
    f[0, :6] = [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -0.0, np.float32(3.3e38), np.nan, -np.inf]
    for window in (16, 8):
        off = (ctypes.c_int64 * 5)()
        total = L.attgnn_serve_stage_layout(g.n, g.nnz, 128, window, off)
        n_pad = off[4]
        assert n_pad == -(-g.n // window) * window and all(o % 256 == 0 for o in off[:4]) and total >= off[3]
        out = np.full(total, 0xAB, np.uint8)
        rc = L.attgnn_serve_stage_host(indptr.ctypes.data, g.indices.ctypes.data, g.n, g.nnz, f.ctypes.data, 128,
                                       window, out.ctypes.data, total)
        assert rc == 0
        p = g.padded(window)
        assert np.array_equal(out[off[0]:off[0] + 4 * (n_pad + 1)].view(np.int32), p.rowptr32())
        assert np.array_equal(out[off[1]:off[1] + 4 * g.nnz].view(np.int32), g.indices)
        d = G.gcn_norm_degree(p)
        d[~np.isfinite(d)] = 0.0
        assert np.array_equal(out[off[2]:off[2] + 4 * n_pad].view(np.float32).view(np.uint32), d.view(np.uint32))
        x = out[off[3]:off[3] + 2 * n_pad * 128].view(np.uint16).reshape(n_pad, 128)
        want = torch.from_numpy(f).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
        nan = np.isnan(f)
        assert np.array_equal(x[:g.n][~nan], want[~nan])
        assert np.all((x[:g.n][nan] & 0x7F80) == 0x7F80) and np.all((x[:g.n][nan] & 0x7F) != 0)   # NaN stays NaN
        assert x[0, 5] == 0xFF80 and x[0, 3] == 0x7F78 and x[0, 1] == 0x3F82 and x[0, 0] == 0x3F80
        assert not x[g.n:].any()
        assert out[off[2] + 4 * 3:off[2] + 4 * 4].view(np.float32)[0] == 0.0
    # features NULL: the rows the caller wrote stay, the padding rows are zeroed
    out = np.full(total, 0x5A, np.uint8)
    assert L.attgnn_serve_stage_host(indptr.ctypes.data, g.indices.ctypes.data, g.n, g.nnz, None, 128, 8,
                                     out.ctypes.data, total) == 0
    x = out[off[3]:off[3] + 2 * off[4] * 128].view(np.uint16).reshape(off[4], 128)
    assert np.all(x[:g.n] == 0x5A5A) and not x[g.n:].any()
    # refused: buffer too small, indptr not ending at nnz, a negative count
    small = np.zeros(16, np.uint8)
    assert L.attgnn_serve_stage_host(indptr.ctypes.data, g.indices.ctypes.data, g.n, g.nnz, f.ctypes.data, 128, 16,
                                     small.ctypes.data, 16) == E_ARG
    out = np.zeros(total, np.uint8)
    assert L.attgnn_serve_stage_host(indptr.ctypes.data, g.indices.ctypes.data, g.n, g.nnz - 1, f.ctypes.data, 128,
                                     8, out.ctypes.data, total) == E_ARG
    assert L.attgnn_serve_stage_layout(-1, 0, 128, 16, off) < 0


def test_prefetch_selection():
    """Server._prefetch starts preparations for the first queued requests
    whose graphs are neither cached nor already in progress, keeps at most
    ``prefetch`` in progress, looks at no more than 4x that many requests and
    stops at the first graph that does not fit the budget."""
    class Fake:
        def __init__(self, prefetch, cached=(), room=99):
            self.prefetch, self.preparing, self.started = prefetch, collections.OrderedDict(), []
            self.cache = types.SimpleNamespace(entries=dict.fromkeys(cached))
            self.room = room

        def _start(self, key):
            if self.room == 0:
                return False
            self.room -= 1
            self.preparing[key] = (None, 1)
            self.started.append(key)
            return True

    q = collections.deque(S.Request(i, 0.0, "gcn", k) for i, k in enumerate([5, 5, 1, 2, 3, 4, 6, 7]))
    f = Fake(3, cached=[1])
    S.Server._prefetch(f, q)
    assert f.started == [5, 2, 3]          # 5 once, 1 cached
    S.Server._prefetch(f, q)
    assert f.started == [5, 2, 3]          # already 3 in progress
    f = Fake(2, room=1)
    S.Server._prefetch(f, q)
    assert f.started == [5]                # the next one does not fit: stop
    f = Fake(1, cached=[5, 1, 2, 3, 4])
    S.Server._prefetch(f, q)
    assert f.started == []                 # only the first 4 requests are looked at
