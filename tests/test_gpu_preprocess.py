"""Device-side window-block preprocessing (attgnn_blocks_plan_device / _fill_device
and the compact count / fill) is bit-identical to the host build, which the
CPU suite pins against the reference's own example.cpp and golden fixtures."""
import numpy as np
import pytest
import torch

from att_amd import ops
from att_amd._lib import AttGnnError
from gpu_util import TOY_COL, TOY_ROW, rand_csr

pytestmark = pytest.mark.gpu

GRID = [(w, wd, ps) for w in (8, 16) for wd in (4, 8, 16, 32) for ps in (1, 2, 32)]
KEYS = ("rp", "col", "val", "window", "atomic")


def _bits(t):
    t = t.detach().cpu()
    if t.dtype in (torch.float16, torch.bfloat16):
        t = t.view(torch.int16)
    elif t.dtype == torch.float32:
        t = t.view(torch.int32)
    return t.numpy()


def _same(host, dev, keys, where):
    for k in keys:
        a, b = _bits(host[k]), _bits(dev[k])
        assert a.shape == b.shape, (where, k, a.shape, b.shape)
        assert np.array_equal(a, b), (where, k, int(np.flatnonzero(a != b)[0]))
    assert host["n_unique"] == dev["n_unique"], where


def _both(rp, col, src, vals, out_dtype, w, wd, ps, emit_empty, trans=False, compact=False, cuda=None):
    host = ops.blocks_build(rp, col, src, vals, out_dtype, w, wd, ps, emit_empty=emit_empty, want_trans=trans,
                            compact=compact)
    rp_d = torch.from_numpy(np.asarray(rp, np.int32)).to(cuda)
    col_d = torch.from_numpy(np.asarray(col, np.int32)).to(cuda)
    vals_d = None
    if src != "map":
        want = torch.float16 if src == "deg_f16" else torch.float32
        vals_d = torch.as_tensor(np.asarray(vals)).to(want).to(cuda)
    dev = ops.blocks_build_device(rp_d, col_d, src, vals_d, out_dtype, w, wd, ps, emit_empty=emit_empty,
                                  want_trans=trans, compact=compact)
    return host, dev


SOURCES = [("deg_f16", torch.float16, False), ("deg_f32", torch.float32, True),
           ("edge_f32", torch.bfloat16, True), ("map", torch.int32, False)]


@pytest.mark.parametrize("graph", ["toy", "random", "hubs"])
def test_device_build_matches_host_grid(graph, cuda):
    rng = np.random.default_rng(11)
    if graph == "toy":
        rp, col = TOY_ROW, TOY_COL
    elif graph == "random":
        rp, col = rand_csr(rng, 333, 24)
    else:
        rp, col = rand_csr(rng, 700, 12, hub_rows=5)
    n = rp.size - 1
    d16 = (rng.random(n) + 0.05).astype(np.float16)
    d32 = rng.random(n).astype(np.float32)
    ev = rng.standard_normal(col.size).astype(np.float32)
    vals = {"deg_f16": d16, "deg_f32": d32, "edge_f32": ev, "map": None}
    for w, wd, ps in GRID:
        for src, dt, ee in SOURCES:
            host, dev = _both(rp, col, src, vals[src], dt, w, wd, ps, ee, trans=(src == "map"), cuda=cuda)
            keys = KEYS + (("map_t",) if src == "map" else ())
            _same(host, dev, keys, (graph, src, w, wd, ps))


@pytest.mark.parametrize("dt", [torch.float32, torch.float16, torch.bfloat16])
def test_edge_values_every_type_and_empty_rule(dt, cuda):
    rng = np.random.default_rng(5)
    rp, col = rand_csr(rng, 500, 40, hub_rows=2)
    ev = rng.standard_normal(col.size).astype(np.float32)
    ev[::7] = 0.0   # zero values stay zero tiles (and are not compact entries)
    for ee in (False, True):
        host, dev = _both(rp, col, "edge_f32", ev, dt, 16, 32, 3, ee, compact=True, cuda=cuda)
        _same(host, dev, KEYS + ("masks", "block_base", "values", "positions"), (dt, ee))


def test_large_windows_bitmap_path(cuda):
    """Windows with more than 4,096 entries spanning more than 131,072 columns
    (several bitmap passes), next to small windows (sorted in LDS)."""
    rng = np.random.default_rng(7)
    n = 300_000
    deg = rng.integers(0, 4, n)
    deg[:40] = rng.integers(400, 3000, 40)        # windows 0-2 at W=16: > 4,096 entries
    deg[1000:1016] = 20                           # a medium window
    rows = [np.unique(rng.integers(0, n, int(k))) for k in deg]
    rp = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    ev = rng.standard_normal(col.size).astype(np.float32)
    for w in (8, 16):
        host, dev = _both(rp, col, "edge_f32", ev, torch.float16, w, 32, 4, True, compact=True, cuda=cuda)
        _same(host, dev, KEYS + ("masks", "block_base", "values", "positions"), ("bitmap", w))
        host, dev = _both(rp, col, "map", None, torch.int32, w, 8, 32, False, trans=True, cuda=cuda)
        _same(host, dev, KEYS + ("map_t",), ("bitmap-map", w))


def test_unsorted_rows_and_duplicates(cuda):
    """Unsorted rows and repeated (row, column) entries: the last entry of a
    repeated pair owns the tile slot, as in the host build."""
    rng = np.random.default_rng(9)
    n = 256
    rows = []
    for i in range(n):
        r = rng.integers(0, n, int(rng.integers(0, 20)))
        if i % 3 == 0:
            r = np.sort(r)                        # sorted with duplicates
        rows.append(r)
    rows[17] = rng.integers(0, n, 5000)            # a long unsorted row with repeats
    rp = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    ev = rng.standard_normal(col.size).astype(np.float32)
    for w, wd, ps in [(8, 8, 2), (16, 32, 1), (16, 4, 32)]:
        host, dev = _both(rp, col, "edge_f32", ev, torch.float32, w, wd, ps, True, cuda=cuda)
        _same(host, dev, KEYS, ("dup-edge", w, wd, ps))
        host, dev = _both(rp, col, "map", None, torch.int32, w, wd, ps, False, trans=True, cuda=cuda)
        _same(host, dev, KEYS + ("map_t",), ("dup-map", w, wd, ps))


def test_map_compact_matches_host(cuda):
    rng = np.random.default_rng(13)
    rp, col = rand_csr(rng, 900, 30, hub_rows=3)
    host, dev = _both(rp, col, "map", None, torch.int32, 16, 32, 6, False, trans=True, compact=True, cuda=cuda)
    _same(host, dev, KEYS + ("map_t", "masks", "block_base", "values", "positions"), "map-compact")


def test_empty_and_degenerate(cuda):
    # no entries at all, fewer rows than a window, trailing rows dropped
    for rp, col in [(np.zeros(33, np.int32), np.zeros(0, np.int32)),
                    (np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32)),
                    (np.array([0, 2, 2, 3, 3, 3, 3, 3, 3, 5], np.int32), np.array([0, 8, 3, 1, 2], np.int32))]:
        n = rp.size - 1
        for ee in (False, True):
            host, dev = _both(rp, col, "deg_f32", np.ones(n, np.float32), torch.float32, 8, 4, 2, ee, cuda=cuda)
            _same(host, dev, KEYS, ("degenerate", n, ee))


def test_malformed_csr_raises(cuda):
    cases = [([0, 2, 1], [0, 1]),        # row pointers decrease
             ([0, 1, 2], [0, 5]),        # column out of range
             ([0, 1, 3], [0, 1]),        # last row pointer != nnz
             ([1, 1, 2], [0, 1])]        # first row pointer != 0
    for rp, col in cases:
        r = torch.tensor(rp, dtype=torch.int32, device=cuda)
        c = torch.tensor(col, dtype=torch.int32, device=cuda)
        with pytest.raises(AttGnnError):
            ops.blocks_build_device(r, c, "deg_f32", torch.ones(2, device=cuda), torch.float32, 8, 8, 2, True)


def test_reference_module_preprocess_gpu_fs(cuda):
    """ATT_Block_gpu.preprocess_gpu_fs returns what ATT_Block.blockProcess_sddmm_balance_gnn_trans
    returns, on the device, plus the build time."""
    from att_amd import compat
    mods = compat.install()
    rng = np.random.default_rng(17)
    rp, col = rand_csr(rng, 160, 10)
    r, c = torch.from_numpy(rp), torch.from_numpy(col)
    got = mods["ATT_Block_gpu"].preprocess_gpu_fs(r, c, rp.size - 1, col.size, 8, 8, 32)
    want = mods["ATT_Block"].blockProcess_sddmm_balance_gnn_trans(r, c, 8, 8, 32)
    assert len(got) == 7 and isinstance(got[6], float) and got[6] > 0
    for a, b in zip(got[:6], want):
        assert a.device.type == "cuda"
        assert torch.equal(a.cpu(), b)


def _tensors_equal(a, b, where):
    a, b = a.detach().cpu(), b.detach().cpu()
    if a.dtype in (torch.float16, torch.bfloat16):
        a, b = a.view(torch.int16), b.view(torch.int16)
    assert a.shape == b.shape and torch.equal(a, b), where


@pytest.mark.parametrize("symmetric", [True, False])
def test_gcn_graph_device_build_equals_host(symmetric, cuda):
    from att_amd import graphs as G
    from att_amd.layers import GCNGraph
    if symmetric:
        g = G.chung_lu(3000, 9, 2.1, seed=4)
    else:
        rng = np.random.default_rng(4)
        rp, col = rand_csr(rng, 2500, 14, hub_rows=2)
        g = G.CSR(rp.astype(np.int64), col, rp.size - 1)
    host = GCNGraph(g, cuda, tile_dtype=torch.bfloat16, build="host")
    dev = GCNGraph(g, cuda, tile_dtype=torch.bfloat16, build="device")
    assert host.symmetric == dev.symmetric == symmetric
    for which in ("fwd", "bwd"):
        h, d = getattr(host, which), getattr(dev, which)
        for i, (a, b) in enumerate(zip(h["arrays"] + h["compact"], d["arrays"] + d["compact"])):
            _tensors_equal(a, b, (which, i))
    X = torch.randn(g.n, 64, device=cuda).bfloat16()
    _tensors_equal(host.spmm(X), dev.spmm(X), "spmm")
    _tensors_equal(host.spmm(X, transpose=True), dev.spmm(X, transpose=True), "spmm^T")


@pytest.mark.parametrize("symmetric", [True, False])
def test_gat_graph_device_build_equals_host(symmetric, cuda):
    from att_amd import graphs as G
    from att_amd.layers import GATGraph
    if symmetric:
        g = G.directed_chung_lu_symmetrized(5000, 60_000, seed=2)
    else:
        rng = np.random.default_rng(6)
        rp, col = rand_csr(rng, 3000, 20, hub_rows=3)
        g = G.CSR(rp.astype(np.int64), col, rp.size - 1)
    host = GATGraph(g, cuda, build="host")
    dev = GATGraph(g, cuda, build="device")
    for which in ("fwd", "edges", "bwd"):
        for i, (a, b) in enumerate(zip(getattr(host, which), getattr(dev, which))):
            _tensors_equal(a, b, (which, i))
    for name in ("t_rowptr", "t_col", "t_perm"):
        _tensors_equal(getattr(host, name), getattr(dev, name), name)


def test_csr_transpose_device_matches_stable_host_sort(cuda):
    """Rectangular, unsorted rows, repeated entries, empty rows and columns:
    the device transpose equals a stable host sort by (column, row)."""
    rng = np.random.default_rng(8)
    for n_rows, n_cols in [(700, 700), (300, 1000), (1000, 50), (5, 1)]:
        rows = [rng.integers(0, n_cols, int(rng.integers(0, 30))) for _ in range(n_rows)]
        rows[0] = rng.integers(0, n_cols, 3000)                 # a long unsorted row with repeats
        if n_rows > 3:
            rows[3] = np.zeros(0, np.int64)
        rp = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
        col = np.concatenate(rows).astype(np.int32)
        t_rowptr, t_col, t_perm = ops.csr_transpose_device(torch.from_numpy(rp).to(cuda), torch.from_numpy(col).to(cuda),
                                                           n_cols)
        src = np.repeat(np.arange(n_rows), np.diff(rp))
        order = np.lexsort((src, col))                           # stable: ties keep CSR order
        want_rp = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=n_cols))])
        assert np.array_equal(t_rowptr.cpu().numpy(), want_rp)
        assert np.array_equal(t_perm.cpu().numpy(), order)
        assert np.array_equal(t_col.cpu().numpy(), src[order])
    for rp, col in ([0, 2, 1], [0, 1]), ([0, 1, 2], [0, 9]), ([0, 1, 3], [0, 1]):
        with pytest.raises(AttGnnError):
            ops.csr_transpose_device(torch.tensor(rp, dtype=torch.int32, device=cuda),
                                     torch.tensor(col, dtype=torch.int32, device=cuda), 2)


@pytest.mark.parametrize("ring_bytes", [None, 1 << 16, 64 << 20])
def test_serving_prepare_device_equals_host(cuda, ring_bytes):
    """Device preparation (values formed on the device, one shared plan) is
    bitwise the host's; through a pinned staging ring too, including a ring so
    small that it wraps and hands larger arrays their own pinned copy."""
    from att_amd import serve as S
    pool = S.GraphPool(6, seed=5, n_max=5000)
    stream = torch.cuda.Stream(cuda)
    ring = S.PinnedRing(ring_bytes) if ring_bytes else None
    for key in range(6):
        csr, feats = pool(key)
        h = S.prepare_host(csr, feats)
        d = S.prepare_device(csr, feats, cuda, stream, ring=ring)
        stream.synchronize()
        assert (h["n"], h["n_pad"], h["nnz"]) == (d["n"], d["n_pad"], d["nnz"])
        for part in ("gcn", "gat"):
            for k in h[part]:
                _tensors_equal(h[part][k], d[part][k], (key, part, k))
        _tensors_equal(h["X"], d["X"], (key, "X"))
