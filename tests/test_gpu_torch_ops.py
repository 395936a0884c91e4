"""torch.ops.attgnn on the GPU: each registered op returns bit for bit what the
ctypes wrapper of the same C-ABI entry point returns, its fake implementation
agrees with the real output (torch.library.opcheck), and a chain of the ops
compiles with torch.compile (fullgraph, no graph breaks) to the eager result."""
import numpy as np
import pytest
import torch

import att_amd.torch_ops  # noqa: F401  (registers torch.ops.attgnn)
from att_amd import graphs as G
from att_amd import ops
from att_amd.layers import GATGraph, GCNGraph

pytestmark = pytest.mark.gpu
A = torch.ops.attgnn


@pytest.fixture(scope="module")
def gcn(cuda):
    g = G.chung_lu(700, 8, 2.1, seed=3)
    return GCNGraph(g, cuda, tile_dtype=torch.float16, compact="auto"), g.n


@pytest.fixture(scope="module")
def gat(cuda):
    g = G.chung_lu(600, 7, 2.1, seed=4)
    return GATGraph(g, cuda), g.n


def _same(a, b):
    assert a.dtype == b.dtype and a.shape == b.shape
    assert torch.equal(a, b)


def test_spmm_ops_match_wrappers(cuda, gcn):
    graph, n = gcn
    rp, col, val, win, at = graph.fwd["arrays"]
    masks, base, values = graph.fwd["compact"]
    torch.manual_seed(0)
    for dim in (32, 128):
        X = torch.randn(n, dim, device=cuda).half()
        for out_dtype in (torch.float32, torch.float16):
            _same(A.spmm_block(rp, col, val, win, at, X, n, 16, 32, "auto", out_dtype, True),
                  ops.spmm_block(rp, col, val, win, at, X, n, 16, 32, out_dtype=out_dtype, rows_covered=True))
            _same(A.spmm_compact(rp, col, masks, base, values, win, at, X, n, 16, "auto", out_dtype, True),
                  ops.spmm_compact(rp, col, masks, base, values, win, at, X, n, 16, out_dtype=out_dtype,
                                   rows_covered=True))


def test_attention_ops_match_wrappers(cuda, gat):
    graph, n = gat
    heads, dim, nnz = 4, 16, graph.num_edges
    rp, col, emap, win, at = graph.fwd
    base, eids, pos = graph.edges
    torch.manual_seed(1)
    X = torch.randn(n, heads, dim, device=cuda).bfloat16()
    a0 = torch.randn(heads, dim, device=cuda)
    a1 = torch.randn(heads, dim, device=cuda)
    s0, s1 = A.node_scores(X, a0, "mhd"), A.node_scores(X, a1, "mhd")
    _same(s0, ops.node_scores(X, a0, "mhd"))
    _same(A.sddmm_edges(rp, col, base, eids, pos, win, X, X, nnz, 16, "mhd", torch.float32, True),
          ops.sddmm_edges(rp, col, base, eids, pos, win, X, X, nnz, 16, layout="mhd", out_dtype=torch.float32,
                          covered=True))
    p, lse = A.edge_softmax_fwd(graph.rowptr, graph.col, s0, s1, 0.2, 0.25, 11, torch.bfloat16)
    p2, lse2 = ops.edge_softmax_fwd(graph.rowptr, graph.col, s0, s1, 0.2, 0.25, 11, p_dtype=torch.bfloat16)
    _same(p, p2)
    _same(lse, lse2)
    for out_layout in ("mhd", "hmd"):
        _same(A.spmm_edge_map(rp, col, emap, p, win, at, X, out_layout, 16, 32, -1, "auto", "mhd", torch.float32,
                              True),
              ops.spmm_edge_map(rp, col, emap, p, win, at, X, out_layout, 16, 32, x_layout="mhd",
                                out_dtype=torch.float32, rows_covered=True))
    dp = torch.randn(heads, nnz, device=cuda)
    dz, ds0 = A.edge_softmax_bwd(graph.rowptr, graph.col, s0, s1, lse, dp, 0.2, 0.25, 11)
    dz2, ds02 = ops.edge_softmax_bwd(graph.rowptr, graph.col, s0, s1, lse, dp, 0.2, 0.25, 11)
    _same(dz, dz2)
    _same(ds0, ds02)
    _same(A.gather_rowsum(graph.t_rowptr, graph.t_perm, dz, n), ops.gather_rowsum(graph.t_rowptr, graph.t_perm, dz, n))
    s0n, s1n = A.node_scores(X, a0, "mhd", "nh"), A.node_scores(X, a1, "mhd", "nh")
    _same(s0n.t(), s0)
    f0, f1 = A.edge_softmax_bwd_sums(graph.rowptr, graph.col, s0n, s1n, lse, dp, 0.2, 0.25, 11, graph.t_rowptr,
                                     graph.t_perm, "nh")
    _same(f0, ds0)
    _same(f1, ops.gather_rowsum(graph.t_rowptr, graph.t_perm, dz, n))
    ds1 = ops.gather_rowsum(graph.t_rowptr, graph.t_perm, dz, n)
    gbase = torch.randn(n, heads, dim, device=cuda)
    got = A.attention_node_bwd(X, X, a0, a1, ds0, ds1, gbase, True)
    ref = ops.attention_node_bwd(X, X, a0, a1, ds0, ds1, base=gbase, combine=True)
    assert got[0].numel() == 0 and ref[0] is None
    for a, b in zip(got[1:], ref[1:]):
        _same(a, b)
    Xt = torch.randn(n, heads, dim, device=cuda).bfloat16()
    got = A.attention_node_bwd(X, Xt, a0, a1, ds0, ds1, None, False)
    ref = ops.attention_node_bwd(X, Xt, a0, a1, ds0, ds1, combine=False)
    for a, b in zip(got, ref):
        _same(a, b)


def test_opcheck_fake_implementations(cuda, gcn, gat):
    """The fake implementations give the real outputs' shapes, dtypes and strides."""
    graph, n = gcn
    rp, col, val, win, at = graph.fwd["arrays"]
    X = torch.randn(n, 64, device=cuda).half()
    util = ("test_schema", "test_faketensor")
    torch.library.opcheck(A.spmm_block, (rp, col, val, win, at, X, n, 16, 32, "auto", torch.float32, True),
                          test_utils=util)
    ggraph, gn = gat
    heads, dim = 2, 32
    Xg = torch.randn(gn, heads, dim, device=cuda).half()
    a0 = torch.randn(heads, dim, device=cuda)
    s0 = ops.node_scores(Xg, a0, "mhd")
    torch.library.opcheck(A.node_scores, (Xg, a0, "mhd"), test_utils=util)
    torch.library.opcheck(A.edge_softmax_fwd, (ggraph.rowptr, ggraph.col, s0, s0, 0.2, 0.0, 0, torch.float16),
                          test_utils=util)
    p, _ = ops.edge_softmax_fwd(ggraph.rowptr, ggraph.col, s0, s0, 0.2, p_dtype=torch.float16)
    rp2, col2, emap, win2, at2 = ggraph.fwd
    torch.library.opcheck(A.spmm_edge_map, (rp2, col2, emap, p, win2, at2, Xg, "mhd", 16, 32, -1, "auto", "mhd",
                                            torch.float32, True), test_utils=util)
    rpt, colt, _, wint, att = ggraph.bwd
    torch.library.opcheck(A.spmm_edge_map_compact_dot, (rpt, colt, *ggraph.cmap_bwd, p, wint, att, Xg, Xg, "mhd", 16,
                                                        -1, "mhd", torch.float32, True), test_utils=util)


def test_compact_edge_map_ops_match_wrappers(cuda, gat):
    graph, n = gat
    heads, dim = 4, 64
    torch.manual_seed(5)
    X = torch.randn(n, heads, dim, device=cuda).bfloat16()
    Xt = torch.randn(n, heads, dim, device=cuda).bfloat16()
    p = torch.rand(heads, graph.num_edges, device=cuda).bfloat16()
    rp, col, _, win, at = graph.fwd
    _same(A.spmm_edge_map_compact(rp, col, *graph.cmap_fwd, p, win, at, X, "mhd", 16, -1, "auto", "mhd",
                                  torch.float32, True),
          ops.spmm_edge_map_compact(rp, col, *graph.cmap_fwd, p, win, at, X, "mhd", 16, x_layout="mhd",
                                    out_dtype=torch.float32, rows_covered=True))
    rp, col, _, win, at = graph.bwd
    y, dot = A.spmm_edge_map_compact_dot(rp, col, *graph.cmap_bwd, p, win, at, X, Xt, "mhd", 16, -1, "mhd",
                                         torch.float32, True)
    y2, dot2 = graph.aggregate_sddmm(p, X, Xt)
    _same(y, y2)
    _same(dot, dot2)


def test_compiled_attention_forward(cuda, gat):
    """node scores -> edge softmax -> aggregation, traced as one graph."""
    graph, n = gat
    heads, dim = 4, 32
    rp, col, emap, win, at = graph.fwd

    def layer(X, a0, a1):
        s0 = A.node_scores(X, a0, "mhd")
        s1 = A.node_scores(X, a1, "mhd")
        p, _ = A.edge_softmax_fwd(graph.rowptr, graph.col, s0, s1, 0.2, 0.0, 0, X.dtype)
        return A.spmm_edge_map(rp, col, emap, p, win, at, X, "mhd", 16, 32, -1, "auto", "mhd", torch.float32, True)

    torch.manual_seed(2)
    X = torch.randn(n, heads, dim, device=cuda).half()
    a0 = torch.randn(heads, dim, device=cuda)
    a1 = torch.randn(heads, dim, device=cuda)
    eager = layer(X, a0, a1)
    compiled = torch.compile(layer, backend="aot_eager", fullgraph=True)
    out = compiled(X, a0, a1)
    assert torch.equal(out, eager)
    # against FP64 message passing on the host
    rows = torch.from_numpy(np.repeat(np.arange(n), np.diff(graph.rowptr.cpu().numpy())).astype(np.int64))
    cols = graph.col.cpu().long()
    Xd = X.double().cpu()
    s0 = (Xd * a0.double().cpu()).sum(-1)
    s1 = (Xd * a1.double().cpu()).sum(-1)
    e = torch.nn.functional.leaky_relu(s0[rows] + s1[cols], 0.2)
    m = torch.full((n, heads), -np.inf, dtype=torch.float64).scatter_reduce(
        0, rows[:, None].expand(-1, heads), e, "amax", include_self=True)
    ex = torch.exp(e - m[rows])
    p = ex / torch.zeros(n, heads, dtype=torch.float64).index_add(0, rows, ex)[rows]
    ref = torch.zeros(n, heads, dim, dtype=torch.float64).index_add(0, rows, p[..., None] * Xd[cols])
    err = ((out.double().cpu() - ref).norm() / ref.norm()).item()
    assert err < 2e-3, err
