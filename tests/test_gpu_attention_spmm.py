"""Attention aggregation with the probabilities rebuilt inside the SpMM
(attgnn_spmm_attention) against the two-step path it replaces: the edge
softmax writes p, the edge-map SpMM reads it.  Both must agree bit for bit --
forward over A's map (row logit s0, column logit s1, row offset lse) and the
backward's A^T aggregation over the transposed map (row logit s1, column
logit s0, column offset lse) -- for every arithmetic kind, head count, slope
and dropout, on symmetric and non-symmetric graphs with hub rows."""
import numpy as np
import pytest
import torch

from att_amd import graphs as G
from att_amd import ops
from att_amd.layers import GATGraph

pytestmark = pytest.mark.gpu


def _nonsymmetric(n, deg, seed):
    g = G.chung_lu(n, deg, 2.1, seed=seed)
    rng = np.random.default_rng(seed)
    keep = rng.random(g.nnz) < 0.7
    rows = g.row_ids()[keep].astype(np.int64)
    cols = g.indices[keep].astype(np.int64)
    keys = np.unique(np.concatenate([rows * n + cols, np.arange(n) * (n + 1)]))
    ip = np.zeros(n + 1, np.int64)
    ip[1:] = np.cumsum(np.bincount(keys // n, minlength=n))
    return G.CSR(ip, (keys % n).astype(np.int32), n)


@pytest.fixture(scope="module", params=["symmetric", "nonsymmetric"])
def graph(request, cuda):
    g = G.chung_lu(900, 10, 2.1, seed=7) if request.param == "symmetric" else _nonsymmetric(900, 10, 8)
    return GATGraph(g, cuda, build="device"), g


@pytest.mark.parametrize("heads,dim", [(1, 64), (2, 32), (4, 16), (8, 8)])
@pytest.mark.parametrize("dtype,mode", [(torch.bfloat16, None), (torch.float16, None), (torch.float32, "f32"),
                                        (torch.float32, "bf16x3")])
def test_fused_equals_softmax_then_edge_map(cuda, graph, heads, dim, dtype, mode):
    gg, g = graph
    n = g.n
    gen = torch.Generator(device="cpu").manual_seed(heads * 10 + dim)
    s0 = torch.randn(n, heads, generator=gen).to(cuda)
    s1 = torch.randn(n, heads, generator=gen).to(cuda)
    X = torch.randn(n, heads, dim, generator=gen).to(dtype).to(cuda)
    dh = torch.randn(n, heads, dim, generator=gen).to(dtype).to(cuda)
    lr, _ = gg.long_row_lists()
    rp, col, emap, win, at = gg.fwd
    rpt, colt, emapt, wint, att = gg.bwd
    for slope, drop in ((0.2, 0.0), (0.2, 0.3), (1.0, 0.1)):
        p, lse = ops.edge_softmax_fwd(gg.rowptr, gg.col, s0, s1, slope, drop, 5, p_dtype=dtype, s_layout="nh",
                                      long_rows=lr)
        _, lse2 = ops.edge_softmax_fwd(gg.rowptr, gg.col, s0, s1, slope, drop, 5, s_layout="nh", long_rows=lr,
                                       lse_only=True)
        assert torch.equal(lse, lse2)
        for out_dtype in (torch.float32, dtype):
            ref = ops.spmm_edge_map(rp, col, emap, p, win, at, X, "mhd", gg.window, gg.wide, m_ori=n, mode=mode,
                                    x_layout="mhd", out_dtype=out_dtype, rows_covered=True)
            got = ops.spmm_attention(rp, col, emap, win, at, X, "mhd", gg.window, gg.wide, s0, s1, row_offset=lse,
                                     slope=slope, drop_p=drop, seed=5, m_ori=n, mode=mode, x_layout="mhd",
                                     out_dtype=out_dtype, rows_covered=True)
            assert torch.equal(ref, got), ("forward", slope, drop, out_dtype)
        ref_t = ops.spmm_edge_map(rpt, colt, emapt, p, wint, att, dh, "mhd", gg.window, gg.wide, m_ori=n, mode=mode,
                                  x_layout="mhd", rows_covered=True)
        got_t = ops.spmm_attention(rpt, colt, emapt, wint, att, dh, "mhd", gg.window, gg.wide, s1, s0,
                                   col_offset=lse, slope=slope, drop_p=drop, seed=5, m_ori=n, mode=mode,
                                   x_layout="mhd", rows_covered=True)
        assert torch.equal(ref_t, got_t), ("transpose", slope, drop)


def test_fused_against_fp64(cuda, graph):
    """The fused aggregation is the softmax-weighted sum (FP64 on the host)."""
    gg, g = graph
    n, heads, dim = g.n, 4, 32
    gen = torch.Generator(device="cpu").manual_seed(3)
    s0 = torch.randn(n, heads, generator=gen).to(cuda)
    s1 = torch.randn(n, heads, generator=gen).to(cuda)
    X = torch.randn(n, heads, dim, generator=gen).to(cuda)
    _, lse = ops.edge_softmax_fwd(gg.rowptr, gg.col, s0, s1, 0.2, 0.0, 0, s_layout="nh", lse_only=True)
    rp, col, emap, win, at = gg.fwd
    y = ops.spmm_attention(rp, col, emap, win, at, X, "mhd", gg.window, gg.wide, s0, s1, row_offset=lse, slope=0.2,
                           m_ori=n, mode="f32", x_layout="mhd", rows_covered=True)
    rows = torch.from_numpy(g.row_ids().astype(np.int64))
    cols = torch.from_numpy(g.indices.astype(np.int64))
    z = s0.double().cpu()[rows] + s1.double().cpu()[cols]
    e = torch.nn.functional.leaky_relu(z, 0.2)
    m = torch.full((n, heads), -np.inf, dtype=torch.float64).scatter_reduce(
        0, rows[:, None].expand(-1, heads), e, "amax", include_self=True)
    ex = torch.exp(e - m[rows])
    p = ex / torch.zeros(n, heads, dtype=torch.float64).index_add(0, rows, ex)[rows]
    ref = torch.zeros(n, heads, dim, dtype=torch.float64).index_add(0, rows, p[..., None] * X.double().cpu()[cols])
    err = ((y.double().cpu() - ref).norm() / ref.norm()).item()
    assert err < 1e-5, err


def test_rejects_bad_terms(cuda, graph):
    gg, g = graph
    rp, col, emap, win, at = gg.fwd
    X = torch.randn(g.n, 2, 16, device=cuda).bfloat16()
    s = torch.randn(g.n, 2, device=cuda)
    with pytest.raises(ValueError):
        ops.spmm_attention(rp, col, emap, win, at, X, "mhd", 16, 32, s, s[:-1].contiguous(), x_layout="mhd")
    with pytest.raises(ValueError):
        ops.spmm_attention(rp, col, emap, win, at, X, "mhd", 16, 32, s, s.double(), x_layout="mhd")


@pytest.mark.parametrize("kind", ["gat", "gatv2"])
def test_layer_fused_equals_stored(cuda, graph, kind):
    """A GAT / reference-GATv2 layer with the probabilities rebuilt in both
    aggregations gives the same output and gradients, bit for bit, as with
    them stored (dropout on, training mode)."""
    from att_amd.layers import GATConv_multi, GATv2Conv_multi
    gg, g = graph
    torch.manual_seed(11)
    layer = (GATConv_multi(48, 32, 0.3, 0.2, 4) if kind == "gat" else GATv2Conv_multi(48, 32, 0.3, 0.2, 4)).to(cuda)
    X = torch.randn(g.n, 48, device=cuda)
    go = torch.randn(g.n, 4, 32, device=cuda).bfloat16()
    results = []
    for fused in (False, True):
        layer.fused = fused
        layer.zero_grad(set_to_none=True)
        out = layer(X, gg, seed=123)
        out.backward(go)
        results.append([out.detach().clone()] + [p.grad.clone() for p in layer.parameters()])
    for a, b in zip(*results):
        assert torch.equal(a, b)


@pytest.mark.parametrize("build", ["host", "device"])
def test_graph_aggregate_compact_equals_dense(cuda, graph, build):
    """GATGraph.aggregate over the compact edge-id lists of A and of A^T
    equals the dense edge-id map tiles bit for bit (host and device builds,
    symmetric and non-symmetric patterns)."""
    _, g = graph
    gg = GATGraph(g, cuda, build=build)
    assert gg.cmap_fwd is not None and gg.cmap_bwd is not None
    gen = torch.Generator(device="cpu").manual_seed(4)
    p = torch.rand(4, g.nnz, generator=gen).bfloat16().to(cuda)
    X = torch.randn(g.n, 4, 32, generator=gen).bfloat16().to(cuda)
    for transpose in (False, True):
        ref = gg.aggregate(p, X, transpose=transpose, compact=False)
        got = gg.aggregate(p, X, transpose=transpose)
        assert torch.equal(ref, got), transpose


@pytest.mark.parametrize("heads,dim", [(1, 64), (2, 32), (4, 64), (8, 64), (2, 128), (2, 16)])
@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("out_dtype", [torch.float32, torch.bfloat16])
def test_aggregate_sddmm_fused_equals_separate(cuda, graph, heads, dim, dtype, out_dtype):
    """The A^T aggregation with the SDDMM <dh[i], Xt[j]> taken from the same
    gathered rows (attgnn_spmm_edge_map_compact_dot) equals the separate
    aggregation and SDDMM bit for bit, on symmetric and non-symmetric
    patterns; 16 features per head run the separate products."""
    gg, g = graph
    assert gg.cmap_bwd is not None
    gen = torch.Generator(device="cpu").manual_seed(heads * 131 + dim)
    p = torch.rand(heads, g.nnz, generator=gen).to(dtype).to(cuda)
    dh = torch.randn(g.n, heads, dim, generator=gen).to(dtype).to(cuda)
    Xt = torch.randn(g.n, heads, dim, generator=gen).to(dtype).to(cuda)
    ref_y = gg.aggregate(p, dh, transpose=True, out_dtype=out_dtype)
    ref_dp = gg.sddmm(dh, Xt)
    y, dp = gg.aggregate_sddmm(p, dh, Xt, out_dtype=out_dtype)
    assert dp.shape == (heads, g.nnz) and dp.dtype == torch.float32
    assert torch.equal(ref_y, y)
    assert torch.equal(ref_dp, dp)
    # every edge's product against float64 on the CPU
    rows = torch.from_numpy(g.row_ids().astype(np.int64))
    cols = torch.from_numpy(g.indices.astype(np.int64))
    want = (dh.double().cpu()[rows] * Xt.double().cpu()[cols]).sum(-1).t()
    assert torch.allclose(dp.double().cpu(), want, rtol=1e-5, atol=1e-5)


def test_aggregate_sddmm_rejects_unsupported(cuda, graph):
    gg, g = graph
    rp, col, _, win, at = gg.bwd
    p = torch.rand(2, g.nnz, device=cuda).bfloat16()
    X = torch.randn(g.n, 2, 48, device=cuda).bfloat16()
    with pytest.raises(RuntimeError):
        ops.spmm_edge_map_compact_dot(rp, col, *gg.cmap_bwd, p, win, at, X, X, "mhd", gg.window, m_ori=g.n,
                                      x_layout="mhd", rows_covered=True)
    with pytest.raises(ValueError):          # dot_x of another width
        ops.spmm_edge_map_compact_dot(rp, col, *gg.cmap_bwd, p, win, at, X, X[:, :, :32].contiguous(), "mhd",
                                      gg.window, m_ori=g.n, x_layout="mhd", rows_covered=True)


@pytest.mark.parametrize("x_layout", ["mhd", "hmd"])
@pytest.mark.parametrize("heads,dim", [(1, 64), (3, 32), (2, 128)])
def test_aggregate_sddmm_dot_rows_differ(cuda, graph, x_layout, heads, dim):
    """dot_x with fewer rows than X (A^T of a rectangular A): its rows are
    addressed with its own head stride in [H, rows, D] and rows past its end
    read as zero, so the products equal those against dot_x padded with zero
    rows, and the aggregation is unchanged."""
    gg, g = graph
    rp, col, _, win, at = gg.bwd
    gen = torch.Generator(device="cpu").manual_seed(heads * 17 + dim)
    cut = g.n - 37
    p = torch.rand(heads, g.nnz, generator=gen).bfloat16().to(cuda)
    dh = torch.randn(g.n, heads, dim, generator=gen).bfloat16().to(cuda)
    Xt = torch.randn(cut, heads, dim, generator=gen).bfloat16().to(cuda)
    pad = torch.cat([Xt, torch.zeros(g.n - cut, heads, dim, dtype=Xt.dtype, device=cuda)])
    if x_layout == "hmd":
        dh, Xt, pad = (t.permute(1, 0, 2).contiguous() for t in (dh, Xt, pad))
    args = (rp, col, *gg.cmap_bwd, p, win, at, dh)
    kw = dict(m_ori=g.n, x_layout=x_layout, rows_covered=True)
    y_cut, dp_cut = ops.spmm_edge_map_compact_dot(*args, Xt, "mhd", gg.window, **kw)
    y_pad, dp_pad = ops.spmm_edge_map_compact_dot(*args, pad, "mhd", gg.window, **kw)
    assert torch.equal(y_cut, y_pad)
    assert torch.equal(dp_cut, dp_pad)
    # the window rows of A^T are A's columns: products of edges (i, j) with j >= cut are zero
    cols = torch.from_numpy(g.indices.astype(np.int64)).to(cuda)
    assert torch.count_nonzero(dp_cut[:, cols >= cut]) == 0
    assert torch.count_nonzero(dp_cut[:, cols < cut]) > 0


@pytest.mark.parametrize("heads,dim", [(1, 64), (4, 32), (2, 128)])
def test_aggregate_sddmm_fused_split_windows(cuda, heads, dim):
    """One block per part splits every window with more than 32 condensed
    columns: the fused kernel's aggregation goes through the split-window
    partials and the ordered combine, its dot products straight to the edges;
    both equal the separate kernels bit for bit."""
    g = G.chung_lu(900, 10, 2.1, seed=7)
    gg = GATGraph(g, cuda, build="device", part_size=1, chunk_blocks=1)
    assert int(gg.fwd[4].sum().item()) > 0          # some tasks belong to split windows
    gen = torch.Generator(device="cpu").manual_seed(dim)
    p = torch.rand(heads, g.nnz, generator=gen).bfloat16().to(cuda)
    dh = torch.randn(g.n, heads, dim, generator=gen).bfloat16().to(cuda)
    Xt = torch.randn(g.n, heads, dim, generator=gen).bfloat16().to(cuda)
    y, dp = gg.aggregate_sddmm(p, dh, Xt)
    assert torch.equal(y, gg.aggregate(p, dh, transpose=True))
    assert torch.equal(dp, gg.sddmm(dh, Xt))


@pytest.mark.parametrize("heads,dim", [(1, 64), (4, 64), (8, 32), (3, 16), (4, 12)])
@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16, torch.float32])
@pytest.mark.parametrize("out_layout", ["hn", "nh"])
def test_node_scores_pair_equals_two_calls(cuda, heads, dim, dtype, out_layout):
    """attgnn_node_scores_pair (one pass over X, vector kernel; generic kernel
    twice for unsupported shapes) equals two attgnn_node_scores calls bit for bit."""
    gen = torch.Generator(device="cpu").manual_seed(heads * 7 + dim)
    X = torch.randn(333, heads, dim, generator=gen).to(dtype).to(cuda)
    a0 = torch.randn(heads, dim, generator=gen).to(dtype).to(cuda)
    a1 = torch.randn(heads, dim, generator=gen).to(dtype).to(cuda)
    s0, s1 = ops.node_scores_pair(X, a0, a1, "mhd", out_layout=out_layout)
    assert torch.equal(s0, ops.node_scores(X, a0, "mhd", out_layout=out_layout))
    assert torch.equal(s1, ops.node_scores(X, a1, "mhd", out_layout=out_layout))
