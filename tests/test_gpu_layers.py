"""GCN / GAT / GATv2 layers on the GPU: forward and exact gradients against a
float64 torch reference on the CPU (dense masked attention), the reference's
own formula pinned by FP64 finite differences, and dropout-mask regeneration."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from att_amd import graphs as G
from att_amd import ops
from att_amd.layers import GATConv_multi, GATGraph, GATv2Conv_multi, GCNConv, GCNGraph

pytestmark = pytest.mark.gpu


def small_graph(n=300, deg=6, seed=0, symmetric=True):
    g = G.chung_lu(n, deg, 2.1, seed=seed)
    if symmetric:
        return g
    rng = np.random.default_rng(seed)
    keep = rng.random(g.nnz) < 0.7
    rows = g.row_ids()[keep].astype(np.int64)
    cols = g.indices[keep].astype(np.int64)
    keys = np.unique(np.concatenate([rows * n + cols, np.arange(n) * (n + 1)]))
    ip = np.zeros(n + 1, np.int64)
    ip[1:] = np.cumsum(np.bincount(keys // n, minlength=n))
    return G.CSR(ip, (keys % n).astype(np.int32), n)


def dense_adj(g):
    A = torch.zeros(g.n, g.n, dtype=torch.bool)
    A[torch.from_numpy(g.row_ids().astype(np.int64)), torch.from_numpy(g.indices.astype(np.int64))] = True
    return A


def _round_through(x, dtype):
    """x rounded to ``dtype`` in the forward, identity in the backward (the
    projection's output is stored rounded; its gradient is taken at that point)."""
    return x if dtype is None else x + (x.to(dtype).double() - x).detach()


def gat_reference(X, Ws, Wt, a0, a1, A, heads, dim, slope, mask=None, q=0.0, v2_alpha=None, round_to=None):
    """FP64 dense reference of the attention layer; returns [N, H, D]."""
    n = X.shape[0]
    Xs = _round_through(X @ Ws.t(), round_to).view(n, heads, dim)
    Xt = Xs if Wt is None else _round_through(X @ Wt.t(), round_to).view(n, heads, dim)
    if v2_alpha is not None:
        Xs = F.leaky_relu(Xs, v2_alpha)
        Xt = F.leaky_relu(Xt, v2_alpha)
    s0 = (Xs * a0.view(1, heads, dim)).sum(-1)      # [N, H]
    s1 = (Xt * a1.view(1, heads, dim)).sum(-1)
    z = s0.t().unsqueeze(2) + s1.t().unsqueeze(1)    # [H, N, N]
    e = torch.where(z > 0, z, slope * z).masked_fill(~A.unsqueeze(0), float("-inf"))
    P = torch.softmax(e, dim=-1)
    P = torch.nan_to_num(P)
    if mask is not None:
        P = P * mask / (1.0 - q)
    return torch.einsum("hij,jhd->ihd", P, Xt)


def test_gcn_layer_forward_backward(cuda):
    for symmetric in (True, False):
        g = small_graph(400, 6, seed=1, symmetric=symmetric)
        d = G.gcn_norm_degree(g)
        vals = G.gcn_norm_values(g, d)
        graph = GCNGraph(g, cuda, tile_dtype=torch.float32)
        torch.manual_seed(0)
        conv = GCNConv(40, 24, dtype=torch.float32).to(cuda)
        X = torch.randn(g.n, 40, device=cuda)
        y = conv(X, graph)
        gy = torch.randn_like(y)
        (y * gy).sum().backward()
        idx = torch.from_numpy(np.stack([g.row_ids(), g.indices]).astype(np.int64))
        A = torch.sparse_coo_tensor(idx, torch.from_numpy(vals.astype(np.float64)), (g.n, g.n)).to_dense()
        W = conv.weights.detach().double().cpu().requires_grad_(True)
        Xd = X.double().cpu()
        ref = A @ (Xd @ W)
        (ref * gy.double().cpu()).sum().backward()
        ferr = ((y.detach().double().cpu() - ref).norm() / ref.norm()).item()
        assert ferr < 1e-5, (symmetric, ferr)
        gerr = (conv.weights.grad.double().cpu() - W.grad).norm() / W.grad.norm()
        assert gerr < 1e-5, (symmetric, gerr)


@pytest.mark.parametrize("heads,dim", [(1, 64), (2, 32), (4, 16), (8, 8), (2, 64), (4, 64), (8, 64)])
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 2e-4), (torch.bfloat16, 3e-2)])
def test_gat_layer_vs_fp64(cuda, heads, dim, dtype, tol):
    g = small_graph(350, 6, seed=heads)
    graph = GATGraph(g, cuda)
    torch.manual_seed(heads)
    layer = GATConv_multi(48, dim, 0.0, 0.2, heads, dtype=dtype).to(cuda)
    X = torch.randn(g.n, 48, device=cuda)
    out = layer(X, graph)
    go = torch.randn(out.shape, device=cuda)
    (out.float() * go).sum().backward()

    A = dense_adj(g)
    # the float64 reference sees the roundings of the GPU path's inputs and
    # stored intermediates: X, W, a0, a1 and the upstream gradient in bf16,
    # and X W stored in bf16 (for fp32 all of these are exact)
    rd = dtype if dtype != torch.float32 else None
    r = (lambda t: t.to(rd).double()) if rd else (lambda t: t.double())
    W = r(layer.fc.weight.detach()).cpu().requires_grad_(True)
    a0 = r(layer.a0.detach()).cpu().view(heads, dim).requires_grad_(True)
    a1 = r(layer.a1.detach()).cpu().view(heads, dim).requires_grad_(True)
    ref = gat_reference(r(X).cpu(), W, None, a0, a1, A, heads, dim, 0.2, round_to=rd)
    (ref * r(go).cpu()).sum().backward()
    rel = lambda a, b: ((a.double().cpu() - b).norm() / b.norm()).item()
    assert rel(out, ref) < tol
    assert rel(layer.fc.weight.grad, W.grad) < tol * 3
    assert rel(layer.a0.grad.view(heads, dim), a0.grad) < tol * 3
    assert rel(layer.a1.grad.view(heads, dim), a1.grad) < tol * 3


@pytest.mark.parametrize("heads,dim,dtype,tol", [(1, 32, torch.float32, 2e-4), (4, 16, torch.float32, 2e-4),
                                                 (1, 64, torch.bfloat16, 3e-2), (4, 64, torch.bfloat16, 3e-2)])
def test_gatv2_layer_vs_fp64(cuda, heads, dim, dtype, tol):
    """The single-head reference file's logits (no min-max map)."""
    g = small_graph(300, 5, seed=7, symmetric=False)
    graph = GATGraph(g, cuda)
    torch.manual_seed(3)
    layer = GATv2Conv_multi(24, dim, 0.0, 0.2, heads, dtype=dtype, minmax=False).to(cuda)
    X = torch.randn(g.n, 24, device=cuda)
    out = layer(X, graph)
    go = torch.randn(out.shape, device=cuda)
    (out.float() * go).sum().backward()
    A = dense_adj(g)
    W0 = layer.fc0.weight.detach().to(dtype).double().cpu().requires_grad_(True)
    W1 = layer.fc1.weight.detach().to(dtype).double().cpu().requires_grad_(True)
    a = layer.a0.detach().double().cpu().view(heads, dim).requires_grad_(True)
    ref = gat_reference(X.to(dtype).double().cpu(), W0, W1, a, a, A, heads, dim, 1.0, v2_alpha=0.2)
    (ref * go.double().cpu()).sum().backward()
    rel = lambda x, y: ((x.double().cpu() - y).norm() / y.norm()).item()
    assert rel(out, ref) < tol
    # a.X0[i] is constant along row i, so the softmax cancels it: dL/dW0 = 0
    # exactly (in the reference's formulation as well)
    assert W0.grad.norm() < 1e-10 * W1.grad.norm()
    assert layer.fc0.weight.grad.norm() < 1e-5 * layer.fc1.weight.grad.norm()
    gtol = 1e-3 if dtype == torch.float32 else 3 * tol
    assert rel(layer.fc1.weight.grad, W1.grad) < gtol
    assert rel(layer.a0.grad.view(heads, dim), a.grad) < gtol


def gatv2_minmax_reference(X, W0, W1, a, rows, cols, n, heads, dim, alpha, at=None):
    """float64 restatement of the multi-head GATv2 forward
    (end2end_head/gatv2_final/mgat_csr/mgat_conv.py:171-197, no dropout):
    X0 = LReLU(X W0), X1 = LReLU(X W1), att = a.X0[i] + a.X1[j] per edge,
    att = (att - min_h) / (max_h - min_h) over each head's edges, p = exp(att),
    h = (sum_j p X1[j]) / (sum_j p); [N, H, D].  ``at``: (edge of the minimum,
    edge of the maximum) per head to take them at instead of searching (a
    subgradient choice at near-ties); returns (h, z)."""
    X0 = F.leaky_relu(X @ W0.t(), alpha).view(n, heads, dim)
    X1 = F.leaky_relu(X @ W1.t(), alpha).view(n, heads, dim)
    s0 = (X0 * a.view(1, heads, dim)).sum(-1)
    s1 = (X1 * a.view(1, heads, dim)).sum(-1)
    z = s0[rows] + s1[cols]                                    # [E, H]
    hs = torch.arange(heads)
    if at is None:
        zmin, zmax = z.min(0).values, z.max(0).values
    else:
        zmin, zmax = z[at[0], hs], z[at[1], hs]
    p = torch.exp((z - zmin) / (zmax - zmin))
    S = torch.zeros(n, heads, dtype=X.dtype).index_add(0, rows, p)
    h = torch.zeros(n, heads, dim, dtype=X.dtype).index_add(0, rows, p.unsqueeze(-1) * X1[cols])
    return h / S.unsqueeze(-1), z


@pytest.mark.parametrize("heads,dim", [(1, 64), (2, 64), (4, 64), (2, 32)])
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 2e-4), (torch.bfloat16, 3e-2)])
@pytest.mark.parametrize("symmetric", [True, False])
def test_gatv2_minmax_layer_vs_fp64(cuda, heads, dim, dtype, tol, symmetric):
    """GATv2Conv_multi built with a head count maps each head's logits to
    [0, 1] by their global min and max, as the reference's multi-head layer
    does; forward and the gradients of W0, W1 and a (W0 reaches the loss only
    through the row terms of the two edges that set min and max) against the
    float64 restatement, with the extreme edges the kernel found (which the
    float64 logits must confirm as extremes up to the input rounding)."""
    g = small_graph(400, 6, seed=heads + dim, symmetric=symmetric)
    graph = GATGraph(g, cuda, build="device")
    torch.manual_seed(heads)
    layer = GATv2Conv_multi(24, dim, 0.0, 0.2, heads, dtype=dtype).to(cuda)
    assert layer.minmax
    X = torch.randn(g.n, 24, device=cuda)
    out = layer(X, graph)
    go = torch.randn(out.shape, device=cuda)
    (out.float() * go).sum().backward()

    # the extreme edges the layer used (same deterministic kernels as inside it)
    with torch.no_grad():
        Xd = X.to(dtype)
        X0 = F.leaky_relu(F.linear(Xd, layer.fc0.weight.to(dtype)), 0.2).view(-1, heads, dim)
        X1 = F.leaky_relu(F.linear(Xd, layer.fc1.weight.to(dtype)), 0.2).view(-1, heads, dim)
        av = layer.a0.view(heads, dim).to(dtype).contiguous()
        s0 = ops.node_scores(X0, av, "mhd", out_layout="nh")
        s1 = ops.node_scores(X1, av, "mhd", out_layout="nh")
        rng, where = ops.edge_score_range(graph.rowptr, graph.col, s0, s1)
    where = where.cpu().numpy()
    rows_np, cols_np = g.row_ids().astype(np.int64), g.indices.astype(np.int64)

    def edge_id(r, c):
        lo, hi = g.indptr[r], g.indptr[r + 1]
        return lo + int(np.searchsorted(g.indices[lo:hi], c))

    e_min = torch.tensor([edge_id(*where[h, 0:2]) for h in range(heads)])
    e_max = torch.tensor([edge_id(*where[h, 2:4]) for h in range(heads)])
    rows, cols = torch.from_numpy(rows_np), torch.from_numpy(cols_np)
    W0 = layer.fc0.weight.detach().to(dtype).double().cpu().requires_grad_(True)
    W1 = layer.fc1.weight.detach().to(dtype).double().cpu().requires_grad_(True)
    a = layer.a0.detach().to(dtype).double().cpu().view(heads, dim).requires_grad_(True)
    ref, z = gatv2_minmax_reference(X.to(dtype).double().cpu(), W0, W1, a, rows, cols, g.n, heads, dim, 0.2,
                                    at=(e_min, e_max))
    (ref * go.double().cpu()).sum().backward()
    zr = z.detach()
    span = (zr.max(0).values - zr.min(0).values)
    hs = torch.arange(heads)
    assert torch.all(zr[e_min, hs] - zr.min(0).values <= tol * span)
    assert torch.all(zr.max(0).values - zr[e_max, hs] <= tol * span)
    assert torch.allclose(rng[:, 2].double().cpu(), 1.0 / (zr[e_max, hs] - zr[e_min, hs]), rtol=3 * tol)
    rel = lambda x, y: ((x.double().cpu() - y).norm() / y.norm()).item()
    assert rel(out.float(), ref) < tol
    gtol = 1e-3 if dtype == torch.float32 else 3 * tol
    assert rel(layer.fc0.weight.grad, W0.grad) < gtol
    assert rel(layer.fc1.weight.grad, W1.grad) < gtol
    assert rel(layer.a0.grad.view(heads, dim), a.grad) < gtol
    assert W0.grad.norm() > 1e-3 * W1.grad.norm()        # the min / max edges do reach W0


def test_edge_score_range_vs_torch(cuda):
    """attgnn_edge_score_range: min, max, temperature and the first extreme
    edge of each head against torch on the per-edge logits; both layouts;
    deterministic; a graph with no edges gives -1 and a zero temperature."""
    g = small_graph(700, 7, seed=3, symmetric=False)
    rowptr = torch.from_numpy(g.rowptr32()).to(cuda)
    col = torch.from_numpy(g.indices).to(cuda)
    gen = torch.Generator().manual_seed(5)
    heads = 3
    s0 = torch.randn(g.n, heads, generator=gen).to(cuda)
    s1 = torch.randn(g.n, heads, generator=gen).to(cuda)
    s1[17, 1] = 1e3        # one column term dominates head 1's maximum (several edges share it)
    rows = torch.from_numpy(g.row_ids().astype(np.int64)).to(cuda)
    z = s0[rows] + s1[col.long()]                              # [E, H]
    for layout in ("nh", "hn"):
        a0, a1 = (s0, s1) if layout == "nh" else (s0.t().contiguous(), s1.t().contiguous())
        rng, where = ops.edge_score_range(rowptr, col, a0, a1, s_layout=layout)
        rng2, where2 = ops.edge_score_range(rowptr, col, a0, a1, s_layout=layout)
        assert torch.equal(rng, rng2) and torch.equal(where, where2)
        assert torch.equal(rng[:, 0], z.min(0).values) and torch.equal(rng[:, 1], z.max(0).values)
        assert torch.allclose(rng[:, 2], 1.0 / (rng[:, 1] - rng[:, 0]))
        for h in range(heads):
            first_min = int((z[:, h] == z[:, h].min()).nonzero()[0])
            first_max = int((z[:, h] == z[:, h].max()).nonzero()[0])
            assert where[h, 0].item() == rows[first_min].item() and where[h, 1].item() == g.indices[first_min]
            assert where[h, 2].item() == rows[first_max].item() and where[h, 3].item() == g.indices[first_max]
    empty = torch.zeros(5, dtype=torch.int32, device=cuda)
    rng, where = ops.edge_score_range(empty, col[:0], s0[:4], s1[:4])
    assert torch.all(where.cpu() == -1) and torch.all(rng[:, 2].cpu() == 0)


def test_edge_score_range_long_rows(cuda):
    """With the long-row list (rows above 256 entries taken by whole
    workgroups) the range is bitwise the one without it and torch's, on a
    graph with hub rows of 300-5,000 entries."""
    rng_np = np.random.default_rng(7)
    n = 6000
    deg = rng_np.integers(0, 12, n)
    deg[[3, 100, 4096, 5999]] = [5000, 300, 1200, 257]
    ip = np.concatenate([[0], np.cumsum(deg)])
    cols = np.concatenate([np.sort(rng_np.choice(n, d, replace=False)) for d in deg]).astype(np.int32)
    g = G.CSR(ip, cols, n)
    rowptr = torch.from_numpy(g.rowptr32()).to(cuda)
    col = torch.from_numpy(g.indices).to(cuda)
    gen = torch.Generator().manual_seed(8)
    heads = 5
    s0 = torch.randn(n, heads, generator=gen).to(cuda)
    s1 = torch.randn(n, heads, generator=gen).to(cuda)
    s0[100, 2] = 50.0       # head 2's maximum lies in a long row
    s0[3, 4] = -50.0        # head 4's minimum lies in the longest row
    lr = ops.long_rows(rowptr)
    assert sorted(lr.cpu().tolist()) == [3, 100, 4096, 5999]
    a = ops.edge_score_range(rowptr, col, s0, s1)
    b = ops.edge_score_range(rowptr, col, s0, s1, long_rows=lr)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    rows = torch.from_numpy(g.row_ids().astype(np.int64)).to(cuda)
    z = s0[rows] + s1[col.long()]
    assert torch.equal(b[0][:, 0], z.min(0).values) and torch.equal(b[0][:, 1], z.max(0).values)
    assert b[1][2, 2].item() == 100 and b[1][4, 0].item() == 3


def _minmax_score_grads_torch(s0, s1, scale, where, ds0s, ds1s):
    """Torch statement of the gradient through the per-head temperature c_h =
    1 / (max_h - min_h) (s0 / s1 fp32 [N, H], ds0s / ds1s of the scaled terms
    fp32 [H, N]): with U_h = sum_i s0[i,h] ds0s[h,i] + s1[i,h] ds1s[h,i] =
    dL/dc_h, dc/dmax = -c^2 and dc/dmin = c^2, ds0 = c ds0s, ds1 = c ds1s, then
    -U c^2 at the row and column of the maximum's edge and +U c^2 at those of
    the minimum's, in that order."""
    c = scale.view(-1, 1)
    U = (s0.t() * ds0s).sum(1) + (s1.t() * ds1s).sum(1)          # [H]
    k = U * scale * scale
    ds0 = ds0s * c
    ds1 = ds1s * c
    for h in range(scale.numel()):
        r_min, c_min, r_max, c_max = (int(v) for v in where[h].tolist())
        if r_max >= 0:
            ds0[h, r_max] -= k[h]
            ds1[h, c_max] -= k[h]
            ds0[h, r_min] += k[h]
            ds1[h, c_min] += k[h]
    return ds0, ds1


@pytest.mark.parametrize("layout", ["nh", "hn"])
def test_minmax_score_grads_vs_torch(cuda, layout):
    """attgnn_minmax_score_grads against the torch statement: equal to ~1e-6
    (U is summed in another order) and bitwise away from the extreme edges'
    rows and columns; a head whose minimum and maximum share a row and a
    column, a head with zero temperature and no edges, and run-to-run
    equality."""
    n, heads = 5000, 3
    gen = torch.Generator().manual_seed(11)
    s0 = torch.randn(n, heads, generator=gen).to(cuda)
    s1 = torch.randn(n, heads, generator=gen).to(cuda)
    ds0s = torch.randn(heads, n, generator=gen).to(cuda)
    ds1s = torch.randn(heads, n, generator=gen).to(cuda)
    rng = torch.tensor([[-2.0, 3.0, 0.2], [-1.0, 1.5, 0.4], [0.0, 0.0, 0.0]], device=cuda)
    where = torch.tensor([[17, 4001, 17, 4001], [3, 9, 2500, 77], [-1, -1, -1, -1]], dtype=torch.int32,
                         device=cuda)
    a0, a1 = (s0, s1) if layout == "nh" else (s0.t().contiguous(), s1.t().contiguous())
    ds0, ds1 = ops.minmax_score_grads(a0, a1, rng, where, ds0s, ds1s, s_layout=layout)
    again = ops.minmax_score_grads(a0, a1, rng, where, ds0s, ds1s, s_layout=layout)
    assert torch.equal(ds0, again[0]) and torch.equal(ds1, again[1])
    r0, r1 = _minmax_score_grads_torch(s0, s1, rng[:, 2], where, ds0s, ds1s)
    assert torch.allclose(ds0, r0, rtol=1e-5, atol=1e-6) and torch.allclose(ds1, r1, rtol=1e-5, atol=1e-6)
    keep0 = torch.ones(heads, n, dtype=torch.bool, device=cuda)
    keep1 = keep0.clone()
    for h, (rmin, cmin, rmax, cmax) in enumerate(where.tolist()):
        if rmin >= 0:
            keep0[h, [rmin, rmax]] = False
            keep1[h, [cmin, cmax]] = False
    assert torch.equal(ds0[keep0], r0[keep0]) and torch.equal(ds1[keep1], r1[keep1])
    assert torch.all(ds0[2] == 0) and torch.all(ds1[2] == 0)
    # head 0: min and max at one row / column -> (c ds - k) + k there
    assert abs(float(ds0[0, 17]) - float(ds0s[0, 17] * 0.2)) < 1e-4


def test_gcn_fp16_two_layers_vs_fp64(cuda):
    """The 2-layer GCN of the benchmark in its storage type -- FP16 tiles,
    features and weights, FP32 accumulation; 128 hidden features (compact
    tiles) and 48 outputs padded to 64 (dense tiles), ReLU -- on a
    20,000-node Chung-Lu graph: the output and both weight gradients against
    float64 on the same FP16-rounded inputs."""
    g = G.chung_lu(20000, 10, 2.1, seed=4)
    graph = GCNGraph(g, cuda, tile_dtype=torch.float16, build="device")
    torch.manual_seed(1)
    c1, c2 = GCNConv(96, 128).to(cuda), GCNConv(128, 48).to(cuda)
    X = torch.randn(g.n, 96, device=cuda).half()
    y = c2(F.relu(c1(X, graph)), graph)
    assert y.dtype == torch.float16 and y.shape == (g.n, 48)
    go = torch.randn(y.shape, device=cuda)
    (y.float() * go).sum().backward()

    d = G.gcn_norm_degree(g)
    vals = torch.from_numpy(G.gcn_norm_values(g, d).astype(np.float32)).half().double()
    idx = torch.from_numpy(np.stack([g.row_ids(), g.indices]).astype(np.int64))
    A = torch.sparse_coo_tensor(idx, vals, (g.n, g.n)).coalesce()
    W1 = c1.weights.detach().half().double().cpu().requires_grad_(True)
    W2 = c2.weights.detach().half().double().cpu().requires_grad_(True)
    h = F.relu(torch.sparse.mm(A, X.double().cpu() @ W1))
    ref = torch.sparse.mm(A, h @ W2)
    (ref * go.double().cpu()).sum().backward()
    rel = lambda x, r: ((x.double().cpu() - r).norm() / r.norm()).item()
    assert rel(y, ref) < 5e-3
    assert rel(c1.weights.grad, W1.grad) < 1e-2
    assert rel(c2.weights.grad, W2.grad) < 1e-2


def test_gat_dropout_mask_regenerated(cuda):
    """With dropout the backward pass regenerates the forward mask: gradients
    match the float64 reference built with the very mask the kernel applied."""
    g = small_graph(300, 6, seed=5)
    graph = GATGraph(g, cuda)
    heads, dim, q, seed = 2, 16, 0.3, 1234
    torch.manual_seed(9)
    layer = GATConv_multi(20, dim, q, 0.2, heads, dtype=torch.float32).to(cuda)
    X = torch.randn(g.n, 20, device=cuda)
    out = layer(X, graph, seed=seed)
    out2 = layer(X, graph, seed=seed)
    assert torch.equal(out, out2)
    go = torch.randn(out.shape, device=cuda)
    (out * go).sum().backward()
    # recover the mask from the fused kernel itself: p_drop / p_keep
    Xp = (X @ layer.fc.weight.t()).view(g.n, heads, dim)
    s0 = ops.node_scores(Xp, layer.a0.detach().view(heads, dim).contiguous(), "mhd")
    s1 = ops.node_scores(Xp, layer.a1.detach().view(heads, dim).contiguous(), "mhd")
    p_keep, _ = ops.edge_softmax_fwd(graph.rowptr, graph.col, s0, s1, 0.2, 0.0, seed)
    p_drop, _ = ops.edge_softmax_fwd(graph.rowptr, graph.col, s0, s1, 0.2, q, seed)
    kept = (p_drop != 0).cpu()
    frac = kept.float().mean().item()
    assert abs(frac - (1 - q)) < 0.05
    mask = torch.zeros(heads, g.n, g.n, dtype=torch.float64)
    rows = torch.from_numpy(g.row_ids().astype(np.int64))
    cols = torch.from_numpy(g.indices.astype(np.int64))
    for h in range(heads):
        mask[h, rows, cols] = kept[h].double()
    A = dense_adj(g)
    W = layer.fc.weight.detach().double().cpu().requires_grad_(True)
    a0 = layer.a0.detach().double().cpu().view(heads, dim).requires_grad_(True)
    a1 = layer.a1.detach().double().cpu().view(heads, dim).requires_grad_(True)
    ref = gat_reference(X.double().cpu(), W, None, a0, a1, A, heads, dim, 0.2, mask=mask, q=q)
    (ref * go.double().cpu()).sum().backward()
    rel = lambda x, y: ((x.double().cpu() - y).norm() / y.norm()).item()
    assert rel(out, ref) < 2e-4
    assert rel(layer.fc.weight.grad, W.grad) < 1e-3
    assert rel(layer.a0.grad.view(heads, dim), a0.grad) < 1e-3
    assert torch.allclose(p_drop.cpu()[kept], p_keep.cpu()[kept] / (1 - q), rtol=1e-6)


@pytest.mark.parametrize("dtype,dim,shared", [(torch.float32, 64, False), (torch.bfloat16, 64, True),
                                              (torch.float16, 32, False), (torch.bfloat16, 6, False),
                                              (torch.float32, 10, True)])
def test_attention_node_bwd_fused(dtype, dim, shared, cuda):
    """attgnn_attention_node_bwd against float64 torch: feature gradients (with the
    aggregation gradient folded in) and attention-vector gradients; dim 6 / 10
    take the scalar path.  Deterministic across runs."""
    gen = torch.Generator().manual_seed(dim)
    n, heads = 3001, 3
    Xs = torch.randn(n, heads, dim, generator=gen).to(dtype)
    Xt = Xs if shared else torch.randn(n, heads, dim, generator=gen).to(dtype)
    a0, a1 = torch.randn(heads, dim, generator=gen), torch.randn(heads, dim, generator=gen)
    ds0, ds1 = torch.randn(heads, n, generator=gen), torch.randn(heads, n, generator=gen)
    base = torch.randn(n, heads, dim, generator=gen)
    Xs_d = Xs.to(cuda)
    Xt_d = Xs_d if shared else Xt.to(cuda)
    out = ops.attention_node_bwd(Xs_d, Xt_d, a0.to(cuda), a1.to(cuda), ds0.to(cuda), ds1.to(cuda),
                                 base=base.to(cuda), combine=shared)
    again = ops.attention_node_bwd(Xs_d, Xt_d, a0.to(cuda), a1.to(cuda), ds0.to(cuda), ds1.to(cuda),
                                   base=base.to(cuda), combine=shared)
    xs, xt = Xs.double(), Xt.double()
    g0 = ds0.double().t().unsqueeze(-1) * a0.double().unsqueeze(0)
    g1 = ds1.double().t().unsqueeze(-1) * a1.double().unsqueeze(0)
    want_da0 = torch.einsum("hi,ihd->hd", ds0.double(), xs)
    want_da1 = torch.einsum("hi,ihd->hd", ds1.double(), xt)
    want_dxt = base.double() + g1 + (g0 if shared else 0)
    rel = lambda a, b: ((a.double().cpu() - b).norm() / b.norm()).item()
    tol_x = 1e-6 if dtype == torch.float32 else (4e-3 if dtype == torch.float16 else 8e-3)
    assert rel(out[1], want_dxt) < tol_x
    if not shared:
        assert rel(out[0], g0) < tol_x
    else:
        assert out[0] is None
    assert rel(out[2], want_da0) < 1e-5
    assert rel(out[3], want_da1) < 1e-5
    for a, b in zip(out, again):
        assert (a is None and b is None) or torch.equal(a, b)


def gatv2_true_reference(X, W0, W1, a, A, heads, dim, slope):
    """FP64 dense GATv2 (the paper's attention): [N, H, D]."""
    n = X.shape[0]
    X0 = (X @ W0.t()).view(n, heads, dim)
    X1 = (X @ W1.t()).view(n, heads, dim)
    U = X0.unsqueeze(1) + X1.unsqueeze(0)                            # [i, j, H, D]
    z = (F.leaky_relu(U, slope) * a.view(1, 1, heads, dim)).sum(-1)  # [i, j, H]
    e = z.permute(2, 0, 1).masked_fill(~A.unsqueeze(0), float("-inf"))
    P = torch.nan_to_num(torch.softmax(e, dim=-1))
    return torch.einsum("hij,jhd->ihd", P, X1)


@pytest.mark.parametrize("heads,dim,symmetric,dtype,tol", [(1, 64, True, torch.float32, 2e-4),
                                                           (4, 16, False, torch.float32, 2e-4),
                                                           (2, 24, True, torch.float32, 2e-4),
                                                           (2, 128, False, torch.float32, 2e-4),
                                                           (4, 32, True, torch.bfloat16, 3e-2)])
def test_true_gatv2_vs_fp64(heads, dim, symmetric, dtype, tol, cuda):
    g = small_graph(200, 5, seed=11, symmetric=symmetric)
    graph = GATGraph(g, cuda, build="device" if symmetric else "host")
    torch.manual_seed(4)
    layer = GATv2Conv_multi(20, dim, 0.0, 0.2, heads, dtype=dtype, true_gatv2=True).to(cuda)
    X = torch.randn(g.n, 20, device=cuda)
    out = layer(X, graph)
    go = torch.randn(out.shape, device=cuda)
    (out.float() * go).sum().backward()
    A = dense_adj(g)
    Xr = X.to(dtype).double().cpu()
    W0 = layer.fc0.weight.detach().to(dtype).double().cpu().requires_grad_(True)
    W1 = layer.fc1.weight.detach().to(dtype).double().cpu().requires_grad_(True)
    a = layer.a0.detach().double().cpu().view(heads, dim).requires_grad_(True)
    ref = gatv2_true_reference(Xr, W0, W1, a, A, heads, dim, 0.2)
    (ref * go.double().cpu()).sum().backward()
    rel = lambda x, y: ((x.double().cpu() - y).norm() / y.norm()).item()
    assert rel(out.float(), ref) < tol
    gtol = 1e-3 if dtype == torch.float32 else 5e-2
    assert rel(layer.fc0.weight.grad, W0.grad) < gtol
    assert rel(layer.fc1.weight.grad, W1.grad) < gtol
    assert rel(layer.a0.grad.view(heads, dim), a.grad) < gtol


def test_true_gatv2_logits_and_dropout(cuda):
    g = small_graph(300, 6, seed=12)
    graph = GATGraph(g, cuda)
    heads, dim = 2, 16
    X0 = torch.randn(g.n, heads, dim, device=cuda)
    X1 = torch.randn(g.n, heads, dim, device=cuda)
    att = torch.randn(heads, dim, device=cuda)
    segs, _ = graph.segments()
    z, lse, p = ops.gatv2_fwd(graph.rowptr, graph.col, segs, X0, X1, att, 0.2)
    rows = torch.from_numpy(g.row_ids().astype(np.int64)).to(cuda)
    cols = torch.from_numpy(g.indices.astype(np.int64)).to(cuda)
    want = (F.leaky_relu(X0[rows].double() + X1[cols].double(), 0.2) * att.double()).sum(-1).t()
    assert ((z.double() - want).abs().max() / want.abs().max()).item() < 1e-6
    # probabilities sum to one per row; dropout keeps a deterministic subset scaled by 1/(1-q)
    sums = torch.zeros(heads, g.n, device=cuda, dtype=torch.float64).index_add_(1, rows, p.double())
    nonempty = torch.from_numpy(np.diff(g.indptr) > 0).to(cuda)
    assert (sums[:, nonempty] - 1).abs().max().item() < 1e-5
    _, _, pd = ops.gatv2_fwd(graph.rowptr, graph.col, segs, X0, X1, att, 0.2, drop_p=0.4, seed=7)
    _, _, pd2 = ops.gatv2_fwd(graph.rowptr, graph.col, segs, X0, X1, att, 0.2, drop_p=0.4, seed=7)
    assert torch.equal(pd, pd2)
    kept = pd != 0
    assert 0.5 < kept.float().mean().item() < 0.7
    assert torch.allclose(pd[kept], p[kept] / 0.6, rtol=1e-5)
