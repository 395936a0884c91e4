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


def gat_reference(X, Ws, Wt, a0, a1, A, heads, dim, slope, mask=None, q=0.0, v2_alpha=None):
    """FP64 dense reference of the attention layer; returns [N, H, D]."""
    n = X.shape[0]
    Xs = (X @ Ws.t()).view(n, heads, dim)
    Xt = Xs if Wt is None else (X @ Wt.t()).view(n, heads, dim)
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


@pytest.mark.parametrize("heads,dim", [(1, 64), (2, 32), (4, 16), (8, 8)])
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
    W = layer.fc.weight.detach().double().cpu().requires_grad_(True)
    a0 = layer.a0.detach().double().cpu().view(heads, dim).requires_grad_(True)
    a1 = layer.a1.detach().double().cpu().view(heads, dim).requires_grad_(True)
    Xd = X.double().cpu()
    if dtype == torch.bfloat16:   # same rounding of the inputs the GPU path sees
        Xd = Xd.to(torch.bfloat16).double()
    ref = gat_reference(Xd, W, None, a0, a1, A, heads, dim, 0.2)
    (ref * go.double().cpu()).sum().backward()
    rel = lambda a, b: ((a.double().cpu() - b).norm() / b.norm()).item()
    assert rel(out, ref) < tol
    assert rel(layer.fc.weight.grad, W.grad) < tol * 3
    assert rel(layer.a0.grad.view(heads, dim), a0.grad) < tol * 3
    assert rel(layer.a1.grad.view(heads, dim), a1.grad) < tol * 3


@pytest.mark.parametrize("heads,dim", [(1, 32), (4, 16)])
def test_gatv2_layer_vs_fp64(cuda, heads, dim):
    g = small_graph(300, 5, seed=7, symmetric=False)
    graph = GATGraph(g, cuda)
    torch.manual_seed(3)
    layer = GATv2Conv_multi(24, dim, 0.0, 0.2, heads, dtype=torch.float32).to(cuda)
    X = torch.randn(g.n, 24, device=cuda)
    out = layer(X, graph)
    go = torch.randn(out.shape, device=cuda)
    (out * go).sum().backward()
    A = dense_adj(g)
    W0 = layer.fc0.weight.detach().double().cpu().requires_grad_(True)
    W1 = layer.fc1.weight.detach().double().cpu().requires_grad_(True)
    a = layer.a0.detach().double().cpu().view(heads, dim).requires_grad_(True)
    ref = gat_reference(X.double().cpu(), W0, W1, a, a, A, heads, dim, 1.0, v2_alpha=0.2)
    (ref * go.double().cpu()).sum().backward()
    rel = lambda x, y: ((x.double().cpu() - y).norm() / y.norm()).item()
    assert rel(out, ref) < 2e-4
    # a.X0[i] is constant along row i, so the softmax cancels it: dL/dW0 = 0
    # exactly (in the reference's formulation as well)
    assert W0.grad.norm() < 1e-10 * W1.grad.norm()
    assert layer.fc0.weight.grad.norm() < 1e-5 * layer.fc1.weight.grad.norm()
    assert rel(layer.fc1.weight.grad, W1.grad) < 1e-3
    assert rel(layer.a0.grad.view(heads, dim), a.grad) < 1e-3


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
