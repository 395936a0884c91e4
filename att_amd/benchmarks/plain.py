"""Plain FP32 PyTorch GPU implementations used as the drop-in baselines:
torch.sparse CSR SpMM for GCN and gather / scatter-softmax / index_add message
passing for the attention layers (autograd for the backward pass)."""
import torch
import torch.nn.functional as F


def csr_tensor(graph_csr, values, device):
    """FP32 CSR tensor with int32 row pointers and column indices (the index
    type of the window-block path; the nnz must stay below 2^31)."""
    import numpy as np
    if graph_csr.nnz >= 2 ** 31:
        raise ValueError("int32 CSR indices need nnz < 2^31")
    return torch.sparse_csr_tensor(torch.from_numpy(graph_csr.indptr.astype(np.int32)),
                                   torch.from_numpy(graph_csr.indices.astype(np.int32)),
                                   torch.from_numpy(np.asarray(values, np.float32)), (graph_csr.n, graph_csr.n)).to(device)


def gat_fp32(X, W, a0, a1, rows, cols, n, heads, dim, slope=0.2, Wt=None, v2_alpha=None, minmax=False):
    """One attention layer in FP32: returns [N, H, D].  ``minmax``: each head's
    logits mapped to (z - min) / (max - min) over all its edges before the
    softmax (the reference's multi-head GATv2)."""
    Xs = (X @ W.t()).view(n, heads, dim)
    Xt = Xs if Wt is None else (X @ Wt.t()).view(n, heads, dim)
    if v2_alpha is not None:
        Xs = F.leaky_relu(Xs, v2_alpha)
        Xt = F.leaky_relu(Xt, v2_alpha)
    s0 = (Xs * a0.view(1, heads, dim)).sum(-1)
    s1 = (Xt * a1.view(1, heads, dim)).sum(-1)
    z = s0[rows] + s1[cols]
    e = F.leaky_relu(z, slope) if slope != 1.0 else z
    if minmax:
        lo, hi = e.min(0).values, e.max(0).values
        e = (e - lo) / (hi - lo)
    idx = rows.unsqueeze(1).expand(-1, heads)
    m = torch.full((n, heads), float("-inf"), device=X.device).scatter_reduce(0, idx, e, "amax", include_self=True)
    ex = torch.exp(e - m[rows])
    s = torch.zeros(n, heads, device=X.device).index_add(0, rows, ex)
    p = ex / s[rows]
    return torch.zeros(n, heads, dim, device=X.device).index_add(0, rows, p.unsqueeze(-1) * Xt[cols])


def gatv2_true_fp32(X, W0, W1, a, rows, cols, n, heads, dim, slope=0.2):
    """GATv2 as its paper defines it, in FP32 message passing: [N, H, D]."""
    X0 = (X @ W0.t()).view(n, heads, dim)
    X1 = (X @ W1.t()).view(n, heads, dim)
    z = (F.leaky_relu(X0[rows] + X1[cols], slope) * a.view(1, heads, dim)).sum(-1)   # [E, H]
    idx = rows.unsqueeze(1).expand(-1, heads)
    m = torch.full((n, heads), float("-inf"), device=X.device).scatter_reduce(0, idx, z, "amax", include_self=True)
    ex = torch.exp(z - m[rows])
    s = torch.zeros(n, heads, device=X.device).index_add(0, rows, ex)
    p = ex / s[rows]
    return torch.zeros(n, heads, dim, device=X.device).index_add(0, rows, p.unsqueeze(-1) * X1[cols])
