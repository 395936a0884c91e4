"""CPU: the float64 dense attention reference used by the GPU layer tests is
pinned by central finite differences on an N=512, degree ~8 Chung-Lu graph."""
import numpy as np
import torch

from tests.test_gpu_layers import dense_adj, gat_reference, small_graph


def test_reference_formula_pinned_by_finite_differences():
    """The float64 reference gradient equals central finite differences (N=512,
    degree ~8), so the GPU gradients compared with it above are exact gradients."""
    g = small_graph(512, 8, seed=11)
    A = dense_adj(g)
    heads, dim = 2, 4
    torch.manual_seed(0)
    X = torch.randn(g.n, 6, dtype=torch.float64)
    W = torch.randn(heads * dim, 6, dtype=torch.float64, requires_grad=True)
    a0 = torch.randn(heads, dim, dtype=torch.float64, requires_grad=True)
    a1 = torch.randn(heads, dim, dtype=torch.float64, requires_grad=True)
    go = torch.randn(g.n, heads, dim, dtype=torch.float64)
    f = lambda W_, a0_, a1_: (gat_reference(X, W_, None, a0_, a1_, A, heads, dim, 0.2) * go).sum()
    f(W, a0, a1).backward()
    eps = 1e-6
    rng = np.random.default_rng(0)
    for t, grad in ((W, W.grad), (a0, a0.grad), (a1, a1.grad)):
        for _ in range(4):
            idx = tuple(int(rng.integers(0, s)) for s in t.shape)
            with torch.no_grad():
                t[idx] += eps
                fp = f(W, a0, a1).item()
                t[idx] -= 2 * eps
                fm = f(W, a0, a1).item()
                t[idx] += eps
            fd = (fp - fm) / (2 * eps)
            assert abs(fd - grad[idx].item()) <= 1e-6 * max(1.0, abs(fd))


def test_gatv2_minmax_reference_pinned_by_finite_differences():
    """The float64 restatement of the multi-head GATv2 (global min-max logit
    map) has autograd gradients equal to central finite differences, W0's
    included (it reaches the loss only through the extreme edges)."""
    from tests.test_gpu_layers import gatv2_minmax_reference
    g = small_graph(512, 8, seed=13, symmetric=False)
    rows = torch.from_numpy(g.row_ids().astype(np.int64))
    cols = torch.from_numpy(g.indices.astype(np.int64))
    heads, dim = 2, 4
    torch.manual_seed(1)
    X = torch.randn(g.n, 6, dtype=torch.float64)
    W0 = torch.randn(heads * dim, 6, dtype=torch.float64, requires_grad=True)
    W1 = torch.randn(heads * dim, 6, dtype=torch.float64, requires_grad=True)
    a = torch.randn(heads, dim, dtype=torch.float64, requires_grad=True)
    go = torch.randn(g.n, heads, dim, dtype=torch.float64)
    f = lambda W0_, W1_, a_: (gatv2_minmax_reference(X, W0_, W1_, a_, rows, cols, g.n, heads, dim, 0.2)[0]
                              * go).sum()
    f(W0, W1, a).backward()
    assert W0.grad.norm() > 0
    eps = 1e-6
    rng = np.random.default_rng(1)
    for t, grad in ((W0, W0.grad), (W1, W1.grad), (a, a.grad)):
        for _ in range(4):
            idx = tuple(int(rng.integers(0, s)) for s in t.shape)
            with torch.no_grad():
                t[idx] += eps
                fp = f(W0, W1, a).item()
                t[idx] -= 2 * eps
                fm = f(W0, W1, a).item()
                t[idx] += eps
            fd = (fp - fm) / (2 * eps)
            assert abs(fd - grad[idx].item()) <= 1e-6 * max(1.0, abs(fd))
