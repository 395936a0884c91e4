"""Host-side pieces of the layers, on the CPU: the split-K weight gradient and
the bias-free linear Function against torch autograd in float64."""
import torch
import torch.nn.functional as F

from att_amd.layers import LinearSplitK, weight_grad


def test_weight_grad_matches_mm():
    gen = torch.Generator().manual_seed(0)
    for n, a, b in [(10007, 37, 19), (4096, 8, 8), (100, 5, 3), (2 * 2048 * 64 + 5, 4, 6)]:
        X = torch.randn(n, a, generator=gen, dtype=torch.float64)
        dY = torch.randn(n, b, generator=gen, dtype=torch.float64)
        got = weight_grad(X, dY)
        assert got.dtype == torch.float64
        assert torch.allclose(got, X.t() @ dY, rtol=0, atol=1e-10)
    Xf = torch.randn(9000, 16, generator=gen)
    assert weight_grad(Xf, Xf).dtype == torch.float32


def test_linear_split_k_gradients():
    gen = torch.Generator().manual_seed(1)
    X = torch.randn(5000, 24, generator=gen, dtype=torch.float64, requires_grad=True)
    W = torch.randn(10, 24, generator=gen, dtype=torch.float64, requires_grad=True)
    g = torch.randn(5000, 10, generator=gen, dtype=torch.float64)
    (LinearSplitK.apply(X, W) * g).sum().backward()
    X2, W2 = X.detach().clone().requires_grad_(True), W.detach().clone().requires_grad_(True)
    (F.linear(X2, W2) * g).sum().backward()
    assert torch.allclose(X.grad, X2.grad, atol=1e-12)
    assert torch.allclose(W.grad, W2.grad, atol=1e-10)
