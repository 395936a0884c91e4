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


def test_csr_segments_partition_rows():
    import numpy as np
    from att_amd.graphs import csr_segments
    rng = np.random.default_rng(3)
    deg = rng.integers(0, 5, 400)
    deg[[3, 4, 200]] = [100, 33, 64]            # long rows, and an exact multiple of the length
    deg[[7, 8, 9]] = 0
    indptr = np.concatenate([[0], np.cumsum(deg)])
    s = csr_segments(indptr, max_len=32)
    n_seg, sp, sr, ss, rs = s["n_seg"], s["seg_ptr"], s["seg_row"], s["seg_slot"], s["row_seg"]
    assert sp[0] == 0 and sp[-1] == indptr[-1] and np.all(np.diff(sp) > 0) and np.all(np.diff(sp) <= 32)
    for i in range(400):
        segs = range(rs[i], rs[i + 1])
        assert len(segs) == -(-deg[i] // 32)
        if len(segs):
            assert sp[segs[0]] == indptr[i] and sp[segs[-1] + 1] == indptr[i + 1]
            assert all(sr[k] == i for k in segs)
            assert all((ss[k] == -1) == (len(segs) == 1) for k in segs)
    slots = ss[ss >= 0]
    assert np.array_equal(slots, np.arange(s["n_slots"])) and s["n_slots"] == 4 + 2 + 2


def test_nll_loss_equals_torch():
    from att_amd.train import nll_loss
    gen = torch.Generator().manual_seed(5)
    logp = torch.log_softmax(torch.randn(1000, 7, generator=gen, dtype=torch.float64), 1)
    y = torch.randint(0, 7, (1000,), generator=gen)
    assert torch.allclose(nll_loss(logp, y).double(), F.nll_loss(logp, y), rtol=1e-6)
