"""Differentiable SpMM / SDDMM primitives vs float64 autograd on the CPU
(non-symmetric patterns, several heads), in fp32 (bf16x3 kernels) and bf16."""
import numpy as np
import pytest
import torch

from att_amd import autograd as AG
from att_amd.layers import GATGraph
from tests.test_gpu_layers import small_graph

pytestmark = pytest.mark.gpu


def dense(g, vals):
    """[H, N, N] float64 from per-edge values [H, nnz]."""
    H = vals.shape[0]
    A = torch.zeros(H, g.n, g.n, dtype=torch.float64)
    r = torch.from_numpy(g.row_ids().astype(np.int64))
    c = torch.from_numpy(g.indices.astype(np.int64))
    for h in range(H):
        A[h, r, c] = vals[h]
    return A


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-4), (torch.bfloat16, 3e-2)])
@pytest.mark.parametrize("heads,dim", [(1, 32), (3, 16)])
def test_spmm_grad(cuda, dtype, tol, heads, dim):
    g = small_graph(260, 6, seed=heads, symmetric=False)
    graph = GATGraph(g, cuda)
    torch.manual_seed(0)
    vals = torch.rand(heads, g.nnz, device=cuda).to(dtype).requires_grad_(True)
    X = torch.randn(g.n, heads, dim, device=cuda).to(dtype).requires_grad_(True)
    Y = AG.spmm(vals, X, graph)
    G = torch.randn(Y.shape, device=cuda)
    (Y.float() * G).sum().backward()
    v64 = vals.detach().double().cpu().requires_grad_(True)
    X64 = X.detach().double().cpu().requires_grad_(True)
    Y64 = torch.einsum("hij,jhd->ihd", dense(g, v64), X64)
    (Y64 * G.double().cpu()).sum().backward()
    rel = lambda a, b: ((a.double().cpu() - b).norm() / b.norm()).item()
    assert rel(Y.detach(), Y64.detach()) < tol
    assert rel(X.grad, X64.grad) < tol
    assert rel(vals.grad, v64.grad) < tol


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-4), (torch.bfloat16, 3e-2)])
def test_sddmm_grad(cuda, dtype, tol):
    g = small_graph(240, 5, seed=4, symmetric=False)
    graph = GATGraph(g, cuda)
    heads, dim = 2, 24
    torch.manual_seed(1)
    A = torch.randn(g.n, heads, dim, device=cuda).to(dtype).requires_grad_(True)
    B = torch.randn(g.n, heads, dim, device=cuda).to(dtype).requires_grad_(True)
    out = AG.sddmm(A, B, graph)
    G = torch.randn(out.shape, device=cuda)
    (out * G).sum().backward()
    A64 = A.detach().double().cpu().requires_grad_(True)
    B64 = B.detach().double().cpu().requires_grad_(True)
    r = torch.from_numpy(g.row_ids().astype(np.int64))
    c = torch.from_numpy(g.indices.astype(np.int64))
    ref = (A64[r] * B64[c]).sum(-1).t()          # [H, nnz]
    (ref * G.double().cpu()).sum().backward()
    rel = lambda a, b: ((a.double().cpu() - b).norm() / b.norm()).item()
    assert rel(out.detach(), ref.detach()) < tol
    assert rel(A.grad, A64.grad) < tol
    assert rel(B.grad, B64.grad) < tol
