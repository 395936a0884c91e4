"""Precision modes of the window-block SpMM against the FP64 oracle on a
band + random matrix (the precision-sweep workload at reduced size), and the
exact-factorisation invariant of the FP8 mode."""
import numpy as np
import pytest
import torch

from att_amd import graphs as G
from att_amd.precision import BlockMatrix
from oracle import dense as O

pytestmark = pytest.mark.gpu

# mode -> bound on the relative Frobenius error vs FP64
BOUNDS = {"f32": 1e-6, "bf16x3": 1e-5, "fp16": 2e-3, "bf16": 1e-2, "fp8": 8e-2}


@pytest.fixture(scope="module")
def band():
    g = G.band_random(1 << 14, 300_000, 2, seed=0)
    X = np.random.default_rng(1).uniform(-1, 1, (g.n, 128)).astype(np.float32)
    return g, X, O.spmm(g.indptr, g.indices, g.values, X)


@pytest.mark.parametrize("mode", list(BOUNDS))
def test_mode_error_vs_fp64(cuda, band, mode):
    g, X, ref = band
    A = BlockMatrix(g, g.values, cuda, mode)
    y = A.matmul(torch.from_numpy(X).to(cuda))
    err = O.rel_fro(y.cpu().numpy(), ref)
    assert err < BOUNDS[mode], (mode, err)


def test_fp8_factorisation_invariant(cuda, band):
    """Y = diag(r) (A' X') diag(c): the kernel result equals the FP64 product of
    the quantised factors up to FP32 accumulation."""
    g, X, _ = band
    A = BlockMatrix(g, g.values, cuda, "fp8")
    Xt = torch.from_numpy(X).to(cuda)
    Xq, c = A.quantize_x(Xt)
    y = A.matmul(None, Xq=Xq, col_scale=c).cpu().numpy()
    r = A.row_scale.cpu().numpy().astype(np.float64)
    vq = torch.from_numpy(g.values / r[g.row_ids()].astype(np.float32)).to(torch.float8_e4m3fn).double().numpy()
    Xd = Xq.double().cpu().numpy()
    ref = r[:, None] * O.spmm(g.indptr, g.indices, vq, Xd) * c.double().cpu().numpy()[None, :]
    assert O.rel_fro(y, ref) < 1e-5   # FP32 accumulation of cancelling sums


def test_modes_deterministic(cuda, band):
    g, X, _ = band
    Xt = torch.from_numpy(X).to(cuda)
    for mode in BOUNDS:
        A = BlockMatrix(g, g.values, cuda, mode, part_size=1)
        assert torch.equal(A.matmul(Xt), A.matmul(Xt))
