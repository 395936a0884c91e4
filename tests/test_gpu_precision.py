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
        A = BlockMatrix(g, g.values, cuda, mode, part_size=1, chunk_blocks=1)
        assert torch.equal(A.matmul(Xt), A.matmul(Xt))


def test_compact_tiles_bitwise_equal_dense(cuda):
    """The compact-tile SpMM multiplies exactly the same numbers as the dense one."""
    from att_amd import ops
    rng = np.random.default_rng(3)
    g = G.chung_lu(3000, 14, 2.1, seed=2)
    v = rng.standard_normal(g.nnz).astype(np.float32)
    for dtype, mode in ((torch.float16, None), (torch.bfloat16, None), (torch.float32, "bf16x3"),
                        (torch.float32, "f32")):
        for ps in (2, 512):
            p = g.padded(16)
            b = ops.blocks_build(p.rowptr32(), p.indices, "edge_f32", v, dtype, 16, 32, ps, emit_empty=True,
                                 compact=True)
            d = {k: b[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic", "masks", "block_base", "values")}
            for dim in (256, 48, 64, 20):
                X = torch.from_numpy(rng.uniform(-1, 1, (g.n, dim)).astype(np.float32)).to(dtype).to(cuda)
                y0 = ops.spmm_block(d["rp"], d["col"], d["val"], d["window"], d["atomic"], X, g.n, 16, 32, mode=mode,
                                    rows_covered=True)
                y1 = ops.spmm_compact(d["rp"], d["col"], d["masks"], d["block_base"], d["values"], d["window"],
                                      d["atomic"], X, g.n, 16, mode=mode, rows_covered=True)
                assert torch.equal(y0, y1), (dtype, mode, ps, dim)
