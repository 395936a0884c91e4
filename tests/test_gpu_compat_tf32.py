"""The reference's TF32 edge-map callers through the drop-in modules with no
``wide=`` keyword: the reference's TF32 SpMM tests
(benchmark-v3/spmm-benchmark/test/spmm_tf32/test_spmm_{h_m,m_h,h_m_ones,m_h_ones}.py,
maps built 4 wide) with their exact calls, every output entry checked instead
of the first and last column, and the op chain of the reference's TF32
multi-head GAT layer (end2end_head/gat_final/mgat32_csr/mgat_conv.py:14-149,
dataset mdataset_tf32.py:39-60) against a float64 evaluation."""
import numpy as np
import pytest
import torch

import att_amd.compat as compat
from att_amd import graphs as G
from oracle import dense as O
from tests.gpu_util import TOY_COL, TOY_ROW

pytestmark = pytest.mark.gpu
M = compat.MODULES


def _scipy_check(row, col, rhs, n):
    """The tests' check(): CSR (ones) of the first n rows times rhs, float64."""
    return O.spmm(row[: n + 1], col[: row[n]], np.ones(row[n]), rhs, n_rows=n)


def _toy_4wide(cuda):
    out = M["ATT_Block"].blockProcess_sddmm_balance_gnn(torch.from_numpy(TOY_ROW), torch.from_numpy(TOY_COL), 8, 4, 32)
    return [t.to(cuda) for t in out]


@pytest.mark.parametrize("layout,head,dimN", [("h_m", 3, 65), ("m_h", 3, 65), ("h_m", 1, 65), ("m_h", 2, 8)])
def test_spmm_tf32_reference_calls(cuda, layout, head, dimN):
    """test_spmm_h_m.py:20-66 / test_spmm_m_h.py:20-84 (tf32): ones values,
    rhs of 30 rows with entries in {1, 2}; exact."""
    rp, col, degrees, window, automic = _toy_4wide(cuda)
    csr_values = torch.ones(head, TOY_COL.size, dtype=torch.float32)
    rhs = torch.randint(low=1, high=3, size=(head, 30, dimN), generator=torch.Generator().manual_seed(dimN)).float()
    result = getattr(M["ATT_SpMM_v3"], f"spmm_tf32_{layout}")(rp, col, degrees, csr_values.to(cuda), window, automic,
                                                              rhs.to(cuda))[0].cpu().double().numpy()
    for h in range(head):
        res = _scipy_check(TOY_ROW, TOY_COL, rhs[h].double().numpy(), 30)
        got = result[h] if layout == "h_m" else result[:, h]
        assert np.array_equal(got, res), h


@pytest.mark.parametrize("layout", ["h_m", "m_h"])
def test_spmm_tf32_rowsum_reference_calls(cuda, layout):
    """test_spmm_h_m_ones.py:20-70 / test_spmm_m_h_ones.py: 8 heads of ones;
    the row sums equal CSR . ones (the row counts)."""
    rp, col, degrees, window, automic = _toy_4wide(cuda)
    head = 8
    csr_values = torch.ones(head, TOY_COL.size, dtype=torch.float32)
    rhs = torch.ones(head, 30, 1)
    result = getattr(M["ATT_SpMM_v3"], f"spmm_tf32_{layout}_rowsum")(rp, degrees, csr_values.to(cuda), window,
                                                                     automic, rhs.size(1))[0].cpu().double().numpy()
    res = _scipy_check(TOY_ROW, TOY_COL, rhs[0].double().numpy(), 30)
    assert result.shape == ((head, 30, 1) if layout == "h_m" else (30, head, 1))
    for h in range(head):
        got = result[h] if layout == "h_m" else result[:, h]
        assert np.array_equal(got, res), h


def test_fp16_ops_reject_a_4_wide_structure(cuda):
    """A 4-wide map read by the fp16 ops (8 wide) is refused, not read at the
    wrong slots: window 0 of the toy graph has 12 condensed columns."""
    rp, col, degrees, window, automic = _toy_4wide(cuda)
    vals = torch.ones(1, TOY_COL.size, dtype=torch.float16, device=cuda)
    rhs = torch.ones(1, 30, 16, dtype=torch.float16, device=cuda)
    with pytest.raises(ValueError, match="wide=8"):
        M["ATT_SpMM_v3"].spmm_fp16_h_m(rp, col, degrees, vals, window, automic, rhs)
    with pytest.raises(ValueError, match="wide=8"):
        M["ATT_SDDMM_v3"].sddmm_fp16_h_nnz(rp, col, degrees, window, rhs, rhs, torch.tensor(1.0), 38)


def _padded_graph(n, deg, seed):
    """mdataset_tf32.py:39-60: rows padded to n + 8 - n % 8 (empty rows), CSR of the edge list."""
    g = G.chung_lu(n, deg, 2.1, seed=seed)
    n_pad = n + 8 - n % 8
    row = np.concatenate([g.indptr, np.full(n_pad - n, g.indptr[-1])]).astype(np.int32)
    return g, row, g.indices.astype(np.int32)


@pytest.mark.parametrize("heads,dim", [(1, 64), (2, 32), (4, 16)])
def test_tf32_gat_layer_op_chain_vs_fp64(cuda, heads, dim):
    """The TF32 GATConv_multi forward (att -> LeakyReLU -> exp -> rowsum ->
    m_h aggregation -> divide, mgat_conv.py:171-197) and the ops of its
    backward (h_m aggregation over map_T, SDDMM, ones broadcast, row and
    column sums, :33-149) through ATT_Block_v3 / ATT_SpMM_v3 / ATT_SDDMM_v3
    on a 4-wide map, each against float64 (bf16x3 arithmetic: <= 1e-5)."""
    g, row, col = _padded_graph(600, 8, seed=heads)
    n = g.n
    rp, col_i, degrees, degrees_trans, window, automic = (
        t.to(cuda) for t in M["ATT_Block_v3"].blockProcess_sddmm_balance_gnn_trans(
            torch.from_numpy(row), torch.from_numpy(col), 8, 4, 32))
    mx = torch.max(rp[2::2] - rp[:-2:2]) / 16
    nnz = col.size
    gen = torch.Generator().manual_seed(dim)
    Xp = torch.randn(heads, n, dim, generator=gen)
    a0 = torch.randn(heads, 1, dim, generator=gen) * 0.2
    a1 = torch.randn(heads, 1, dim, generator=gen) * 0.2
    S, A = M["ATT_SpMM_v3"], M["ATT_SDDMM_v3"]
    Xd, a0d, a1d = Xp.to(cuda), a0.to(cuda), a1.to(cuda)

    att = A.att_tf32_h_nnz(rp, col_i, degrees, window, Xd, a0d, a1d, mx, nnz)[0]
    att = torch.exp(torch.nn.functional.leaky_relu(att, 0.2))
    rows_sum = S.spmm_tf32_m_h_rowsum(rp, degrees, att, window, automic, n)[0]
    h = S.spmm_tf32_m_h(rp, col_i, degrees, att, window, automic, Xd)[0]
    h = h.div(rows_sum)

    X64, a064, a164 = Xp.double().numpy(), a0.double().numpy(), a1.double().numpy()
    e = O.att(g.indptr, g.indices, X64, a064, a164)
    p = np.exp(np.where(e > 0, e, 0.2 * e))
    rs = O.rowsum(g.indptr, p)
    want = O.spmm_heads(g.indptr, g.indices, p, X64) / rs[:, :, None]
    assert O.rel_fro(h.permute(1, 0, 2).cpu().double().numpy(), want) < 1e-5
    assert O.rel_fro(att.cpu().double().numpy(), p) < 1e-5

    # backward ops (MGATSpmm_multi.backward, MGATSpmm1_multi.backward, MGATFunction_multi.backward)
    gh = torch.randn(n, heads, dim, generator=gen)
    gd = gh.permute(1, 0, 2).contiguous().to(cuda)
    dX = S.spmm_tf32_h_m(rp, col_i, degrees_trans, att, window, automic, gd)[0]
    d_att = A.sddmm_tf32_h_nnz(rp, col_i, degrees, window, gd, Xd, mx, nnz)[0]
    grs = torch.randn(n, heads, 1, generator=gen)
    d_att1 = A.sddmm_tf32_ones_h_nnz(rp, col_i, degrees, window, grs.to(cuda), mx, nnz)[0]
    ag = torch.randn(heads, nnz, generator=gen)
    a0_tmp = S.spmm_tf32_h_m_rowsum(rp, degrees_trans, ag.to(cuda), window, automic, n)[0]
    a1_tmp = S.spmm_tf32_h_m_rowsum(rp, degrees, ag.to(cuda), window, automic, n)[0]

    G64 = gh.permute(1, 0, 2).double().numpy()
    rows = np.repeat(np.arange(n), np.diff(g.indptr))
    order = np.lexsort((rows, g.indices))       # edges by (column, row): k-th is edge (j, i) of A's k-th (i, j)
    pT = p[:, order]                            # value of edge (j, i) at the position of edge (i, j)
    want_dX = O.spmm_heads(g.indptr, g.indices, pT, G64)   # symmetric pattern: A^T's CSR is A's
    assert O.rel_fro(dX.cpu().double().numpy(), want_dX) < 1e-5
    assert O.rel_fro(d_att.cpu().double().numpy(), O.sddmm(g.indptr, g.indices, G64, X64)) < 1e-5
    assert np.array_equal(d_att1.cpu().double().numpy(), O.ones(g.indptr, grs[:, :, 0].double().numpy()))
    ag64 = ag.double().numpy()
    assert O.rel_fro(a1_tmp[:, :, 0].cpu().double().numpy(), O.rowsum(g.indptr, ag64)) < 1e-6
    assert O.rel_fro(a0_tmp[:, :, 0].cpu().double().numpy(), O.rowsum(g.indptr, ag64[:, order])) < 1e-6
