"""SDDMM, attention terms, edge-map SpMM and row sums on the GPU: the
reference tests' known answers (small integers, exact) and random graphs vs
the FP64 oracle."""
import numpy as np
import pytest
import torch

import att_amd.compat as compat
from att_amd import ops
from oracle import dense as O
from tests.gpu_util import TOY_COL, TOY_ROW, csr_rows, rand_csr

pytestmark = pytest.mark.gpu
M = compat.MODULES


def toy_blocks(cuda, window=8, wide=16, ps=2):
    out = M["ATT_Block"].blockProcess_sddmm_balance_gnn_trans(torch.from_numpy(TOY_ROW), torch.from_numpy(TOY_COL),
                                                              window, wide, ps)
    return [t.to(cuda) for t in out]


def test_sddmm_known_answer(cuda):
    """benchmark-v3/sddmm-benchmark/test/sddmm_fp16/test_sddmm_fp16.py:24-69, every edge checked."""
    rp, col, emap, emap_t, win, at = toy_blocks(cuda)
    for prec, dt in (("fp16", torch.float16), ("tf32", torch.float32)):
        rhs = torch.randint(1, 3, (4, 30, 125), generator=torch.Generator().manual_seed(0)).to(dt)
        res = getattr(M["ATT_SDDMM_v3"], f"sddmm_{prec}_h_nnz")(rp, col, emap, win, rhs.to(cuda), rhs.to(cuda),
                                                                 torch.tensor(1.0), 38, wide=16)[0]
        ref = O.sddmm(TOY_ROW[:31], TOY_COL[:TOY_ROW[30]], rhs.double().numpy(), rhs.double().numpy())
        got = res.cpu().double().numpy()[:, :TOY_ROW[30]]
        assert np.array_equal(got, ref), prec
        assert res[0][0].item() == torch.dot(rhs[0][0].double(), rhs[0][1].double()).item()


def test_att_known_answer_v4_layout(cuda):
    """test_sddmm_fp16_att.py:32-75: X in [N, H, D], a0 = a1 = ones."""
    rp, col, emap, emap_t, win, at = toy_blocks(cuda)
    rhs = torch.randint(1, 3, (30, 4, 126), generator=torch.Generator().manual_seed(1)).half()
    w = torch.ones(4, 1, 126, dtype=torch.float16)
    res = M["ATT_SDDMM_v4"].att_fp16_h_nnz(rp, col, emap, win, rhs.to(cuda), w.to(cuda), w.to(cuda),
                                           torch.tensor(1.0), 38, wide=16)[0].float().cpu()
    X = rhs.permute(1, 0, 2).double().numpy()
    ref = O.att(TOY_ROW[:31], TOY_COL[:TOY_ROW[30]], X, w.double().numpy(), w.double().numpy())
    assert np.array_equal(res.numpy()[:, :TOY_ROW[30]], ref)
    assert res[0][0] == (rhs[0][0].double() @ w[0][0].double() + rhs[1][0].double() @ w[0][0].double())


def test_attv2_and_ones_known_answers(cuda):
    rp, col, emap, emap_t, win, at = toy_blocks(cuda)
    g = torch.Generator().manual_seed(2)
    X0 = torch.randint(1, 3, (4, 30, 126), generator=g).half()
    X1 = torch.randint(1, 3, (4, 30, 126), generator=g).half()
    a = torch.ones(4, 1, 126, dtype=torch.float16)
    res = M["ATT_SDDMM_v3"].attv2_fp16_h_nnz(rp, col, emap, win, X0.to(cuda), X1.to(cuda), a.to(cuda),
                                             torch.tensor(1.0), 38, wide=16)[0].cpu().double().numpy()
    ref = O.attv2(TOY_ROW[:31], TOY_COL[:TOY_ROW[30]], X0.double().numpy(), X1.double().numpy(), a.double().numpy())
    assert np.array_equal(res[:, :TOY_ROW[30]], ref)
    gv = torch.randint(1, 5, (30, 4, 1), generator=g).half()
    res = M["ATT_SDDMM_v3"].sddmm_fp16_ones_h_nnz(rp, col, emap, win, gv.to(cuda), torch.tensor(1.0), 38,
                                                  wide=16)[0].cpu()
    assert res[0][0] == gv[0][0] and res[1][6] == gv[1][1]       # test_sddmm_fp16_ones.py:62-68
    ref = O.ones(TOY_ROW[:31], gv[:, :, 0].double().numpy())
    assert np.array_equal(res.double().numpy()[:, :TOY_ROW[30]], ref)


@pytest.mark.parametrize("layout", ["h_m", "m_h"])
def test_spmm_edge_map_known_answer(cuda, layout):
    """test_spmm_h_m.py / test_spmm_m_h.py: ones values, integer X, heads 1 and 3."""
    rp, col, emap, emap_t, win, at = toy_blocks(cuda, 8, 8, 32)
    for head, dimN in ((1, 66), (3, 65), (4, 2)):
        vals = torch.ones(head, 38, dtype=torch.float16)
        rhs = torch.randint(1, 3, (head, 30, dimN), generator=torch.Generator().manual_seed(head)).half()
        res = getattr(M["ATT_SpMM_v3"], f"spmm_fp16_{layout}")(rp, col, emap, vals.to(cuda), win, at,
                                                                rhs.to(cuda))[0].cpu().double().numpy()
        for h in range(head):
            ref = O.spmm(TOY_ROW, TOY_COL, np.ones(38), rhs[h].double().numpy(), n_rows=30)
            got = res[h] if layout == "h_m" else res[:, h]
            assert np.array_equal(got, ref)


def test_rowsum_known_answer(cuda):
    rp, col, emap, emap_t, win, at = toy_blocks(cuda, 8, 8, 32)
    vals = torch.arange(3 * 38, dtype=torch.float32).reshape(3, 38) % 7 + 1
    s = M["ATT_SpMM_v3"].spmm_fp16_h_m_rowsum(rp, emap, vals.half().to(cuda), win, at, 30)[0].cpu()
    ref = O.rowsum(TOY_ROW[:31], vals[:, :TOY_ROW[30]].double().numpy())
    assert s.shape == (3, 30, 1) and np.array_equal(s[:, :, 0].double().numpy(), ref)
    s = M["ATT_SpMM_v3"].spmm_fp16_m_h_rowsum(rp, emap, vals.half().to(cuda), win, at, 30)[0].cpu()
    assert s.shape == (30, 3, 1) and np.array_equal(s[:, :, 0].double().numpy().T, ref)
    # the transposed map sums columns (symmetric pattern not needed: absent transposes are skipped)
    s1 = M["ATT_SpMM_v1"].spmm_fp16_h_m_rowsum(rp, emap, vals[:1].half().to(cuda), win, at, 30)[0]
    assert s1.shape == (30, 1)


@pytest.mark.parametrize("window,wide,ps", [(8, 8, 32), (16, 32, 2), (8, 16, 1)])
@pytest.mark.parametrize("heads,dim", [(1, 64), (4, 64), (2, 40), (3, 13)])
def test_sddmm_random_vs_oracle(cuda, window, wide, ps, heads, dim):
    rng = np.random.default_rng(heads * 100 + dim + window)
    n = 300
    rp, col = rand_csr(rng, n, 20, hub_rows=2)
    b = ops.blocks_build(rp, col, "map", None, torch.int32, window, wide, ps, emit_empty=False)
    s = [b[k].to(cuda) for k in ("rp", "col", "val", "window")]
    A = rng.standard_normal((heads, n, dim)).astype(np.float32)
    Bm = rng.standard_normal((heads, n, dim)).astype(np.float32)
    covered = rp[n // window * window]
    ref = O.sddmm(rp, col, A, Bm)[:, :covered]
    for dt, mode, tol in ((torch.float32, "bf16x3", 1e-5), (torch.float32, "f32", 2e-6),
                          (torch.bfloat16, "native", 2e-6), (torch.float16, "native", 2e-6)):
        At = torch.from_numpy(A).to(dt)
        Bt = torch.from_numpy(Bm).to(dt)
        out = ops.sddmm_block(*s, At.to(cuda), Bt.to(cuda), col.size, window, wide, out_dtype=torch.float32,
                              mode=mode)
        r = ref if dt == torch.float32 else O.sddmm(rp, col, At.double().numpy(), Bt.double().numpy())[:, :covered]
        err = O.rel_fro(out.cpu().numpy()[:, :covered], r)
        assert err < tol, (dt, mode, err)


@pytest.mark.parametrize("heads,dim", [(1, 64), (4, 64), (8, 32)])
def test_edge_map_spmm_random_vs_oracle(cuda, heads, dim):
    rng = np.random.default_rng(heads + dim)
    n = 500
    rp, col = rand_csr(rng, n, 30, hub_rows=3)
    b = ops.blocks_build(rp, col, "map", None, torch.int32, 16, 32, 2, emit_empty=True)
    s = [b[k].to(cuda) for k in ("rp", "col", "val")]
    wa = [b[k].to(cuda) for k in ("window", "atomic")]
    ev = rng.random((heads, col.size)).astype(np.float32)
    X = rng.uniform(-1, 1, (heads, n, dim)).astype(np.float32)
    ref = O.spmm_heads(rp, col, ev, X)
    ref[:, n // 16 * 16:] = 0
    for dt, mode, tol in ((torch.float32, "bf16x3", 1e-5), (torch.bfloat16, "native", 1e-2),
                          (torch.float16, "native", 2e-3)):
        y = ops.spmm_edge_map(*s, torch.from_numpy(ev).to(dt).to(cuda), *wa, torch.from_numpy(X).to(dt).to(cuda),
                              "hmd", 16, 32, mode=mode)
        assert O.rel_fro(y.cpu().numpy(), ref) < tol, (dt, mode)
    rs = ops.rowsum_edge_map(b["rp"].to(cuda), b["val"].to(cuda), b["window"].to(cuda),
                             torch.from_numpy(ev).to(cuda), n, "hmd", 16, 32)
    rref = O.rowsum(rp, ev)
    rref[:, n // 16 * 16:] = 0
    assert O.rel_fro(rs.cpu().numpy(), rref) < 1e-6


@pytest.mark.parametrize("heads,dim", [(1, 64), (4, 64), (8, 32), (2, 40), (3, 13)])
@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float16, torch.float32])
def test_sddmm_edge_list_equals_dense_map(cuda, heads, dim, dt):
    """The compact edge-list SDDMM writes exactly the dense-map SDDMM's values."""
    rng = np.random.default_rng(heads * 7 + dim)
    n = 700
    rp, col = rand_csr(rng, n, 40, hub_rows=3)
    for window in (16, 8):
        b = ops.blocks_build(rp, col, "map", None, torch.int32, window, 32, 2, emit_empty=True, compact=True)
        d = {k: b[k].to(cuda) for k in ("rp", "col", "val", "window", "block_base", "values", "positions")}
        A = torch.from_numpy(rng.standard_normal((n, heads, dim)).astype(np.float32)).to(dt).to(cuda)
        B = torch.from_numpy(rng.standard_normal((n, heads, dim)).astype(np.float32)).to(dt).to(cuda)
        ref = ops.sddmm_block(d["rp"], d["col"], d["val"], d["window"], A, B, col.size, window, 32, layout="mhd",
                              out_dtype=torch.float32)
        got = ops.sddmm_edges(d["rp"], d["col"], d["block_base"], d["values"], d["positions"], d["window"], A, B,
                              col.size, window, layout="mhd", out_dtype=torch.float32)
        assert torch.equal(ref, got), (window, dt)
