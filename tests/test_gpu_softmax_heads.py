"""All-heads edge softmax, fused backward row/column sums and the vector node
scores on the GPU.

The all-heads kernels (1, 2, 4 or 8 heads, one 16-lane group per row) repeat
the per-(row, head) kernels' arithmetic step for step, so their results must be
bitwise equal to them: with 3 heads the per-(row, head) kernels run, and their
first two heads must equal a 2-head run on the same scores (the dropout hash
depends on (seed, head, edge) only).  Node-major and head-major scores must
give the same bits, the fused backward sums must equal edge_softmax_bwd +
gather_rowsum bit for bit, and everything is checked against FP64 on the host.
"""
import numpy as np
import pytest
import torch

from att_amd import ops
from tests.gpu_util import csr_rows, rand_csr

pytestmark = pytest.mark.gpu


def _graph(cuda, n, max_deg, hub_rows, seed, symmetric):
    rng = np.random.default_rng(seed)
    rp, col = rand_csr(rng, n, max_deg, hub_rows=hub_rows)
    if symmetric:
        rows = csr_rows(rp)
        keys = np.unique(np.concatenate([rows * n + col, col.astype(np.int64) * n + rows]))
        rp = np.zeros(n + 1, np.int32)
        rp[1:] = np.cumsum(np.bincount(keys // n, minlength=n))
        col = (keys % n).astype(np.int32)
    rowptr = torch.from_numpy(rp).to(cuda)
    colt = torch.from_numpy(col).to(cuda)
    t_rowptr, _, t_perm = ops.csr_transpose_device(rowptr, colt, n)
    return rp, col, rowptr, colt, t_rowptr, t_perm


def _softmax_fp64(rp, col, s0, s1, slope):
    """p [H, nnz] and lse [H, N] in FP64 (no dropout); s0 / s1 [H, N]."""
    rows = csr_rows(rp)
    z = s0[:, rows] + s1[:, col]
    v = np.where(z > 0, z, z * slope)
    heads, n = s0.shape
    lse = np.zeros((heads, n))
    p = np.zeros_like(v)
    for i in range(n):
        b, e = rp[i], rp[i + 1]
        if e > b:
            m = v[:, b:e].max(1, keepdims=True)
            lse[:, i] = (m + np.log(np.exp(v[:, b:e] - m).sum(1, keepdims=True)))[:, 0]
            p[:, b:e] = np.exp(v[:, b:e] - lse[:, i:i + 1])
    return p, lse


@pytest.mark.parametrize("use_long", [False, True])
@pytest.mark.parametrize("symmetric", [True, False])
def test_heads_kernels_equal_per_head_kernels(cuda, symmetric, use_long):
    n = 700
    rp, col, rowptr, colt, t_rowptr, t_perm = _graph(cuda, n, 20, 3, seed=5, symmetric=symmetric)
    lr = ops.long_rows(rowptr) if use_long else None
    tlr = ops.long_rows(t_rowptr) if use_long else None
    if use_long:
        assert lr.numel() == 3     # the hub rows (350+ entries)
    nnz = col.size
    g = torch.Generator(device="cpu").manual_seed(0)
    s0 = torch.randn(3, n, generator=g).to(cuda)
    s1 = torch.randn(3, n, generator=g).to(cuda)
    dp = torch.randn(3, nnz, generator=g).to(cuda)
    for drop in (0.0, 0.3):
        p3, lse3 = ops.edge_softmax_fwd(rowptr, colt, s0, s1, 0.2, drop, 9, p_dtype=torch.float32, long_rows=lr)
        p2, lse2 = ops.edge_softmax_fwd(rowptr, colt, s0[:2].contiguous(), s1[:2].contiguous(), 0.2, drop, 9,
                                        p_dtype=torch.float32, long_rows=lr)
        assert torch.equal(p3[:2], p2) and torch.equal(lse3[:2], lse2)
        dz3, ds03 = ops.edge_softmax_bwd(rowptr, colt, s0, s1, lse3, dp, 0.2, drop, 9, long_rows=lr)
        dz2, ds02 = ops.edge_softmax_bwd(rowptr, colt, s0[:2].contiguous(), s1[:2].contiguous(), lse2,
                                         dp[:2].contiguous(), 0.2, drop, 9, long_rows=lr)
        assert torch.equal(dz3[:2], dz2) and torch.equal(ds03[:2], ds02)
        c3 = ops.gather_rowsum(t_rowptr, t_perm, dz3, n, long_rows=tlr)
        c2 = ops.gather_rowsum(t_rowptr, t_perm, dz2, n, long_rows=tlr)
        assert torch.equal(c3[:2], c2)


def test_long_rows_list_and_workgroup_path(cuda):
    """ops.long_rows lists the rows above 256 entries; with the list, the short
    rows' results are unchanged bit for bit, the long rows' agree to FP32
    rounding, and extra short or repeated rows in the list change nothing."""
    n = 2000
    rp, col, rowptr, colt, t_rowptr, t_perm = _graph(cuda, n, 12, 6, seed=11, symmetric=True)
    deg = np.diff(rp)
    lr = ops.long_rows(rowptr)
    assert lr.dtype == torch.int32 and np.array_equal(lr.cpu().numpy(), np.flatnonzero(deg > 256))
    assert lr.numel() >= 6
    heads, nnz = 4, col.size
    g = torch.Generator(device="cpu").manual_seed(1)
    s0 = torch.randn(n, heads, generator=g).to(cuda)
    s1 = torch.randn(n, heads, generator=g).to(cuda)
    dp = torch.randn(heads, nnz, generator=g).to(cuda)
    tlr = ops.long_rows(t_rowptr)
    padded = torch.cat([lr, lr[:2], torch.tensor([0, 1, 2], dtype=torch.int32, device=cuda)])
    short = torch.from_numpy(deg <= 256).to(cuda)
    short_e = torch.from_numpy(np.repeat(deg <= 256, deg)).to(cuda)
    for drop in (0.0, 0.2):
        p0, l0 = ops.edge_softmax_fwd(rowptr, colt, s0, s1, 0.2, drop, 4, s_layout="nh")
        p1, l1 = ops.edge_softmax_fwd(rowptr, colt, s0, s1, 0.2, drop, 4, s_layout="nh", long_rows=lr)
        p2, l2 = ops.edge_softmax_fwd(rowptr, colt, s0, s1, 0.2, drop, 4, s_layout="nh", long_rows=padded)
        assert torch.equal(p1, p2) and torch.equal(l1, l2)
        assert torch.equal(p0[:, short_e], p1[:, short_e]) and torch.equal(l0[:, short], l1[:, short])
        assert torch.allclose(p0, p1, rtol=1e-5, atol=1e-7) and torch.allclose(l0, l1, rtol=1e-5, atol=1e-5)
        d0, r0 = ops.edge_softmax_bwd(rowptr, colt, s0, s1, l1, dp, 0.2, drop, 4, s_layout="nh")
        d1, r1 = ops.edge_softmax_bwd(rowptr, colt, s0, s1, l1, dp, 0.2, drop, 4, s_layout="nh", long_rows=lr)
        assert torch.equal(d0[:, short_e], d1[:, short_e]) and torch.equal(r0[:, short], r1[:, short])
        assert torch.allclose(d0, d1, rtol=1e-4, atol=1e-6) and torch.allclose(r0, r1, rtol=1e-4, atol=1e-5)
        c1 = ops.gather_rowsum(t_rowptr, t_perm, d1, n, long_rows=tlr)
        f0, f1 = ops.edge_softmax_bwd_sums(rowptr, colt, s0, s1, l1, dp, 0.2, drop, 4, t_rowptr, t_perm,
                                           s_layout="nh", long_rows=lr, t_long_rows=tlr)
        assert torch.equal(f0, r1) and torch.equal(f1, c1)
        c0 = ops.gather_rowsum(t_rowptr, t_perm, d1, n)
        assert torch.allclose(c0, c1, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("heads", [1, 2, 3, 4, 8])
@pytest.mark.parametrize("dp_dtype", [torch.float32, torch.bfloat16])
def test_layouts_and_fused_sums(cuda, heads, dp_dtype):
    n = 900
    rp, col, rowptr, colt, t_rowptr, t_perm = _graph(cuda, n, 24, 2, seed=heads, symmetric=heads % 2 == 0)
    lr, tlr = ops.long_rows(rowptr), ops.long_rows(t_rowptr)
    nnz = col.size
    g = torch.Generator(device="cpu").manual_seed(heads)
    s0 = torch.randn(heads, n, generator=g).to(cuda)
    s1 = torch.randn(heads, n, generator=g).to(cuda)
    dp = torch.randn(heads, nnz, generator=g).to(dp_dtype).to(cuda)
    s0n, s1n = s0.t().contiguous(), s1.t().contiguous()
    for drop in (0.0, 0.25):
        p, lse = ops.edge_softmax_fwd(rowptr, colt, s0, s1, 0.2, drop, 3, p_dtype=torch.bfloat16)
        pn, lsen = ops.edge_softmax_fwd(rowptr, colt, s0n, s1n, 0.2, drop, 3, p_dtype=torch.bfloat16, s_layout="nh")
        assert torch.equal(p, pn) and torch.equal(lse, lsen)
        dz, ds0 = ops.edge_softmax_bwd(rowptr, colt, s0, s1, lse, dp, 0.2, drop, 3)
        dzn, ds0n = ops.edge_softmax_bwd(rowptr, colt, s0n, s1n, lse, dp, 0.2, drop, 3, s_layout="nh")
        assert torch.equal(dz, dzn) and torch.equal(ds0, ds0n)
        ds1 = ops.gather_rowsum(t_rowptr, t_perm, dz, n)
        for layout, a, b in (("hn", s0, s1), ("nh", s0n, s1n)):
            f0, f1 = ops.edge_softmax_bwd_sums(rowptr, colt, a, b, lse, dp, 0.2, drop, 3, t_rowptr, t_perm,
                                               s_layout=layout)
            assert torch.equal(f0, ds0) and torch.equal(f1, ds1), layout
        # with the long-row lists: head-major and node-major scores, split and fused backward, same bits
        pl, lsel = ops.edge_softmax_fwd(rowptr, colt, s0, s1, 0.2, drop, 3, p_dtype=torch.bfloat16, long_rows=lr)
        pln, lseln = ops.edge_softmax_fwd(rowptr, colt, s0n, s1n, 0.2, drop, 3, p_dtype=torch.bfloat16,
                                          s_layout="nh", long_rows=lr)
        assert torch.equal(pl, pln) and torch.equal(lsel, lseln)
        dzl, ds0l = ops.edge_softmax_bwd(rowptr, colt, s0, s1, lsel, dp, 0.2, drop, 3, long_rows=lr)
        ds1l = ops.gather_rowsum(t_rowptr, t_perm, dzl, n, long_rows=tlr)
        for layout, a, b in (("hn", s0, s1), ("nh", s0n, s1n)):
            f0, f1 = ops.edge_softmax_bwd_sums(rowptr, colt, a, b, lsel, dp, 0.2, drop, 3, t_rowptr, t_perm,
                                               s_layout=layout, long_rows=lr, t_long_rows=tlr)
            assert torch.equal(f0, ds0l) and torch.equal(f1, ds1l), layout
    # FP64 on the host (no dropout), through the long-row path
    p, lse = ops.edge_softmax_fwd(rowptr, colt, s0n, s1n, 0.2, 0.0, 0, p_dtype=torch.float32, s_layout="nh",
                                  long_rows=lr)
    pr, lse64 = _softmax_fp64(rp, col, s0.double().cpu().numpy(), s1.double().cpu().numpy(), 0.2)
    assert np.abs(p.cpu().numpy() - pr).max() < 1e-5
    assert np.abs(lse.cpu().numpy() - lse64).max() < 1e-4
    ds0, ds1 = ops.edge_softmax_bwd_sums(rowptr, colt, s0n, s1n, lse, dp, 0.2, 0.0, 0, t_rowptr, t_perm,
                                         s_layout="nh", long_rows=lr, t_long_rows=tlr)
    g64 = dp.double().cpu().numpy()
    rows = csr_rows(rp)
    z = s0.double().cpu().numpy()[:, rows] + s1.double().cpu().numpy()[:, col]
    S = np.zeros((heads, n))
    np.add.at(S.T, rows, (pr * g64).T)
    dz = pr * (g64 - S[:, rows]) * np.where(z > 0, 1.0, 0.2)
    r0 = np.zeros((heads, n))
    r1 = np.zeros((heads, n))
    np.add.at(r0.T, rows, dz.T)
    np.add.at(r1.T, col, dz.T)
    scale = np.abs(dz).max() * 30
    assert np.abs(ds0.cpu().numpy() - r0).max() < 1e-5 * scale
    assert np.abs(ds1.cpu().numpy() - r1).max() < 1e-5 * scale


def test_fused_sums_empty_rows_and_bad_perm(cuda):
    """Rows without entries give zero sums; out-of-range transpose positions are skipped."""
    n = 64
    rp = np.zeros(n + 1, np.int32)
    rp[10:] = 3
    rp[40:] = 7
    col = np.array([5, 9, 40, 1, 2, 3, 9], np.int32)
    rowptr = torch.from_numpy(rp).cuda()
    colt = torch.from_numpy(col).cuda()
    t_rowptr, _, t_perm = ops.csr_transpose_device(rowptr, colt, n)
    s = torch.randn(n, 4, device=cuda)
    dp = torch.randn(4, col.size, device=cuda)
    p, lse = ops.edge_softmax_fwd(rowptr, colt, s, s, 0.2, 0.0, 0, s_layout="nh")
    ds0, ds1 = ops.edge_softmax_bwd_sums(rowptr, colt, s, s, lse, dp, 0.2, 0.0, 0, t_rowptr, t_perm, s_layout="nh")
    dz, r0 = ops.edge_softmax_bwd(rowptr, colt, s, s, lse, dp, 0.2, 0.0, 0, s_layout="nh")
    assert torch.equal(ds0, r0)
    empty = torch.from_numpy(np.diff(rp) == 0).cuda()
    assert bool((ds0[:, empty] == 0).all()) and bool((lse[:, empty] == 0).all())
    bad = t_perm.clone()
    bad[0] = -1
    bad[1] = 10 ** 9
    _, c = ops.edge_softmax_bwd_sums(rowptr, colt, s, s, lse, dp, 0.2, 0.0, 0, t_rowptr, bad, s_layout="nh")
    ref = ops.gather_rowsum(t_rowptr, bad, dz, n)
    assert torch.equal(c, ref)


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16, torch.float32])
@pytest.mark.parametrize("heads,dim", [(1, 64), (2, 40), (3, 16), (4, 64), (8, 8), (4, 12)])
def test_node_scores_vector_and_layouts(cuda, dtype, heads, dim):
    n = 1000
    g = torch.Generator(device="cpu").manual_seed(heads * 100 + dim)
    X = torch.randn(n, heads, dim, generator=g).to(dtype).to(cuda)
    a = torch.randn(heads, dim, generator=g).to(cuda)
    ref = (X.double() * a.double().view(1, heads, dim)).sum(-1).t().cpu()
    for vec in (a.to(dtype), a):
        s = ops.node_scores(X, vec, "mhd")
        sn = ops.node_scores(X, vec, "mhd", out_layout="nh")
        assert s.shape == (heads, n) and sn.shape == (n, heads)
        assert torch.equal(sn.t(), s)
        tol = 1e-5 if vec.dtype == torch.float32 and dtype == torch.float32 else 2e-2
        err = (s.double().cpu() - ref).abs().max() / ref.abs().max()
        assert err < tol, (vec.dtype, err)
    # [H, N, D] input
    Xh = X.permute(1, 0, 2).contiguous()
    assert torch.allclose(ops.node_scores(Xh, a.to(dtype), "hmd"), ops.node_scores(X, a.to(dtype), "mhd"),
                          rtol=1e-5, atol=1e-5)
