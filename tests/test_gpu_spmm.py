"""Window-block SpMM on the GPU against the FP64 oracle (all modes), the
reference's known-answer test, torch.sparse, and run-to-run determinism."""
import numpy as np
import pytest
import torch

import att_amd.compat as compat
from att_amd import graphs as G
from att_amd import ops
from att_amd.layers import GCNGraph
from oracle import dense as O
from tests.gpu_util import TOY_COL, TOY_ROW, csr_rows, rand_csr

pytestmark = pytest.mark.gpu


def test_reference_balance_known_answer(cuda):
    """gcn/SpMM/test/mgcn/test_balance.py:20-62 with partSize 1 (every window
    split), small-integer X so FP16 products are exact; all columns compared."""
    m = compat.MODULES
    dd = torch.ones(32, dtype=torch.float16)
    rp, col, val, win, at = m["Block_sr_bcrs"].blockProcess_fp16_balance(
        torch.from_numpy(TOY_ROW), torch.from_numpy(TOY_COL), dd, 8, 8, 1)
    for dimN in (16, 64, 66, 2):
        rhs = torch.randint(1, 3, (32, dimN), generator=torch.Generator().manual_seed(dimN)).half()
        y = m["SpMM_sr_bcrs"].forward_fp16_gnn(rp.to(cuda), col.to(cuda), val.to(cuda), win.to(cuda), at.to(cuda),
                                               rhs.to(cuda), 32, dimN, 30)[0]
        ref = O.spmm(TOY_ROW, TOY_COL, np.ones(38), rhs.double().numpy(), n_rows=30)
        assert y.shape == (30, dimN)
        assert np.array_equal(y.cpu().double().numpy(), ref)
        # forward_fp16_gnn_acc (benchmark.cpp:494): the same product under the second name
        y2 = m["SpMM_sr_bcrs"].forward_fp16_gnn_acc(rp.to(cuda), col.to(cuda), val.to(cuda), win.to(cuda),
                                                    at.to(cuda), rhs.to(cuda), 32, dimN, 30)[0]
        assert torch.equal(y, y2)


def test_reference_balance_known_answer_tf32(cuda):
    """gcn/SpMM/test/mgcntf32/test_balance.py:20-62 (4-wide tf32 blocks, no
    width keyword), through forward_tf32_gnn and forward_tf32_gnn_acc."""
    m = compat.MODULES
    dd = torch.ones(32, dtype=torch.float32)
    rp, col, val, win, at = m["Block_sr_bcrs"].blockProcess_tf32_balance(
        torch.from_numpy(TOY_ROW), torch.from_numpy(TOY_COL), dd, 8, 4, 2)
    for dimN in (32, 65):
        rhs = torch.randint(1, 3, (30, dimN), generator=torch.Generator().manual_seed(dimN)).float()
        args = (rp.to(cuda), col.to(cuda), val.to(cuda), win.to(cuda), at.to(cuda), rhs.to(cuda), 32, dimN, 30)
        y = m["SpMM_sr_bcrs"].forward_tf32_gnn(*args)[0]
        ref = O.spmm(TOY_ROW, TOY_COL, np.ones(38), rhs.double().numpy(), n_rows=30)
        assert np.array_equal(y.cpu().double().numpy(), ref)
        assert torch.equal(y, m["SpMM_sr_bcrs"].forward_tf32_gnn_acc(*args)[0])


def _tile_values_f64(rp_csr, col_csr, d, dtype):
    v = (d[csr_rows(rp_csr)].astype(np.float32) * d[col_csr].astype(np.float32))
    return torch.from_numpy(v).to(dtype).double().numpy()


@pytest.mark.parametrize("window,wide,ps", [(8, 8, 32), (8, 8, 1), (16, 32, 2), (16, 32, 1000), (8, 4, 2),
                                            (16, 16, 1), (16, 8, 3), (8, 32, 2)])
@pytest.mark.parametrize("dim", [16, 64, 48, 256, 66, 7])
def test_spmm_native_vs_oracle(cuda, window, wide, ps, dim):
    rng = np.random.default_rng(window * 1000 + wide * 10 + ps + dim)
    n = 333
    rp, col = rand_csr(rng, n, 24, hub_rows=3)
    d = (rng.random(n) + 0.2).astype(np.float32)
    for xdt, ydt in ((torch.float16, torch.float16), (torch.bfloat16, torch.bfloat16)):
        src = "deg_f16" if xdt == torch.float16 else "deg_f32"
        dv = d.astype(np.float16) if xdt == torch.float16 else d
        b = ops.blocks_build(rp, col, src, dv, ydt, window, wide, ps, emit_empty=(window == 16))
        X = torch.from_numpy(rng.uniform(-1, 1, (n, dim)).astype(np.float32)).to(xdt)
        y = ops.spmm_block(b["rp"].to(cuda), b["col"].to(cuda), b["val"].to(cuda), b["window"].to(cuda),
                           b["atomic"].to(cuda), X.to(cuda), n, window, wide)
        # oracle with the same quantised tile values and X (error = FP32 accumulation only)
        if xdt == torch.float16:
            vals = torch.from_numpy(dv[csr_rows(rp)].astype(np.float32) * dv[col].astype(np.float32)).half().double().numpy()
        else:
            vals = _tile_values_f64(rp, col, d, torch.bfloat16)
        covered = n // window * window
        ref = O.spmm(rp, col, vals, X.double().numpy())
        ref[covered:] = 0.0                     # trailing rows are dropped by the windowing
        err = O.rel_fro(y.cpu().numpy(), ref)
        assert err < 2e-6, (xdt, err)


@pytest.mark.parametrize("mode,tol", [("bf16x3", 1e-5), ("f32", 2e-6)])
@pytest.mark.parametrize("window,wide,ps", [(8, 4, 32), (16, 32, 2), (16, 8, 1)])
def test_spmm_fp32_modes_vs_fp64(cuda, mode, tol, window, wide, ps):
    rng = np.random.default_rng(7)
    n = 400
    rp, col = rand_csr(rng, n, 30, hub_rows=2)
    vals = rng.standard_normal(col.size).astype(np.float32)
    b = ops.blocks_build(rp, col, "edge_f32", vals, torch.float32, window, wide, ps, emit_empty=True)
    for dim in (64, 128, 20):
        X = rng.uniform(-1, 1, (n, dim)).astype(np.float32)
        y = ops.spmm_block(b["rp"].to(cuda), b["col"].to(cuda), b["val"].to(cuda), b["window"].to(cuda),
                           b["atomic"].to(cuda), torch.from_numpy(X).to(cuda), n, window, wide, mode=mode)
        ref = O.spmm(rp, col, vals, X)
        ref[n // window * window:] = 0.0
        err = O.rel_fro(y.cpu().numpy(), ref)
        assert err < tol, (mode, dim, err)


def test_tf32_dropin_and_torch_sparse(cuda):
    """SpMM_sr_bcrs.forward_tf32_gnn on the 16,384-node Chung-Lu graph (mean
    degree 12) vs torch.sparse.mm in FP32 on the same device."""
    g = G.chung_lu(16384, 12, 2.1, seed=0)
    d = G.gcn_norm_degree(g)
    rp, col, val, win, at = compat.MODULES["Block_sr_bcrs"].blockProcess_tf32_balance(
        torch.from_numpy(g.rowptr32()), torch.from_numpy(g.indices), torch.from_numpy(d), 8, 4, 32)
    X = torch.from_numpy(np.random.default_rng(1).uniform(-1, 1, (g.n, 64)).astype(np.float32)).to(cuda)
    y = compat.MODULES["SpMM_sr_bcrs"].forward_tf32_gnn(rp.to(cuda), col.to(cuda), val.to(cuda), win.to(cuda),
                                                         at.to(cuda), X, g.n, 64, g.n)[0]
    A = torch.sparse_csr_tensor(torch.from_numpy(g.indptr), torch.from_numpy(g.indices.astype(np.int64)),
                                torch.from_numpy(G.gcn_norm_values(g, d)), (g.n, g.n)).to(cuda)
    ref = torch.sparse.mm(A, X)
    err = (y - ref).norm() / ref.norm()
    assert err.item() < 1e-5
    ref64 = O.spmm(g.indptr, g.indices, G.gcn_norm_values(g, d), X.cpu().numpy())
    assert O.rel_fro(y.cpu().numpy(), ref64) < 1e-5


def test_determinism_run_to_run(cuda):
    rng = np.random.default_rng(9)
    n = 2000
    rp, col = rand_csr(rng, n, 40, hub_rows=4)
    vals = rng.standard_normal(col.size).astype(np.float32)
    b = ops.blocks_build(rp, col, "edge_f32", vals, torch.bfloat16, 16, 32, 1, emit_empty=True)
    args = [b[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic")]
    assert int(b["atomic"].sum()) > 0
    X = torch.from_numpy(rng.uniform(-1, 1, (n, 128)).astype(np.float32)).to(torch.bfloat16).to(cuda)
    y0 = ops.spmm_block(*args, X, n, 16, 32)
    for _ in range(3):
        y1 = ops.spmm_block(*args, X, n, 16, 32)
        assert torch.equal(y0, y1)


def _tasks_np(rp, win, chunk_cols, tmax):
    """Restatement of the canonical task table: each window's columns [start,
    padded end) cut into chunks of chunk_cols from its start (one task for an
    empty window), running (real, pad) pairs; slots past the total get
    window -1 and empty ranges at the structure's end."""
    t_rp, t_win, t_at = [], [], []
    P, p = len(win), 0
    while p < P:
        w = int(win[p])
        if w < 0:
            p += 1
            continue
        q = p
        while q + 1 < P and win[q + 1] == w:
            q += 1
        s0, real_end, pad_end = int(rp[2 * p]), int(rp[2 * q + 1]), int(rp[2 * q + 2])
        c = 1 if pad_end <= s0 else -(-(pad_end - s0) // chunk_cols)
        for k in range(c):
            s = s0 + k * chunk_cols
            e = min(s + chunk_cols, pad_end)
            t_rp += [s, max(s, min(real_end, e))]
            t_win.append(w)
            t_at.append(int(c > 1))
        p = q + 1
    end = int(rp[2 * P])
    while len(t_win) < tmax:
        t_rp += [end, end]
        t_win.append(-1)
        t_at.append(0)
    return np.array(t_rp + [end], np.int32), np.array(t_win, np.int32), np.array(t_at, np.int32)


@pytest.mark.parametrize("window,wide,ps,chunk", [(16, 32, 1, 1), (16, 32, 2, 3), (16, 32, 32, 4), (8, 8, 1, 2),
                                                  (8, 4, 3, 5), (16, 32, 1 << 20, 2)])
@pytest.mark.parametrize("emit_empty", [True, False])
def test_canonical_tasks_vs_restatement(cuda, window, wide, ps, chunk, emit_empty):
    rng = np.random.default_rng(window + wide + ps + chunk)
    n = 600
    rp, col = rand_csr(rng, n, 50, hub_rows=4)
    b = ops.blocks_build(rp, col, "map", None, torch.int32, window, wide, ps, emit_empty=emit_empty)
    rpd, wind = b["rp"].to(cuda), b["window"].to(cuda)
    t_rp, t_win, t_at = ops.canonical_tasks(rpd, wind, b["col"].numel(), wide, chunk)
    want = _tasks_np(b["rp"].numpy(), b["window"].numpy(), chunk * wide, t_win.numel())
    assert np.array_equal(t_rp.cpu().numpy(), want[0])
    assert np.array_equal(t_win.cpu().numpy(), want[1])
    assert np.array_equal(t_at.cpu().numpy(), want[2])
    assert ops.canonical_tasks(rpd, wind, b["col"].numel(), wide, chunk)[0] is t_rp      # kept on rp


@pytest.mark.parametrize("dim", [64, 256])
@pytest.mark.parametrize("build", ["host", "device"])
def test_split_factor_invariance_native(cuda, dim, build):
    """The SpMM of GCNGraph is bitwise the same for part sizes 1, 2, 32 and
    unbounded: it runs over the canonical task table, so every window's sum
    is grouped by the chunk size alone (dense and compact tiles)."""
    g = G.chung_lu(4000, 60, 2.1, seed=3)
    X = torch.randn(g.n, dim, generator=torch.Generator().manual_seed(dim)).half().to(cuda)
    ys = []
    for ps in (1, 2, 32, 1 << 20):
        gg = GCNGraph(g, cuda, part_size=ps, tile_dtype=torch.float16, build=build)
        assert int(gg.fwd["tasks"][2].sum().item()) > 0          # split windows in the task table
        ys.append(gg.spmm(X, out_dtype=torch.float32))
    assert all(torch.equal(ys[0], y) for y in ys[1:])


@pytest.mark.parametrize("prec", ["fp16", "tf32"])
def test_split_factor_invariance_dropin(cuda, prec):
    """SpMM_sr_bcrs.forward_{fp16,tf32}_gnn on structures built by the
    reference-named preprocessing with partSize 1, 2, 32 and unbounded: the
    same result bit for bit (and exact on small-integer inputs)."""
    m = compat.MODULES
    g = G.chung_lu(3000, 40, 2.1, seed=5)
    wide = 8 if prec == "fp16" else 4
    dt = torch.float16 if prec == "fp16" else torch.float32
    dd = torch.from_numpy(G.gcn_norm_degree(g)).to(dt)
    rhs = torch.randn(g.n, 64, generator=torch.Generator().manual_seed(1)).to(dt).to(cuda)
    build = getattr(m["Block_sr_bcrs"], f"blockProcess_{prec}_balance")
    fwd = getattr(m["SpMM_sr_bcrs"], f"forward_{prec}_gnn")
    ys = []
    for ps in (1, 2, 32, 1 << 20):
        rp, col, val, win, at = build(torch.from_numpy(g.rowptr32()), torch.from_numpy(g.indices), dd, 8, wide, ps)
        if ps == 1:
            assert int(at.sum()) > 0
        ys.append(fwd(rp.to(cuda), col.to(cuda), val.to(cuda), win.to(cuda), at.to(cuda), rhs, g.n, 64, g.n)[0])
    assert all(torch.equal(ys[0], y) for y in ys[1:])


def test_split_factor_invariance_attention(cuda):
    """GATGraph's aggregations (A and A^T, compact edge lists) and the fused
    backward product are the same bit for bit for every part size."""
    from att_amd.layers import GATGraph
    g = G.chung_lu(3000, 40, 2.1, seed=6)
    gen = torch.Generator().manual_seed(2)
    p = torch.rand(2, g.nnz, generator=gen).bfloat16().to(cuda)
    X = torch.randn(g.n, 2, 64, generator=gen).bfloat16().to(cuda)
    outs = []
    for ps in (1, 2, 32, 1 << 20):
        gg = GATGraph(g, cuda, build="device", part_size=ps, chunk_blocks=2)
        outs.append((gg.aggregate(p, X), *gg.aggregate_sddmm(p, X, X)))
    for o in outs[1:]:
        assert all(torch.equal(a, b) for a, b in zip(outs[0], o))


def test_empty_and_degenerate(cuda):
    # graph with no edges: every row of Y is zero
    rp = np.zeros(33, np.int32)
    col = np.zeros(0, np.int32)
    b = ops.blocks_build(rp, col, "deg_f32", np.ones(32, np.float32), torch.float32, 8, 4, 2, emit_empty=False)
    assert b["window"].numel() == 0
    X = torch.ones(32, 16, device=cuda)
    y = ops.spmm_block(*[b[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic")], X, 32, 8, 4,
                       mode="f32")
    assert torch.count_nonzero(y) == 0


@pytest.mark.parametrize("window,wide,ps", [(16, 32, 1), (16, 32, 3), (16, 32, 1000), (8, 8, 2), (16, 8, 5)])
@pytest.mark.parametrize("dim", [8, 48, 64, 256, 66])
def test_pipeline_schedules_bitwise_equal(cuda, window, wide, ps, dim):
    """The k-loop schedules (none / one block / two blocks per iteration, with
    their column-id and compact-mask prefetches) only reorder loads: dense,
    compact and edge-map results are bit for bit the same, for every
    arithmetic kind, odd block counts and split windows included."""
    rng = np.random.default_rng(window * 100 + wide + ps * 7 + dim)
    n = 700
    rp, col = rand_csr(rng, n, 60, hub_rows=3)
    vals = rng.standard_normal(col.size).astype(np.float32)
    for dt, mode in ((torch.float16, None), (torch.bfloat16, None), (torch.float32, "f32"),
                     (torch.float32, "bf16x3")):
        b = ops.blocks_build(rp, col, "edge_f32", vals, dt, window, wide, ps, emit_empty=True,
                             compact=(wide == 32))
        args = [b[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic")]
        X = torch.from_numpy(rng.uniform(-1, 1, (n, dim)).astype(np.float32)).to(dt).to(cuda)
        ys = [ops.spmm_block(*args, X, n, window, wide, mode=mode, pipe=p) for p in ("none", "one", "pair")]
        assert all(torch.equal(ys[0], y) for y in ys[1:]), (dt, mode)
        if wide == 32:
            cargs = [b[k].to(cuda) for k in ("masks", "block_base", "values")]
            cs = [ops.spmm_compact(args[0], args[1], *cargs, args[3], args[4], X, n, window, mode=mode, pipe=p)
                  for p in ("none", "one", "pair")]
            assert all(torch.equal(ys[0], c) for c in cs), (dt, mode, "compact")
    # edge-map SpMM, 4 heads
    m = ops.blocks_build(rp, col, "map", None, torch.int32, window, wide, ps, emit_empty=True)
    margs = [m[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic")]
    ev = torch.rand(4, col.size, device=cuda).bfloat16()
    Xh = torch.randn(n, 4, dim, device=cuda).bfloat16()
    # 700 rows leave a partial last window, so Y is zero-filled (rows_covered=False)
    es = [ops.spmm_edge_map(margs[0], margs[1], margs[2], ev, margs[3], margs[4], Xh, "mhd", window, wide,
                            x_layout="mhd", pipe=p) for p in ("none", "one", "pair")]
    assert all(torch.equal(es[0], e) for e in es[1:])


def _with_padding_parts(rp, win, atomic, every):
    """Insert a padding part (window -1, atomic 0, no columns) before every
    `every`-th part that starts a new window, and three after the last."""
    rp, win, atomic = rp.cpu().numpy(), win.cpu().numpy(), atomic.cpu().numpy()
    nrp, nwin, nat = [rp[0]], [], []
    starts = 0
    for p in range(win.size):
        if p == 0 or win[p] != win[p - 1]:
            starts += 1
            if starts % every == 0:
                nrp += [rp[2 * p], rp[2 * p]]
                nwin.append(-1)
                nat.append(0)
        nrp += [rp[2 * p + 1], rp[2 * p + 2]]
        nwin.append(win[p])
        nat.append(atomic[p])
    for _ in range(3):
        nrp += [rp[-1], rp[-1]]
        nwin.append(-1)
        nat.append(0)
    return [torch.tensor(np.array(x, np.int32)) for x in (nrp, nwin, nat)]


@pytest.mark.parametrize("ps", [1, 2, 1000])
def test_padding_parts_skipped(cuda, ps):
    """Parts with a negative window (the padding of capacity-sized buffers) do
    no work and store nothing: dense, compact and edge-map results with
    padding parts between windows and after the last equal the unpadded ones
    bit for bit, with and without rows_covered, split windows included."""
    rng = np.random.default_rng(ps)
    n = 1024
    rp, col = rand_csr(rng, n, 50, hub_rows=3)
    vals = rng.standard_normal(col.size).astype(np.float32)
    b = ops.blocks_build(rp, col, "edge_f32", vals, torch.bfloat16, 16, 32, ps, emit_empty=True, compact=True)
    if ps == 1:
        assert int(b["atomic"].sum()) > 0
    prp, pwin, pat = (t.to(cuda) for t in _with_padding_parts(b["rp"], b["window"], b["atomic"], 3))
    assert int((pwin < 0).sum()) > 3
    brp, bcol, bval, bwin, bat = (b[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic"))
    X = torch.from_numpy(rng.uniform(-1, 1, (n, 96)).astype(np.float32)).bfloat16().to(cuda)
    for cov in (False, True):
        ref = ops.spmm_block(brp, bcol, bval, bwin, bat, X, n, 16, 32, rows_covered=cov)
        got = ops.spmm_block(prp, bcol, bval, pwin, pat, X, n, 16, 32, rows_covered=cov)
        assert torch.equal(ref, got), ("dense", cov)
        cargs = [b[k].to(cuda) for k in ("masks", "block_base", "values")]
        got = ops.spmm_compact(prp, bcol, *cargs, pwin, pat, X, n, 16, rows_covered=cov)
        assert torch.equal(ref, got), ("compact", cov)
    m = ops.blocks_build(rp, col, "map", None, torch.int32, 16, 32, ps, emit_empty=True)
    mrp, mcol, mmap, mwin, mat = (m[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic"))
    qrp, qwin, qat = (t.to(cuda) for t in _with_padding_parts(m["rp"], m["window"], m["atomic"], 2))
    ev = torch.rand(4, col.size, device=cuda).bfloat16()
    Xh = torch.randn(n, 4, 32, device=cuda).bfloat16()
    ref = ops.spmm_edge_map(mrp, mcol, mmap, ev, mwin, mat, Xh, "mhd", 16, 32, x_layout="mhd", rows_covered=True)
    got = ops.spmm_edge_map(qrp, mcol, mmap, ev, qwin, qat, Xh, "mhd", 16, 32, x_layout="mhd", rows_covered=True)
    assert torch.equal(ref, got)


@pytest.mark.parametrize("window,ps", [(16, 1), (16, 4), (8, 2), (16, 1000)])
@pytest.mark.parametrize("heads,dim", [(1, 64), (4, 32), (2, 48), (8, 8)])
def test_edge_map_compact_equals_dense_map(cuda, window, ps, heads, dim):
    """The compact edge map (row masks + packed edge ids) only changes where a
    lane finds its edge ids: results equal the dense-map SpMM bit for bit for
    every arithmetic kind, both layouts, every k-loop schedule, split windows
    and padding parts, on host- and device-built maps."""
    rng = np.random.default_rng(window * 1000 + ps * 10 + heads + dim)
    n = 900
    rp, col = rand_csr(rng, n, 45, hub_rows=3)
    m = ops.blocks_build(rp, col, "map", None, torch.int32, window, 32, ps, emit_empty=True, compact=True)
    md = ops.blocks_build_device(torch.from_numpy(rp).to(cuda), torch.from_numpy(col).to(cuda), "map", None,
                                 torch.int32, window, 32, ps, emit_empty=True, compact=True)
    for k in ("rp", "col", "val", "window", "atomic", "masks", "block_base"):
        assert torch.equal(m[k], md[k].cpu()), k
    nv = int(m["block_base"][-1])
    assert torch.equal(m["values"][:nv], md["values"][:nv].cpu())
    rp_, col_, emap, win, at, masks, base, ids = (m[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic",
                                                                          "masks", "block_base", "values"))
    for dt, mode in ((torch.bfloat16, None), (torch.float16, None), (torch.float32, "f32"),
                     (torch.float32, "bf16x3")):
        ev = torch.rand(heads, col.size, device=cuda).to(dt)
        for xl, ol in (("mhd", "mhd"), ("hmd", "hmd"), ("mhd", "hmd")):
            Xs = (n, heads, dim) if xl == "mhd" else (heads, n, dim)
            X = torch.randn(*Xs, device=cuda).to(dt)
            ref = ops.spmm_edge_map(rp_, col_, emap, ev, win, at, X, ol, window, 32, mode=mode, x_layout=xl)
            for pipe in (None, "none", "one", "pair"):
                got = ops.spmm_edge_map_compact(rp_, col_, masks, base, ids, ev, win, at, X, ol, window, mode=mode,
                                                x_layout=xl, pipe=pipe)
                assert torch.equal(ref, got), (dt, mode, xl, ol, pipe)
    prp, pwin, pat = (t.to(cuda) for t in _with_padding_parts(m["rp"], m["window"], m["atomic"], 2))
    ev = torch.rand(heads, col.size, device=cuda).bfloat16()
    X = torch.randn(n, heads, dim, device=cuda).bfloat16()
    ref = ops.spmm_edge_map(rp_, col_, emap, ev, win, at, X, "mhd", window, 32, x_layout="mhd", out_dtype=torch.bfloat16)
    got = ops.spmm_edge_map_compact(prp, col_, masks, base, ids, ev, pwin, pat, X, "mhd", window, x_layout="mhd",
                                    out_dtype=torch.bfloat16)
    assert torch.equal(ref, got)


@pytest.mark.parametrize("heads,dim", [(1, 256), (2, 256), (4, 32), (8, 8), (3, 128), (1, 512)])
def test_xcd_slice_mapping_bitwise_equal(cuda, heads, dim):
    """The slice-per-XCD task mapping only reorders waves: every SpMM variant
    gives the same bits with it on and off, for single and multi-slice heads."""
    rng = np.random.default_rng(heads * 100 + dim)
    n = 800
    rp, col = rand_csr(rng, n, 50, hub_rows=2)
    m = ops.blocks_build(rp, col, "map", None, torch.int32, 16, 32, 2, emit_empty=True, compact=True)
    a = [m[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic", "masks", "block_base", "values")]
    ev = torch.rand(heads, col.size, device=cuda).bfloat16()
    X = torch.randn(n, heads, dim, device=cuda).bfloat16()
    for xl, ol in (("mhd", "mhd"), ("mhd", "hmd")):
        on = ops.spmm_edge_map(*a[:3], ev, a[3], a[4], X, ol, 16, 32, x_layout=xl)
        off = ops.spmm_edge_map(*a[:3], ev, a[3], a[4], X, ol, 16, 32, x_layout=xl, xcd=False)
        assert torch.equal(on, off), (xl, ol)
        c = ops.spmm_edge_map_compact(a[0], a[1], a[5], a[6], a[7], ev, a[3], a[4], X, ol, 16, x_layout=xl)
        assert torch.equal(on, c), (xl, ol)
    s0 = torch.randn(n, heads, device=cuda)
    s1 = torch.randn(n, heads, device=cuda)
    fa = [ops.spmm_attention(*a[:5], X, "mhd", 16, 32, s0, s1, x_layout="mhd", xcd=x) for x in (True, False)]
    assert torch.equal(fa[0], fa[1])
    if heads == 1:
        vals = rng.standard_normal(col.size).astype(np.float32)
        b = ops.blocks_build(rp, col, "edge_f32", vals, torch.float16, 16, 32, 2, emit_empty=True, compact=True)
        args = [b[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic")]
        X2 = X.reshape(n, dim).half()
        on = ops.spmm_block(*args, X2, n, 16, 32)
        assert torch.equal(on, ops.spmm_block(*args, X2, n, 16, 32, xcd=False))
        cargs = [b[k].to(cuda) for k in ("masks", "block_base", "values")]
        assert torch.equal(on, ops.spmm_compact(args[0], args[1], *cargs, args[3], args[4], X2, n, 16, xcd=False))
