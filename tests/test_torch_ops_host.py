"""torch.ops.attgnn registration on the host: every op is registered, and its
fake implementation gives the output shapes and dtypes under FakeTensorMode
(fake cuda tensors need no device), so the ops trace without a GPU."""
import pytest
import torch
from torch._subclasses.fake_tensor import FakeTensorMode

import att_amd.torch_ops as T


def test_all_ops_registered():
    for name in T.OPS:
        op = getattr(torch.ops.attgnn, name)
        assert op.default._schema.name == f"attgnn::{name}"


def _i32(n):
    return torch.empty(n, dtype=torch.int32, device="cuda")


@pytest.fixture
def fake():
    with FakeTensorMode(allow_non_fake_inputs=False) as m:
        yield m


def test_spmm_shapes(fake):
    parts, n, dim, m = 7, 100, 48, 90
    rp, col, win, at = _i32(2 * parts + 1), _i32(64), _i32(parts), _i32(parts)
    val = torch.empty(64 * 16, dtype=torch.float16, device="cuda")
    X = torch.empty(n, dim, dtype=torch.float16, device="cuda")
    y = torch.ops.attgnn.spmm_block(rp, col, val, win, at, X, m, 16, 32, "auto", torch.float32, True)
    assert y.shape == (m, dim) and y.dtype == torch.float32 and y.device.type == "cuda"
    masks = torch.empty(64, dtype=torch.int16, device="cuda")
    base = torch.empty(3, dtype=torch.int64, device="cuda")
    y = torch.ops.attgnn.spmm_compact(rp, col, masks, base, val, win, at, X, m, 16, "auto", torch.bfloat16, True)
    assert y.shape == (m, dim) and y.dtype == torch.bfloat16


@pytest.mark.parametrize("x_layout,out_layout", [("hmd", "hmd"), ("mhd", "mhd"), ("mhd", "hmd")])
def test_spmm_edge_map_shapes(fake, x_layout, out_layout):
    heads, n, dim, nnz = 4, 100, 16, 500
    X = torch.empty((heads, n, dim) if x_layout == "hmd" else (n, heads, dim), dtype=torch.float16, device="cuda")
    ev = torch.empty(heads, nnz, dtype=torch.float16, device="cuda")
    args = (_i32(9), _i32(64), _i32(64 * 16), ev, _i32(4), _i32(4), X, out_layout, 16, 32)
    y = torch.ops.attgnn.spmm_edge_map(*args, -1, "auto", x_layout, torch.float32, True)
    assert y.shape == ((heads, n, dim) if out_layout == "hmd" else (n, heads, dim))
    y = torch.ops.attgnn.spmm_edge_map(*args, 80, "auto", x_layout, torch.float16, True)
    assert y.shape[0 if out_layout == "mhd" else 1] == 80 and y.dtype == torch.float16
    s = torch.empty(n, heads, device="cuda")
    lse = torch.empty(heads, n, device="cuda")
    y = torch.ops.attgnn.spmm_attention(*args[:3], args[4], args[5], X, out_layout, 16, 32, s, s, lse, None, 0.2,
                                        0.0, 0, -1, "auto", x_layout, torch.bfloat16, True)
    assert y.shape == ((heads, n, dim) if out_layout == "hmd" else (n, heads, dim)) and y.dtype == torch.bfloat16


def test_spmm_edge_map_compact_shapes(fake):
    heads, n, dim, nnz, parts = 4, 100, 64, 500, 4
    X = torch.empty(n, heads, dim, dtype=torch.bfloat16, device="cuda")
    ev = torch.empty(heads, nnz, dtype=torch.bfloat16, device="cuda")
    masks = torch.empty(64, dtype=torch.int16, device="cuda")
    base = torch.empty(3, dtype=torch.int64, device="cuda")
    args = (_i32(2 * parts + 1), _i32(64), masks, base, _i32(nnz), ev, _i32(parts), _i32(parts), X)
    y = torch.ops.attgnn.spmm_edge_map_compact(*args, "mhd", 16, -1, "auto", "mhd", torch.float32, True)
    assert y.shape == (n, heads, dim) and y.dtype == torch.float32
    y, dot = torch.ops.attgnn.spmm_edge_map_compact_dot(*args, X, "hmd", 16, 80, "mhd", torch.bfloat16, True)
    assert y.shape == (heads, 80, dim) and y.dtype == torch.bfloat16
    assert dot.shape == (heads, nnz) and dot.dtype == torch.float32


def test_attention_shapes(fake):
    heads, n, dim, nnz = 2, 50, 8, 300
    X = torch.empty(n, heads, dim, dtype=torch.bfloat16, device="cuda")
    vec = torch.empty(heads, dim, dtype=torch.float32, device="cuda")
    s = torch.ops.attgnn.node_scores(X, vec, "mhd")
    assert s.shape == (heads, n) and s.dtype == torch.float32
    sn = torch.ops.attgnn.node_scores(X, vec, "mhd", "nh")
    assert sn.shape == (n, heads)
    p, lse = torch.ops.attgnn.edge_softmax_fwd(_i32(n + 1), _i32(nnz), sn, sn, 0.2, 0.0, 0, torch.float16, "nh")
    assert p.shape == (heads, nnz) and lse.shape == (heads, n)
    t_perm = torch.empty(nnz, dtype=torch.int64, device="cuda")
    d0, d1 = torch.ops.attgnn.edge_softmax_bwd_sums(_i32(n + 1), _i32(nnz), sn, sn, lse, p, 0.2, 0.0, 0,
                                                    _i32(n + 1), t_perm, "nh")
    assert d0.shape == (heads, n) and d1.shape == (heads, n) and d1.dtype == torch.float32
    e = torch.ops.attgnn.sddmm_edges(_i32(5), _i32(32), torch.empty(3, dtype=torch.int64, device="cuda"),
                                     _i32(nnz), torch.empty(nnz, dtype=torch.int16, device="cuda"), _i32(2), X, X,
                                     nnz, 16, "mhd", torch.float32, True)
    assert e.shape == (heads, nnz) and e.dtype == torch.float32
    rowptr, col = _i32(n + 1), _i32(nnz)
    p, lse = torch.ops.attgnn.edge_softmax_fwd(rowptr, col, s, s, 0.2, 0.1, 7, torch.bfloat16)
    assert p.shape == (heads, nnz) and p.dtype == torch.bfloat16 and lse.shape == (heads, n)
    dz, ds0 = torch.ops.attgnn.edge_softmax_bwd(rowptr, col, s, s, lse, p, 0.2, 0.1, 7)
    assert dz.shape == (heads, nnz) and dz.dtype == torch.float32 and ds0.shape == (heads, n)
    cs = torch.ops.attgnn.gather_rowsum(rowptr, torch.empty(nnz, dtype=torch.int64, device="cuda"), dz, n)
    assert cs.shape == (heads, n) and cs.dtype == torch.float32
    base = torch.empty(n, heads, dim, dtype=torch.float32, device="cuda")
    dXs, dXt, da0, da1 = torch.ops.attgnn.attention_node_bwd(X, X, vec, vec, s, s, base, True)
    assert dXs.numel() == 0 and dXt.shape == X.shape and dXt.dtype == X.dtype
    assert da0.shape == (heads, dim) and da1.dtype == torch.float32
    dXs, dXt, _, _ = torch.ops.attgnn.attention_node_bwd(X, X.clone(), vec, vec, s, s, None, False)
    assert dXs.shape == X.shape
