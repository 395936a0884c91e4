"""The allocation-free window-block SpMM (attgnn_spmm_block_noalloc, scratch in a
fixed device array) against attgnn_spmm_block with a caller workspace, both
through the C ABI: bitwise equal results on two calls in a row, the second on
a smaller problem (so the second reads a scratch the first left filled), with
split windows in both; a problem larger than the scratch is refused."""
import ctypes

import numpy as np
import pytest
import torch

from att_amd import _lib
from att_amd import graphs as G
from att_amd import ops

pytestmark = pytest.mark.gpu

P = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)


def _problem(cuda, n, deg, dim, part_size, seed):
    g = G.chung_lu(n, deg, 2.1, seed=seed)
    p = g.padded(16)
    d = G.gcn_norm_degree(p)
    d[~np.isfinite(d)] = 0.0
    b = ops.blocks_build(p.rowptr32(), p.indices, "edge_f32", G.gcn_norm_values(p, d), torch.float16, 16, 32,
                         part_size, emit_empty=True)
    arrs = [b[k].to(cuda) for k in ("rp", "col", "val", "window", "atomic")]
    X = torch.randn(p.n, dim, generator=torch.Generator().manual_seed(seed)).half().to(cuda)
    return arrs, X, p.n


def _args(arrs, X, m, Y, flags=0):
    rp, col, val, win, at = arrs
    return (P(rp), P(col), P(val), _lib.F16, P(win), P(at), win.numel(), P(X), _lib.F16, X.shape[0], P(Y), _lib.F32,
            m, X.shape[1], 16, 32, _lib.MODE_NATIVE, None, None)


def test_noalloc_equals_workspace_entry_point(cuda):
    L = _lib.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    for n, deg, dim, ps, seed in ((3000, 14, 128, 2, 1), (700, 9, 64, 1, 2)):
        arrs, X, m = _problem(cuda, n, deg, dim, ps, seed)
        parts = arrs[3].numel()
        assert int(arrs[4].sum().item()) > 0                      # some windows are split
        need = L.attgnn_spmm_workspace_elems(parts, 16, dim)
        assert need <= 16 * 1024 * 1024
        ws = torch.empty(need, dtype=torch.float32, device=cuda)
        Y_ref = torch.full((m, dim), float("nan"), device=cuda)
        Y = torch.full((m, dim), float("nan"), device=cuda)
        for flags in (0, _lib.SPMM_ROWS_COVERED):
            rc = L.attgnn_spmm_block(*_args(arrs, X, m, Y_ref), P(ws), need, flags, stream)
            assert rc == 0, L.attgnn_status_string(rc)
            rc = L.attgnn_spmm_block_noalloc(*_args(arrs, X, m, Y), flags, stream)
            assert rc == 0, L.attgnn_status_string(rc)
            torch.cuda.synchronize()
            assert torch.equal(Y, Y_ref), (n, flags)
            assert not torch.isnan(Y).any()                        # every row written


def test_noalloc_refuses_problems_above_its_scratch():
    """Checked on the host before any device work (null pointers never read)."""
    L = _lib.lib()
    nul = ctypes.c_void_p(0)
    parts = 16 * 1024 * 1024 // (16 * 256) + 1                   # one part too many at 256 features
    rc = L.attgnn_spmm_block_noalloc(nul, nul, nul, _lib.F16, nul, nul, parts, nul, _lib.F16, 0, nul, _lib.F32,
                                     0, 256, 16, 32, _lib.MODE_NATIVE, nul, nul, 0, nul)
    assert rc == -5
    assert L.attgnn_status_string(rc).decode().startswith("larger than the fixed scratch")
