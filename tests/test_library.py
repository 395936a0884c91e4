"""The C-ABI library loads and exports every entry point include/attgnn.h
declares; argument validation returns errors before touching a device."""
import ctypes
import os
import re

from att_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "attgnn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(attgnn_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(_lib.EXPORTED), set(names) - set(_lib.EXPORTED)


def test_status_strings():
    L = _lib.lib()
    assert L.attgnn_status_string(0) == b"ok"
    assert L.attgnn_status_string(-1) == b"invalid argument"


def test_spmm_rejects_bad_arguments_without_device():
    L = _lib.lib()
    nul = ctypes.c_void_p(0)
    # window 12 is not supported: rejected during host validation
    rc = L.attgnn_spmm_block(nul, nul, nul, _lib.F16, nul, nul, 0, nul, _lib.F16, 0, nul, _lib.F32, 0, 16, 12, 8,
                             _lib.MODE_NATIVE, nul, nul, nul, 0, 0, nul)
    assert rc == -1
    # mismatched tile / X types
    rc = L.attgnn_spmm_block(nul, nul, nul, _lib.F16, nul, nul, 0, nul, _lib.BF16, 0, nul, _lib.F32, 0, 16, 8, 8,
                             _lib.MODE_NATIVE, nul, nul, nul, 0, 0, nul)
    assert rc == -4
    rc = L.attgnn_sddmm_block(nul, nul, nul, nul, 1, nul, 0, 0, nul, 0, 0, _lib.F16, 1, 1, 8, nul, _lib.F16, 1,
                              8, 7, 0, nul)
    assert rc == -1


def test_fused_dot_rejects_bad_arguments_without_device():
    """attgnn_spmm_edge_map_compact_dot validates on the host: a negative
    dot_rows is an argument error, an unsupported width (48) is unsupported."""
    L = _lib.lib()
    nul = ctypes.c_void_p(0)

    def call(dim, dot_rows):
        return L.attgnn_spmm_edge_map_compact_dot(nul, nul, nul, nul, nul, nul, _lib.BF16, 0, nul, nul, 0, nul,
                                                  _lib.BF16, _lib.LAYOUT_HMD, 1, 0, ctypes.c_void_p(256), _lib.F32,
                                                  0, dim, _lib.LAYOUT_HMD, 16, nul, dot_rows, nul, nul, 0, 0, nul)
    assert call(64, -1) == -1
    assert call(48, 0) == -4


def test_range_and_minmax_reject_bad_arguments_without_device():
    """attgnn_edge_score_range / attgnn_minmax_score_grads: sizes, missing
    buffers, a workspace below the stated size and a long-row count without
    its list are refused before any launch."""
    L = _lib.lib()
    nul = ctypes.c_void_p(0)
    fake = ctypes.c_void_p(0x1000)    # never dereferenced: validation fails first
    ws = L.attgnn_edge_score_range_workspace_elems(4)
    assert ws == 4 * 2 * 8192 and L.attgnn_edge_score_range_workspace_elems(0) == 0
    # heads out of range
    assert L.attgnn_edge_score_range(fake, fake, 10, 20, 0, fake, fake, 1, 4, fake, ws, fake, fake, nul, 0, nul) == -1
    assert L.attgnn_edge_score_range(fake, fake, 10, 20, 257, fake, fake, 1, 4, fake, ws, fake, fake, nul, 0, nul) == -1
    # no range / where outputs
    assert L.attgnn_edge_score_range(fake, fake, 10, 20, 4, fake, fake, 1, 4, fake, ws, nul, fake, nul, 0, nul) == -1
    # workspace too small
    assert L.attgnn_edge_score_range(fake, fake, 10, 20, 4, fake, fake, 1, 4, fake, ws - 1, fake, fake, nul, 0,
                                     nul) == -1
    # long-row count without a list
    assert L.attgnn_edge_score_range(fake, fake, 10, 20, 4, fake, fake, 1, 4, fake, ws, fake, fake, nul, 3, nul) == -1
    mg = L.attgnn_minmax_grads_workspace_elems(4)
    assert mg == 4 * (256 + 1)
    assert L.attgnn_minmax_score_grads(fake, fake, 10, 0, 1, 4, fake, fake, fake, fake, fake, fake, fake, mg,
                                       nul) == -1
    assert L.attgnn_minmax_score_grads(fake, fake, 10, 4, 1, 4, fake, fake, fake, fake, fake, fake, fake, mg - 1,
                                       nul) == -1
    assert L.attgnn_minmax_score_grads(fake, fake, 10, 4, 1, 4, fake, nul, fake, fake, fake, fake, fake, mg,
                                       nul) == -1
    # nothing to do for zero nodes: ok without touching the buffers
    assert L.attgnn_minmax_score_grads(nul, nul, 0, 4, 1, 4, nul, nul, nul, nul, nul, nul, nul, 0, nul) == 0


def test_nll_log_softmax_rejects_bad_arguments_without_device():
    """attgnn_nll_log_softmax: logit type, class count against the row
    stride, an empty mean, missing buffers (neither loss nor gradient, no
    logits, a loss without its workspace), a short workspace and more
    selected rows than rows without a list are refused before any launch."""
    L = _lib.lib()
    nul = ctypes.c_void_p(0)
    fake = ctypes.c_void_p(0x1000)    # never dereferenced: validation fails first
    ws = L.attgnn_nll_log_softmax_workspace_elems(1000)
    assert ws == 63 and L.attgnn_nll_log_softmax_workspace_elems(1) == 1   # 16 rows per workgroup
    assert L.attgnn_nll_log_softmax_workspace_elems(10 ** 6) == 2048      # at most 2048 workgroups
    assert L.attgnn_nll_log_softmax_workspace_elems(-1) == -1

    def call(t=1, n=1000, ld=48, c=41, rows=nul, sel=1000, loss=fake, grad=fake, wsp=fake, ws=ws, logits=fake):
        return L.attgnn_nll_log_softmax(logits, t, n, ld, c, fake, rows, sel, loss, grad, nul, wsp, ws, nul)
    assert call(t=3) == -1                 # an int32 type
    assert call(c=49) == -1                # classes beyond the row stride
    assert call(c=0) == -1
    assert call(sel=0, ws=1) == -1         # the mean of no rows
    assert call(sel=1001) == -1            # more rows than there are, without a list
    assert call(loss=nul, grad=nul) == -1  # nothing to compute
    assert call(logits=nul) == -1
    assert call(wsp=nul) == -1 and call(ws=ws - 1) == -1   # a loss needs its workspace


def test_relu_dropout_rejects_bad_arguments_without_device():
    """attgnn_relu_dropout_fwd / _bwd: element counts not a multiple of 8,
    p outside [0, 1], unaligned or missing buffers and a missing mask are
    refused before any launch."""
    L = _lib.lib()
    nul = ctypes.c_void_p(0)
    al = ctypes.c_void_p(0x1000)
    mis = ctypes.c_void_p(0x1008)
    assert L.attgnn_relu_dropout_fwd(al, _lib.F16, 12, 0.5, 7, al, al, nul) == -1      # 12 % 8
    assert L.attgnn_relu_dropout_fwd(al, _lib.F16, 16, 1.5, 7, al, al, nul) == -1      # p > 1
    assert L.attgnn_relu_dropout_fwd(al, _lib.F16, 16, -0.1, 7, al, al, nul) == -1
    assert L.attgnn_relu_dropout_fwd(mis, _lib.F16, 16, 0.5, 7, al, al, nul) == -1     # unaligned input
    assert L.attgnn_relu_dropout_fwd(al, _lib.F16, 16, 0.5, 7, al, nul, nul) == -1     # no mask
    assert L.attgnn_relu_dropout_fwd(al, _lib.FP8E4M3, 16, 0.5, 7, al, al, nul) == -1  # not fp32/16/bf16
    assert L.attgnn_relu_dropout_fwd(nul, _lib.F16, 0, 0.5, 7, nul, nul, nul) == 0     # nothing to do
    assert L.attgnn_relu_dropout_bwd(al, _lib.BF16, 16, 0.5, nul, al, nul) == -1       # no mask
    assert L.attgnn_relu_dropout_bwd(al, _lib.BF16, 16, 0.5, al, mis, nul) == -1       # unaligned output
