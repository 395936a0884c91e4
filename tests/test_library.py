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
