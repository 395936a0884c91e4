"""ctypes binding of libattgnn.so (the C ABI declared in include/attgnn.h).

The library is loaded once; a missing library raises immediately -- there is
no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libattgnn.so")

# element types / modes / sources / layouts (attgnn.h)
F32, F16, BF16, FP8E4M3, I32 = 0, 1, 2, 3, 4
MODE_NATIVE, MODE_BF16X3, MODE_F32 = 0, 1, 2
SRC_DEG_F16, SRC_DEG_F32, SRC_EDGE_F32, SRC_MAP = 0, 1, 2, 3
LAYOUT_HMD, LAYOUT_MHD = 0, 1
SPMM_ROWS_COVERED = 1
SPMM_NO_XCD_SLICES = 1 << 6

_c = ctypes
_vp, _i64, _int = _c.c_void_p, _c.c_int64, _c.c_int


class Segments(ctypes.Structure):
    """attgnn_segments (include/attgnn.h)."""
    _fields_ = [("seg_ptr", _c.c_void_p), ("seg_row", _c.c_void_p), ("seg_slot", _c.c_void_p),
                ("row_seg", _c.c_void_p), ("n_seg", _c.c_int64), ("n_slots", _c.c_int64)]


_SEGP = _c.POINTER(Segments)


class AttnTerms(ctypes.Structure):
    """attgnn_attn_terms (include/attgnn.h)."""
    _fields_ = [("row_logit", _c.c_void_p), ("col_logit", _c.c_void_p), ("row_offset", _c.c_void_p),
                ("col_offset", _c.c_void_p), ("logit_head_stride", _c.c_int64), ("logit_node_stride", _c.c_int64),
                ("offset_head_stride", _c.c_int64), ("offset_node_stride", _c.c_int64), ("n_nodes", _c.c_int64),
                ("slope", _c.c_float), ("drop_p", _c.c_float), ("seed", _c.c_uint64)]


_ATTNP = _c.POINTER(AttnTerms)

_SIGNATURES = {
    "attgnn_status_string": (_c.c_char_p, [_int]),
    "attgnn_blocks_build": (_int, [_vp, _i64, _vp, _i64, _int, _vp, _int, _int, _int, _int, _int, _int, _int,
                                   _c.POINTER(_vp)]),
    "attgnn_blocks_sizes": (_int, [_vp, _c.POINTER(_i64), _c.POINTER(_i64), _c.POINTER(_i64),
                                   _c.POINTER(_i64), _c.POINTER(_i64), _c.POINTER(_int)]),
    "attgnn_blocks_export": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "attgnn_blocks_free": (None, [_vp]),
    "attgnn_spmm_block": (_int, [_vp, _vp, _vp, _int, _vp, _vp, _i64, _vp, _int, _i64, _vp, _int, _i64, _i64,
                                 _int, _int, _int, _vp, _vp, _vp, _i64, _int, _vp]),
    "attgnn_spmm_workspace_elems": (_i64, [_i64, _int, _i64]),
    "attgnn_spmm_striped_workspace_elems": (_i64, [_i64, _int, _i64, _int]),
    "attgnn_stripe_segments_device": (_int, [_vp, _vp, _vp, _i64, _int, _i64, _vp, _vp]),
    "attgnn_spmm_compact_striped": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _vp, _vp, _i64, _vp, _int, _vp, _int,
                                           _i64, _vp, _int, _i64, _i64, _int, _int, _vp, _i64, _int, _vp]),
    "attgnn_blocks_tasks_max": (_i64, [_i64, _i64, _int, _int]),
    "attgnn_blocks_tasks_workspace_bytes": (_i64, [_i64]),
    "attgnn_blocks_tasks_device": (_int, [_vp, _vp, _i64, _i64, _int, _int, _vp, _vp, _vp, _i64, _vp, _i64, _vp,
                                          _vp]),
    "attgnn_spmm_block_noalloc": (_int, [_vp, _vp, _vp, _int, _vp, _vp, _i64, _vp, _int, _i64, _vp, _int, _i64,
                                         _i64, _int, _int, _int, _vp, _vp, _int, _vp]),
    "attgnn_spmm_compact": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _vp, _vp, _i64, _vp, _int, _i64, _vp, _int, _i64,
                                   _i64, _int, _int, _vp, _vp, _vp, _i64, _int, _vp]),
    "attgnn_blocks_compact_sizes": (_int, [_vp, _c.POINTER(_i64), _c.POINTER(_i64)]),
    "attgnn_blocks_compact_export": (_int, [_vp, _vp, _vp, _vp, _vp]),
    "attgnn_sddmm_edges": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _int, _i64,
                                  _i64, _i64, _vp, _int, _i64, _int, _int, _vp]),
    "attgnn_spmm_edge_map": (_int, [_vp, _vp, _vp, _vp, _int, _i64, _vp, _vp, _i64, _vp, _int, _int, _i64, _i64,
                                    _vp, _int, _i64, _i64, _int, _int, _int, _int, _vp, _i64, _int, _vp]),
    "attgnn_spmm_edge_map_compact": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _int, _i64, _vp, _vp, _i64, _vp, _int,
                                            _int, _i64, _i64, _vp, _int, _i64, _i64, _int, _int, _int, _vp, _i64,
                                            _int, _vp]),
    "attgnn_spmm_edge_map_compact_dot": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _int, _i64, _vp, _vp, _i64, _vp,
                                                 _int, _int, _i64, _i64, _vp, _int, _i64, _i64, _int, _int, _vp,
                                                 _i64, _vp, _vp, _i64, _int, _vp]),
    "attgnn_spmm_attention": (_int, [_vp, _vp, _vp, _vp, _vp, _i64, _ATTNP, _vp, _int, _int, _i64, _i64, _vp, _int,
                                     _i64, _i64, _int, _int, _int, _int, _vp, _i64, _int, _vp]),
    "attgnn_rowsum_edge_map": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _i64, _i64, _vp, _int, _i64, _int, _int,
                                      _vp]),
    "attgnn_sddmm_block": (_int, [_vp, _vp, _vp, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _int, _i64, _i64,
                                  _i64, _vp, _int, _i64, _int, _int, _int, _vp]),
    "attgnn_node_scores": (_int, [_vp, _int, _i64, _i64, _vp, _int, _i64, _i64, _i64, _vp, _i64, _i64, _vp]),
    "attgnn_node_scores_pair": (_int, [_vp, _int, _i64, _i64, _vp, _vp, _int, _i64, _i64, _i64, _vp, _vp, _i64, _i64,
                                       _vp]),
    "attgnn_edge_from_nodes": (_int, [_vp, _vp, _vp, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _int,
                                      _i64, _i64, _vp, _int, _i64, _int, _int, _vp]),
    "attgnn_long_rows": (_i64, [_vp, _i64, _vp]),
    "attgnn_edge_softmax_fwd": (_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _i64, _c.c_float, _c.c_float,
                                       _c.c_uint64, _vp, _int, _vp, _vp, _i64, _vp]),
    "attgnn_edge_softmax_bwd": (_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _vp, _int,
                                       _c.c_float, _c.c_float, _c.c_uint64, _vp, _vp, _vp, _i64, _vp]),
    "attgnn_edge_softmax_bwd_sums_workspace_elems": (_i64, [_i64, _i64]),
    "attgnn_edge_softmax_bwd_sums": (_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _vp, _int,
                                            _c.c_float, _c.c_float, _c.c_uint64, _vp, _vp, _vp, _i64, _vp, _i64,
                                            _vp, _vp, _vp, _i64, _vp]),
    "attgnn_gather_rowsum": (_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp]),
    "attgnn_edge_score_range_workspace_elems": (_i64, [_i64]),
    "attgnn_minmax_grads_workspace_elems": (_i64, [_i64]),
    "attgnn_minmax_score_grads": (_int, [_vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                         _vp]),
    "attgnn_edge_score_range": (_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _vp, _vp,
                                       _i64, _vp]),
    "attgnn_blocks_device_workspace_bytes": (_i64, [_i64, _i64, _int]),
    "attgnn_blocks_plan_device": (_int, [_vp, _i64, _vp, _i64, _int, _int, _int, _int, _vp, _i64, _vp, _vp]),
    "attgnn_blocks_fill_device": (_int, [_vp, _i64, _vp, _i64, _int, _vp, _int, _int, _int, _int, _vp, _i64, _vp,
                                         _vp, _vp, _vp, _vp, _vp, _vp]),
    "attgnn_blocks_compact_workspace_bytes": (_i64, [_i64]),
    "attgnn_blocks_compact_count_device": (_int, [_vp, _int, _i64, _int, _vp, _vp, _vp, _i64, _vp]),
    "attgnn_blocks_compact_fill_device": (_int, [_vp, _int, _i64, _int, _vp, _vp, _vp, _vp, _vp]),
    "attgnn_map_transpose_perm_device": (_int, [_vp, _vp, _i64, _i64, _vp, _vp]),
    "attgnn_attention_node_bwd_workspace_elems": (_i64, [_i64, _i64, _i64]),
    "attgnn_attention_node_bwd": (_int, [_vp, _vp, _int, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _i64, _vp]),
    "attgnn_gatv2_fwd": (_int, [_vp, _vp, _SEGP, _i64, _i64, _i64, _i64, _vp, _vp, _int, _vp, _c.c_float,
                                _c.c_float, _c.c_uint64, _vp, _vp, _vp, _int, _vp]),
    "attgnn_gatv2_bwd_workspace_elems": (_i64, [_i64, _i64, _SEGP, _SEGP]),
    "attgnn_gatv2_bwd": (_int, [_vp, _vp, _SEGP, _vp, _vp, _SEGP, _i64, _i64, _i64, _i64, _vp, _vp, _int, _vp,
                                _c.c_float, _c.c_float, _c.c_uint64, _vp, _vp, _vp, _int, _vp, _vp, _vp, _vp, _vp,
                                _vp, _i64, _vp]),
    "attgnn_csr_transpose_workspace_bytes": (_i64, [_i64, _i64, _i64]),
    "attgnn_csr_transpose_device": (_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "attgnn_map_remap_device": (_int, [_vp, _i64, _vp, _i64, _vp, _vp]),
    "attgnn_pack_segments": (_int, [_vp, _i64, _vp, _i64, _i64, _vp]),
    "attgnn_pack_segment_bytes": (_i64, []),
    "attgnn_relu_dropout_fwd": (_int, [_vp, _int, _i64, _c.c_float, _c.c_uint64, _vp, _vp, _vp]),
    "attgnn_relu_dropout_bwd": (_int, [_vp, _int, _i64, _c.c_float, _vp, _vp, _vp]),
    "attgnn_nll_log_softmax_workspace_elems": (_i64, [_i64]),
    "attgnn_nll_log_softmax": (_int, [_vp, _int, _i64, _i64, _int, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp]),
    "attgnn_serve_stage_layout": (_i64, [_i64, _i64, _i64, _int, _vp]),
    "attgnn_serve_stage_host": (_int, [_vp, _vp, _i64, _i64, _vp, _i64, _int, _vp, _i64]),
    "attgnn_serve_stage": (_int, [_vp, _vp, _i64, _i64, _vp, _i64, _int, _vp, _i64, _vp, _vp]),
}

# Every symbol include/attgnn.h declares (checked by the CPU test suite).
EXPORTED = tuple(_SIGNATURES)

_lib = None


class AttGnnError(RuntimeError):
    pass


def lib():
    """Load libattgnn.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AttGnnError(
                f"{LIB_PATH} is missing: build it with `python -m att_amd.build` "
                "(there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().attgnn_status_string(rc).decode()
        raise AttGnnError(f"{what} failed: {msg} (status {rc})")
