"""torch.library registration of the device ops as ``torch.ops.attgnn.*``.

Each op calls the same C-ABI entry point as :mod:`att_amd.ops` and has a fake
(shape / dtype) implementation, so the ops trace under FakeTensorMode and
torch.compile (no graph breaks) and can be used from TorchScript-free
serialised graphs.  The op names, argument order and argument meaning follow
the library's C ABI (include/attgnn.h); the reference-named list-returning
modules live in :mod:`att_amd.compat`.

    import att_amd.torch_ops
    y = torch.ops.attgnn.spmm_block(rp, col, val, win, atomic, X, m_ori, 16, 32, "auto", torch.float16, True)
"""
from typing import Optional, Tuple

import torch

from . import ops

_LIB = "attgnn"


def _mode(mode: str):
    return None if mode == "auto" else mode


def _hmd(X, layout):
    if X.dim() == 2:
        return 1, X.shape[0], X.shape[1]
    if layout == "hmd":
        return X.shape[0], X.shape[1], X.shape[2]
    return X.shape[1], X.shape[0], X.shape[2]


# ---------------------------------------------------------------- SpMM

@torch.library.custom_op(f"{_LIB}::spmm_block", mutates_args=())
def spmm_block(rp: torch.Tensor, col: torch.Tensor, val: torch.Tensor, window_of_part: torch.Tensor,
               atomic: torch.Tensor, X: torch.Tensor, m_ori: int, window: int, wide: int, mode: str,
               out_dtype: torch.dtype, rows_covered: bool) -> torch.Tensor:
    """Y[m_ori, D] = A . X over dense window tiles (attgnn_spmm_block)."""
    return ops.spmm_block(rp, col, val, window_of_part, atomic, X, m_ori, window, wide, mode=_mode(mode),
                          out_dtype=out_dtype, rows_covered=rows_covered)


@spmm_block.register_fake
def _(rp, col, val, window_of_part, atomic, X, m_ori, window, wide, mode, out_dtype, rows_covered):
    return X.new_empty((m_ori, X.shape[1]), dtype=out_dtype)


@torch.library.custom_op(f"{_LIB}::spmm_compact", mutates_args=())
def spmm_compact(rp: torch.Tensor, col: torch.Tensor, masks: torch.Tensor, block_base: torch.Tensor,
                 values: torch.Tensor, window_of_part: torch.Tensor, atomic: torch.Tensor, X: torch.Tensor,
                 m_ori: int, window: int, mode: str, out_dtype: torch.dtype, rows_covered: bool) -> torch.Tensor:
    """Y[m_ori, D] = A . X over compact row-mask tiles (attgnn_spmm_compact)."""
    return ops.spmm_compact(rp, col, masks, block_base, values, window_of_part, atomic, X, m_ori, window,
                            mode=_mode(mode), out_dtype=out_dtype, rows_covered=rows_covered)


@spmm_compact.register_fake
def _(rp, col, masks, block_base, values, window_of_part, atomic, X, m_ori, window, mode, out_dtype, rows_covered):
    return X.new_empty((m_ori, X.shape[1]), dtype=out_dtype)


@torch.library.custom_op(f"{_LIB}::spmm_edge_map", mutates_args=())
def spmm_edge_map(rp: torch.Tensor, col: torch.Tensor, emap: torch.Tensor, edge_val: torch.Tensor,
                  window_of_part: torch.Tensor, atomic: torch.Tensor, X: torch.Tensor, out_layout: str, window: int,
                  wide: int, m_ori: int, mode: str, x_layout: str, out_dtype: torch.dtype,
                  rows_covered: bool) -> torch.Tensor:
    """Edge-map SpMM (attgnn_spmm_edge_map); m_ori < 0 means the rows of X."""
    return ops.spmm_edge_map(rp, col, emap, edge_val, window_of_part, atomic, X, out_layout, window, wide,
                             m_ori=None if m_ori < 0 else m_ori, mode=_mode(mode), x_layout=x_layout,
                             out_dtype=out_dtype, rows_covered=rows_covered)


@spmm_edge_map.register_fake
def _(rp, col, emap, edge_val, window_of_part, atomic, X, out_layout, window, wide, m_ori, mode, x_layout, out_dtype,
      rows_covered):
    heads, n, dim = _hmd(X, x_layout)
    m = n if m_ori < 0 else m_ori
    shape = (heads, m, dim) if out_layout == "hmd" else (m, heads, dim)
    return X.new_empty(shape, dtype=out_dtype)


@torch.library.custom_op(f"{_LIB}::spmm_edge_map_compact", mutates_args=())
def spmm_edge_map_compact(rp: torch.Tensor, col: torch.Tensor, masks: torch.Tensor, block_base: torch.Tensor,
                          ids: torch.Tensor, edge_val: torch.Tensor, window_of_part: torch.Tensor, atomic: torch.Tensor,
                          X: torch.Tensor, out_layout: str, window: int, m_ori: int, mode: str, x_layout: str,
                          out_dtype: torch.dtype, rows_covered: bool) -> torch.Tensor:
    """Edge-map SpMM over the compact edge-id list (attgnn_spmm_edge_map_compact);
    m_ori < 0 means the rows of X."""
    return ops.spmm_edge_map_compact(rp, col, masks, block_base, ids, edge_val, window_of_part, atomic, X, out_layout,
                                     window, m_ori=None if m_ori < 0 else m_ori, mode=_mode(mode), x_layout=x_layout,
                                     out_dtype=out_dtype, rows_covered=rows_covered)


@spmm_edge_map_compact.register_fake
def _(rp, col, masks, block_base, ids, edge_val, window_of_part, atomic, X, out_layout, window, m_ori, mode, x_layout,
      out_dtype, rows_covered):
    heads, n, dim = _hmd(X, x_layout)
    m = n if m_ori < 0 else m_ori
    shape = (heads, m, dim) if out_layout == "hmd" else (m, heads, dim)
    return X.new_empty(shape, dtype=out_dtype)


@torch.library.custom_op(f"{_LIB}::spmm_edge_map_compact_dot", mutates_args=())
def spmm_edge_map_compact_dot(rp: torch.Tensor, col: torch.Tensor, masks: torch.Tensor, block_base: torch.Tensor,
                              ids: torch.Tensor, edge_val: torch.Tensor, window_of_part: torch.Tensor,
                              atomic: torch.Tensor, X: torch.Tensor, dot_x: torch.Tensor, out_layout: str,
                              window: int, m_ori: int, x_layout: str, out_dtype: torch.dtype,
                              rows_covered: bool) -> Tuple[torch.Tensor, torch.Tensor]:
    """The compact edge-map SpMM with each edge's dot product <dot_x[window row], X[column]>
    (attgnn_spmm_edge_map_compact_dot): (Y, dot [H, nnz] fp32); m_ori < 0 means the rows of X."""
    return ops.spmm_edge_map_compact_dot(rp, col, masks, block_base, ids, edge_val, window_of_part, atomic, X, dot_x,
                                         out_layout, window, m_ori=None if m_ori < 0 else m_ori, x_layout=x_layout,
                                         out_dtype=out_dtype, rows_covered=rows_covered)


@spmm_edge_map_compact_dot.register_fake
def _(rp, col, masks, block_base, ids, edge_val, window_of_part, atomic, X, dot_x, out_layout, window, m_ori,
      x_layout, out_dtype, rows_covered):
    heads, n, dim = _hmd(X, x_layout)
    m = n if m_ori < 0 else m_ori
    shape = (heads, m, dim) if out_layout == "hmd" else (m, heads, dim)
    nnz = edge_val.shape[-1]
    return X.new_empty(shape, dtype=out_dtype), X.new_empty((heads, nnz), dtype=torch.float32)


@torch.library.custom_op(f"{_LIB}::spmm_attention", mutates_args=())
def spmm_attention(rp: torch.Tensor, col: torch.Tensor, emap: torch.Tensor, window_of_part: torch.Tensor,
                   atomic: torch.Tensor, X: torch.Tensor, out_layout: str, window: int, wide: int,
                   row_logit: torch.Tensor, col_logit: torch.Tensor, row_offset: Optional[torch.Tensor],
                   col_offset: Optional[torch.Tensor], slope: float, drop_p: float, seed: int, m_ori: int, mode: str,
                   x_layout: str, out_dtype: torch.dtype, rows_covered: bool) -> torch.Tensor:
    """Attention aggregation with the probabilities rebuilt in the kernel (attgnn_spmm_attention);
    logits node-major [N, H], offsets [H, N]; m_ori < 0 means the rows of X."""
    return ops.spmm_attention(rp, col, emap, window_of_part, atomic, X, out_layout, window, wide, row_logit,
                              col_logit, row_offset=row_offset, col_offset=col_offset, slope=slope, drop_p=drop_p,
                              seed=seed, m_ori=None if m_ori < 0 else m_ori, mode=_mode(mode), x_layout=x_layout,
                              out_dtype=out_dtype, rows_covered=rows_covered)


@spmm_attention.register_fake
def _(rp, col, emap, window_of_part, atomic, X, out_layout, window, wide, row_logit, col_logit, row_offset,
      col_offset, slope, drop_p, seed, m_ori, mode, x_layout, out_dtype, rows_covered):
    heads, n, dim = _hmd(X, x_layout)
    m = n if m_ori < 0 else m_ori
    shape = (heads, m, dim) if out_layout == "hmd" else (m, heads, dim)
    return X.new_empty(shape, dtype=out_dtype)


# ---------------------------------------------------------------- SDDMM and attention

@torch.library.custom_op(f"{_LIB}::sddmm_edges", mutates_args=())
def sddmm_edges(rp: torch.Tensor, col: torch.Tensor, block_base: torch.Tensor, edge_ids: torch.Tensor,
                positions: torch.Tensor, window_of_part: torch.Tensor, A: torch.Tensor, B: torch.Tensor, nnz: int,
                window: int, layout: str, out_dtype: torch.dtype, covered: bool) -> torch.Tensor:
    """out[h, e=(i,j)] = <A_h[i], B_h[j]> over a compact edge list (attgnn_sddmm_edges)."""
    return ops.sddmm_edges(rp, col, block_base, edge_ids, positions, window_of_part, A, B, nnz, window, layout=layout,
                           out_dtype=out_dtype, covered=covered)


@sddmm_edges.register_fake
def _(rp, col, block_base, edge_ids, positions, window_of_part, A, B, nnz, window, layout, out_dtype, covered):
    heads, _, _ = _hmd(A, layout)
    return A.new_empty((heads, nnz), dtype=out_dtype)


@torch.library.custom_op(f"{_LIB}::node_scores", mutates_args=())
def node_scores(X: torch.Tensor, vec: torch.Tensor, layout: str, out_layout: str = "hn") -> torch.Tensor:
    """s[h, i] = X_h[i] . vec_h (attgnn_node_scores): [H, N] ("hn") or [N, H] ("nh")."""
    return ops.node_scores(X, vec, layout, out_layout=out_layout)


@node_scores.register_fake
def _(X, vec, layout, out_layout="hn"):
    heads, n, _ = _hmd(X, layout)
    return X.new_empty((heads, n) if out_layout == "hn" else (n, heads), dtype=torch.float32)


def _sdims(s0, s_layout):
    return tuple(s0.shape) if s_layout == "hn" else tuple(s0.shape)[::-1]


@torch.library.custom_op(f"{_LIB}::edge_softmax_fwd", mutates_args=())
def edge_softmax_fwd(rowptr: torch.Tensor, col: torch.Tensor, s0: torch.Tensor, s1: torch.Tensor, slope: float,
                     drop_p: float, seed: int, p_dtype: torch.dtype, s_layout: str = "hn",
                     long_rows: Optional[torch.Tensor] = None) -> tuple[torch.Tensor, torch.Tensor]:
    """Fused LeakyReLU + row softmax + dropout (attgnn_edge_softmax_fwd): (p, lse)."""
    return ops.edge_softmax_fwd(rowptr, col, s0, s1, slope, drop_p, seed, p_dtype=p_dtype, s_layout=s_layout,
                                long_rows=long_rows)


@edge_softmax_fwd.register_fake
def _(rowptr, col, s0, s1, slope, drop_p, seed, p_dtype, s_layout="hn", long_rows=None):
    heads, n = _sdims(s0, s_layout)
    return s0.new_empty((heads, col.shape[0]), dtype=p_dtype), s0.new_empty((heads, n))


@torch.library.custom_op(f"{_LIB}::edge_softmax_bwd", mutates_args=())
def edge_softmax_bwd(rowptr: torch.Tensor, col: torch.Tensor, s0: torch.Tensor, s1: torch.Tensor, lse: torch.Tensor,
                     dp: torch.Tensor, slope: float, drop_p: float, seed: int, s_layout: str = "hn",
                     long_rows: Optional[torch.Tensor] = None) -> tuple[torch.Tensor, torch.Tensor]:
    """Backward of edge_softmax_fwd (attgnn_edge_softmax_bwd): (dz, ds0)."""
    return ops.edge_softmax_bwd(rowptr, col, s0, s1, lse, dp, slope, drop_p, seed, s_layout=s_layout,
                                long_rows=long_rows)


@edge_softmax_bwd.register_fake
def _(rowptr, col, s0, s1, lse, dp, slope, drop_p, seed, s_layout="hn", long_rows=None):
    heads, n = _sdims(s0, s_layout)
    return s0.new_empty((heads, col.shape[0])), s0.new_empty((heads, n))


@torch.library.custom_op(f"{_LIB}::edge_softmax_bwd_sums", mutates_args=())
def edge_softmax_bwd_sums(rowptr: torch.Tensor, col: torch.Tensor, s0: torch.Tensor, s1: torch.Tensor,
                          lse: torch.Tensor, dp: torch.Tensor, slope: float, drop_p: float, seed: int,
                          t_rowptr: torch.Tensor, t_perm: torch.Tensor, s_layout: str = "hn",
                          long_rows: Optional[torch.Tensor] = None,
                          t_long_rows: Optional[torch.Tensor] = None) -> tuple[torch.Tensor, torch.Tensor]:
    """Row and column sums of the softmax backward's dz (attgnn_edge_softmax_bwd_sums): (ds0, ds1)."""
    return ops.edge_softmax_bwd_sums(rowptr, col, s0, s1, lse, dp, slope, drop_p, seed, t_rowptr, t_perm,
                                     s_layout=s_layout, long_rows=long_rows, t_long_rows=t_long_rows)


@edge_softmax_bwd_sums.register_fake
def _(rowptr, col, s0, s1, lse, dp, slope, drop_p, seed, t_rowptr, t_perm, s_layout="hn", long_rows=None,
      t_long_rows=None):
    heads, n = _sdims(s0, s_layout)
    return s0.new_empty((heads, n)), s0.new_empty((heads, n))


@torch.library.custom_op(f"{_LIB}::gather_rowsum", mutates_args=())
def gather_rowsum(rowptr: torch.Tensor, perm: torch.Tensor, vals: torch.Tensor, n: int,
                  long_rows: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Column sums through the transpose permutation (attgnn_gather_rowsum)."""
    return ops.gather_rowsum(rowptr, perm, vals, n, long_rows=long_rows)


@gather_rowsum.register_fake
def _(rowptr, perm, vals, n, long_rows=None):
    return vals.new_empty((vals.shape[0], n), dtype=torch.float32)


@torch.library.custom_op(f"{_LIB}::attention_node_bwd", mutates_args=())
def attention_node_bwd(Xs: torch.Tensor, Xt: torch.Tensor, a0: torch.Tensor, a1: torch.Tensor, ds0: torch.Tensor,
                       ds1: torch.Tensor, base: Optional[torch.Tensor],
                       combine: bool) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Fused node-term backward (attgnn_attention_node_bwd): (dXs, dXt, da0, da1);
    dXs is an empty tensor when combine is set (its gradient is inside dXt)."""
    dXs, dXt, da0, da1 = ops.attention_node_bwd(Xs, Xt, a0, a1, ds0, ds1, base=base, combine=combine)
    return (Xs.new_empty(0) if dXs is None else dXs), dXt, da0, da1


@attention_node_bwd.register_fake
def _(Xs, Xt, a0, a1, ds0, ds1, base, combine):
    n, heads, dim = Xs.shape
    dXs = Xs.new_empty(0) if combine else torch.empty_like(Xs)
    return dXs, torch.empty_like(Xt), Xs.new_empty((heads, dim), dtype=torch.float32), \
        Xs.new_empty((heads, dim), dtype=torch.float32)


OPS = ("spmm_block", "spmm_compact", "spmm_edge_map", "spmm_edge_map_compact", "spmm_edge_map_compact_dot",
       "spmm_attention", "sddmm_edges", "node_scores", "edge_softmax_fwd", "edge_softmax_bwd", "edge_softmax_bwd_sums",
       "gather_rowsum", "attention_node_bwd")
