"""Differentiable sparse primitives on the window-block kernels.

``spmm(values, X, graph)``  Y[i] = sum_j A[i,j] X[j] with A's values given per
    edge in CSR order ([H, nnz]); X and Y are [N, H, D].
    backward: dX = A^T dY (edge-map SpMM over the A^T map, ids into A's edge
    order) and dvalues[e=(i,j)] = <dY[i], X[j]> (SDDMM).
``sddmm(A, B, graph)``  out[h, e=(i,j)] = <A[i,h], B[j,h]>, [H, nnz].
    backward: dA = G . B and dB = G^T . A -- two SpMMs with the incoming
    gradient G as edge values, over the A and A^T maps.

``graph`` is an ``att_amd.layers.GATGraph`` (edge-id maps of A and A^T).
Reference counterparts: ATT_SpMM_v3.spmm_*_m_h and ATT_SDDMM_v3.sddmm_*_h_nnz
as used by end2end_head/gat_final/mgat_csr/mgat_conv.py:72-119, whose backward
relies on A being symmetric; these use the transposed map, so any pattern works.
"""
import torch

from . import ops


def _emap(graph, transpose, values, X, n_out):
    rp, col, emap, win, at = graph.bwd if transpose else graph.fwd
    return ops.spmm_edge_map(rp, col, emap, values.contiguous(), win, at, X.contiguous(), "mhd", graph.window,
                             graph.wide, m_ori=n_out, x_layout="mhd")


def _sddmm(graph, A, B):
    return graph.sddmm(A, B)


class SpMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, X, graph):
        ctx.save_for_backward(values, X)
        ctx.graph = graph
        return _emap(graph, False, values.to(X.dtype), X, X.shape[0]).to(X.dtype)

    @staticmethod
    def backward(ctx, dY):
        values, X = ctx.saved_tensors
        g = ctx.graph
        dY = dY.to(X.dtype).contiguous()
        dX = dvals = None
        if ctx.needs_input_grad[1]:
            dX = _emap(g, True, values.to(X.dtype), dY, X.shape[0]).to(X.dtype)
        if ctx.needs_input_grad[0]:
            dvals = _sddmm(g, dY, X).to(values.dtype)
        return dvals, dX, None


class SDDMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B, graph):
        ctx.save_for_backward(A, B)
        ctx.graph = graph
        return _sddmm(graph, A, B)

    @staticmethod
    def backward(ctx, G):
        A, B = ctx.saved_tensors
        g = ctx.graph
        Gv = G.to(A.dtype).contiguous()
        dA = dB = None
        if ctx.needs_input_grad[0]:
            dA = _emap(g, False, Gv, B, A.shape[0]).to(A.dtype)
        if ctx.needs_input_grad[1]:
            dB = _emap(g, True, Gv, A, B.shape[0]).to(B.dtype)
        return dA, dB, None


def spmm(values, X, graph):
    return SpMM.apply(values, X, graph)


def sddmm(A, B, graph):
    return SDDMM.apply(A, B, graph)
