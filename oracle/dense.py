"""TEST INFRASTRUCTURE ONLY -- FP64 host oracle (numpy / scipy) for the kernels
and layers on the reference's sparse GNN path.

Every function is a closed form of what the reference computes, with the
reference call site it follows:

  spmm            Y = A.X                   gcn/SpMM/src/spmmKernel.cu:158-265
                                            (known answer gcn/SpMM/test/mgcn/test_balance.py:8-18)
  spmm_heads      Y_h = A_h.X_h             ATT_SpMM_v3.spmm_*_h_m / m_h
                                            (test_spmm_h_m.py:8-18,45-66)
  rowsum          s_h[i] = sum_j A_h[i,j]   ATT_SpMM_v3.*_rowsum (test_spmm_h_m_ones.py:45-70)
  sddmm           out_h[e] = <A_h[i],B_h[j]> ATT_SDDMM_v3.sddmm_*_h_nnz (test_sddmm_fp16.py:62-69)
  att             a0.X[i] + a1.X[j]         ATT_SDDMM_v3.att_*_h_nnz (test_sddmm_fp16_att.py:68-75)
  attv2           a.X0[i] + a.X1[j]         ATT_SDDMM_v3.attv2_*_h_nnz (test_sddmm_fp16_att_v2.py:70-77)
  ones            out_h[e] = g[i,h]         ATT_SDDMM_v3.sddmm_*_ones_h_nnz (test_sddmm_fp16_ones.py:62-68)
  gcn_layer       relu-free A.(X.W)         end2end_ori/gcn_final/mgat_csr/mgat_conv.py:29-55
  gat_layer       softmax_j(lrelu(a0.x_i + a1.x_j)) x_j
                                            end2end_head/gat_final/mgat_csr/mgat_conv.py:174-199
                  (softmax with max subtraction: mathematically the reference's
                  exp / rowsum, mgat_conv.py:189-197)

Pinned by tests/test_oracle_dense.py against scipy.sparse products, brute-force
loops on tiny inputs, and the reference tests' small-integer known answers.
"""
import numpy as np
import scipy.sparse as sp


def rel_fro(x, ref):
    """Relative Frobenius error ||x - ref|| / ||ref|| (FP64)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (den if den > 0 else 1.0))


def csr_matrix(indptr, indices, values, shape):
    return sp.csr_matrix((np.asarray(values, np.float64), np.asarray(indices), np.asarray(indptr)), shape=shape)


def spmm(indptr, indices, values, X, n_rows=None):
    """Y = A.X in FP64; A given as CSR with the given values."""
    n = len(indptr) - 1 if n_rows is None else n_rows
    X = np.asarray(X, np.float64)
    A = csr_matrix(indptr[: n + 1], indices[: indptr[n]], values[: indptr[n]], (n, X.shape[0]))
    return A @ X


def spmm_heads(indptr, indices, edge_values, X):
    """Y[h] = A_h.X[h]; edge_values [H, nnz]; X [H, N, D] -> [H, M, D]."""
    ev = np.asarray(edge_values, np.float64).reshape(X.shape[0], -1)
    return np.stack([spmm(indptr, indices, ev[h], X[h]) for h in range(X.shape[0])])


def rowsum(indptr, edge_values):
    """s[h, i] = sum of edge values of row i, [H, M]."""
    ev = np.asarray(edge_values, np.float64)
    ev = ev.reshape(1, -1) if ev.ndim == 1 else ev
    n = len(indptr) - 1
    rows = np.repeat(np.arange(n), np.diff(indptr))
    out = np.zeros((ev.shape[0], n))
    for h in range(ev.shape[0]):
        np.add.at(out[h], rows, ev[h])
    return out


def sddmm(indptr, indices, A, B):
    """out[h, e] = <A[h, i], B[h, j]> for e = (i, j); A, B [H, N, D] -> [H, nnz]."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    cols = np.asarray(indices)
    return np.einsum("hed,hed->he", A[:, rows], B[:, cols])


def att(indptr, indices, X, a0, a1):
    """out[h, e] = a0_h.X_h[i] + a1_h.X_h[j]; X [H,N,D], a0/a1 [H,1,D]."""
    X = np.asarray(X, np.float64)
    s0 = np.einsum("hnd,hd->hn", X, np.asarray(a0, np.float64).reshape(X.shape[0], -1))
    s1 = np.einsum("hnd,hd->hn", X, np.asarray(a1, np.float64).reshape(X.shape[0], -1))
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    return s0[:, rows] + s1[:, np.asarray(indices)]


def attv2(indptr, indices, X0, X1, a):
    X0 = np.asarray(X0, np.float64)
    X1 = np.asarray(X1, np.float64)
    av = np.asarray(a, np.float64).reshape(X0.shape[0], -1)
    s0 = np.einsum("hnd,hd->hn", X0, av)
    s1 = np.einsum("hnd,hd->hn", X1, av)
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    return s0[:, rows] + s1[:, np.asarray(indices)]


def ones(indptr, g):
    """out[h, e] = g[i, h] for e in row i; g [N, H] -> [H, nnz]."""
    g = np.asarray(g, np.float64).reshape(len(indptr) - 1 if np.asarray(g).shape[0] == len(indptr) - 1 else -1, -1)
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    return g[rows].T


def leaky_relu(x, slope=0.2):
    return np.where(x > 0, x, slope * x)


def edge_softmax(indptr, logits):
    """Row-wise softmax of per-edge logits [H, nnz] (max-subtracted, FP64)."""
    lg = np.asarray(logits, np.float64)
    n = len(indptr) - 1
    rows = np.repeat(np.arange(n), np.diff(indptr))
    out = np.empty_like(lg)
    for h in range(lg.shape[0]):
        mx = np.full(n, -np.inf)
        np.maximum.at(mx, rows, lg[h])
        ex = np.exp(lg[h] - mx[rows])
        s = np.zeros(n)
        np.add.at(s, rows, ex)
        out[h] = ex / s[rows]
    return out


def gat_layer(indptr, indices, Xp, a0, a1, slope=0.2):
    """Single GAT layer after the projection: Xp [H, N, D] -> [N, H, D]."""
    e = leaky_relu(att(indptr, indices, Xp, a0, a1), slope)
    p = edge_softmax(indptr, e)
    out = spmm_heads(indptr, indices, p, Xp)          # [H, N, D]
    return out.transpose(1, 0, 2)


def gcn_layer(indptr, indices, values, X, W):
    return spmm(indptr, indices, values, np.asarray(X, np.float64) @ np.asarray(W, np.float64))
