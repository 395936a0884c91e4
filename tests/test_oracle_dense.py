"""Pins the FP64 oracle (oracle/dense.py) independently of its own formulas
(CPU only).

Every function is checked against brute-force Python loops over the edges of
small asymmetric inputs -- A != B, a0 != a1, X0 != X1, 1-3 heads, a
non-symmetric pattern with empty rows, a long row and a rectangular shape --
and against the reference tests' own formulation where one exists (row sums
as CSR . ones, test_spmm_h_m_ones.py:58-70; scipy products,
test_spmm_h_m.py:8-18).  The edge softmax is also checked against a dense
masked softmax in torch float64.  Each pin is then shown to fail for the
mistakes it exists to catch (a swapped operand, row/column or vector, a
dropped term, a transposed result), planted in copies of the oracle here.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import dense as O


def _graph(seed, n=23, n_cols=None, long_row=True):
    """Random non-symmetric CSR with empty rows (every 5th), one long row and
    sorted, duplicate-free columns; n_cols > n gives a rectangular matrix."""
    rng = np.random.default_rng(seed)
    n_cols = n_cols or n
    rows = []
    for i in range(n):
        if i % 5 == 3:
            rows.append(np.zeros(0, np.int64))
        elif long_row and i == 1:
            rows.append(np.arange(n_cols))
        else:
            rows.append(np.unique(rng.integers(0, n_cols, size=int(rng.integers(1, 6)))))
    indptr = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int64)
    indices = np.concatenate(rows).astype(np.int64)
    return indptr, indices, n_cols


def _edges(indptr, indices):
    for i in range(len(indptr) - 1):
        for e in range(indptr[i], indptr[i + 1]):
            yield e, i, int(indices[e])


def _brute_spmm(indptr, indices, vals, X, n_rows):
    Y = np.zeros((n_rows, X.shape[1]))
    for e, i, j in _edges(indptr[: n_rows + 1], indices):
        for f in range(X.shape[1]):
            Y[i, f] += vals[e] * X[j, f]
    return Y


def test_graph_is_asymmetric_and_has_empty_and_long_rows():
    indptr, indices, _ = _graph(0)
    A = sp.csr_matrix((np.ones(indices.size), indices, indptr), shape=(23, 23))
    assert (A != A.T).nnz > 0
    assert np.any(np.diff(indptr) == 0) and np.diff(indptr).max() == 23


@pytest.mark.parametrize("seed", [0, 1])
def test_spmm_vs_loops(seed):
    indptr, indices, n_cols = _graph(seed, n_cols=31)
    rng = np.random.default_rng(seed + 10)
    vals = rng.standard_normal(indices.size)
    X = rng.standard_normal((n_cols, 5))
    assert np.allclose(O.spmm(indptr, indices, vals, X), _brute_spmm(indptr, indices, vals, X, 23), atol=1e-12)
    # the first rows only (the windowing's trailing rows, n_rows < M)
    assert np.allclose(O.spmm(indptr, indices, vals, X, n_rows=17), _brute_spmm(indptr, indices, vals, X, 17),
                       atol=1e-12)
    # planted mistake: A^T . X instead of A . X (square case) fails the pin
    indptr, indices, _ = _graph(seed)
    vals = rng.standard_normal(indices.size)
    X = rng.standard_normal((23, 5))
    At = sp.csr_matrix((vals, indices, indptr), shape=(23, 23)).T.tocsr()
    assert not np.allclose(O.spmm(At.indptr, At.indices, At.data, X), _brute_spmm(indptr, indices, vals, X, 23))


@pytest.mark.parametrize("heads", [1, 2, 3])
def test_spmm_heads_vs_loops(heads):
    indptr, indices, _ = _graph(heads)
    rng = np.random.default_rng(heads)
    ev = rng.standard_normal((heads, indices.size))
    X = rng.standard_normal((heads, 23, 4))
    got = O.spmm_heads(indptr, indices, ev, X)
    for h in range(heads):
        assert np.allclose(got[h], _brute_spmm(indptr, indices, ev[h], X[h], 23), atol=1e-12)
    if heads > 1:   # planted mistake: head 0's values for every head
        bad = np.stack([O.spmm(indptr, indices, ev[0], X[h]) for h in range(heads)])
        assert not np.allclose(bad, got)


def test_rowsum_is_csr_times_ones():
    """test_spmm_h_m_ones.py:58-70: the row sum of per-edge values is the
    CSR product with a ones column (with 1-D and [H, nnz] values)."""
    indptr, indices, _ = _graph(3)
    rng = np.random.default_rng(3)
    ev = rng.standard_normal((3, indices.size))
    got = O.rowsum(indptr, ev)
    for h in range(3):
        want = O.spmm(indptr, indices, ev[h], np.ones((23, 1)))[:, 0]
        assert np.allclose(got[h], want, atol=1e-12)
        loops = np.zeros(23)
        for e, i, j in _edges(indptr, indices):
            loops[i] += ev[h, e]
        assert np.allclose(got[h], loops, atol=1e-12)
    assert got[0][3] == 0.0                                   # an empty row sums to zero
    assert np.allclose(O.rowsum(indptr, ev[1]), got[1:2])
    # planted mistake: column sums instead of row sums
    colsum = np.zeros(23)
    np.add.at(colsum, indices, ev[0])
    assert not np.allclose(colsum, got[0])


@pytest.mark.parametrize("heads", [1, 3])
def test_sddmm_vs_loops(heads):
    indptr, indices, n_cols = _graph(heads + 4, n_cols=29)
    rng = np.random.default_rng(heads + 4)
    A = rng.standard_normal((heads, 23, 6))
    B = rng.standard_normal((heads, n_cols, 6))
    got = O.sddmm(indptr, indices, A, B)
    assert got.shape == (heads, indices.size)
    for h in range(heads):
        for e, i, j in _edges(indptr, indices):
            assert abs(got[h, e] - sum(A[h, i, f] * B[h, j, f] for f in range(6))) < 1e-12
    # planted mistake: rows and columns swapped (<A[j], B[i]>) on a square pattern
    indptr, indices, _ = _graph(heads + 4)
    A = rng.standard_normal((heads, 23, 6))
    B = rng.standard_normal((heads, 23, 6))
    rows = np.repeat(np.arange(23), np.diff(indptr))
    swapped = np.einsum("hed,hed->he", A[:, indices], B[:, rows])
    assert not np.allclose(swapped, O.sddmm(indptr, indices, A, B))


@pytest.mark.parametrize("heads", [1, 2, 3])
def test_att_vs_loops(heads):
    indptr, indices, _ = _graph(heads + 7)
    rng = np.random.default_rng(heads + 7)
    X = rng.standard_normal((heads, 23, 5))
    a0 = rng.standard_normal((heads, 1, 5))
    a1 = rng.standard_normal((heads, 1, 5))
    got = O.att(indptr, indices, X, a0, a1)
    for h in range(heads):
        for e, i, j in _edges(indptr, indices):
            want = sum(a0[h, 0, f] * X[h, i, f] + a1[h, 0, f] * X[h, j, f] for f in range(5))
            assert abs(got[h, e] - want) < 1e-12
    assert not np.allclose(O.att(indptr, indices, X, a1, a0), got)       # planted: a0 / a1 swapped


@pytest.mark.parametrize("heads", [1, 3])
def test_attv2_vs_loops(heads):
    indptr, indices, _ = _graph(heads + 11)
    rng = np.random.default_rng(heads + 11)
    X0 = rng.standard_normal((heads, 23, 5))
    X1 = rng.standard_normal((heads, 23, 5))
    a = rng.standard_normal((heads, 1, 5))
    got = O.attv2(indptr, indices, X0, X1, a)
    for h in range(heads):
        for e, i, j in _edges(indptr, indices):
            want = sum(a[h, 0, f] * (X0[h, i, f] + X1[h, j, f]) for f in range(5))
            assert abs(got[h, e] - want) < 1e-12
    assert not np.allclose(O.attv2(indptr, indices, X1, X0, a), got)     # planted: X0 / X1 swapped


@pytest.mark.parametrize("heads", [1, 3])
def test_ones_vs_loops(heads):
    indptr, indices, _ = _graph(heads + 13)
    g = np.random.default_rng(heads).standard_normal((23, heads))
    got = O.ones(indptr, g)
    assert got.shape == (heads, indices.size)
    for h in range(heads):
        for e, i, j in _edges(indptr, indices):
            assert got[h, e] == g[i, h]
    bad = g[indices].T                                                     # planted: the column's value
    assert not np.allclose(bad, got)


def test_leaky_relu():
    x = np.array([-3.0, -0.5, 0.0, 0.25, 4.0])
    assert np.array_equal(O.leaky_relu(x), [-0.6000000000000001, -0.1, 0.0, 0.25, 4.0])
    assert np.array_equal(O.leaky_relu(x, 0.01), [-0.03, -0.005, 0.0, 0.25, 4.0])


def _dense_masked_softmax(indptr, indices, logits):
    """torch float64: dense [H, N, N] logits, -inf off the pattern, softmax over
    the last axis; empty rows give no entries."""
    n = len(indptr) - 1
    h = logits.shape[0]
    dense = torch.full((h, n, n), float("-inf"), dtype=torch.float64)
    rows = np.repeat(np.arange(n), np.diff(indptr))
    dense[:, rows, indices] = torch.from_numpy(logits)
    sm = torch.softmax(dense, dim=-1)
    return sm[:, rows, indices].numpy()


@pytest.mark.parametrize("heads", [1, 3])
def test_edge_softmax_vs_dense_masked_softmax(heads):
    indptr, indices, _ = _graph(heads + 17)
    rng = np.random.default_rng(heads + 17)
    logits = rng.standard_normal((heads, indices.size)) * 4.0
    logits[:, indptr[1]:indptr[2]] += 30.0                # a long row of large logits: max subtraction matters
    got = O.edge_softmax(indptr, logits)
    assert np.allclose(got, _dense_masked_softmax(indptr, indices, logits), atol=1e-14)
    sums = O.rowsum(indptr, got)
    nonempty = np.diff(indptr) > 0
    assert np.allclose(sums[:, nonempty], 1.0) and np.all(sums[:, ~nonempty] == 0)
    # planted mistake: normalised over columns instead of rows
    t_indptr, t_vals, order = _transpose(indptr, indices, logits)
    colwise = np.empty_like(got)
    colwise[:, order] = O.edge_softmax(t_indptr, t_vals)
    assert not np.allclose(colwise, got)


def _transpose(indptr, indices, vals):
    A = sp.csr_matrix((np.arange(indices.size) + 1, indices, indptr), shape=(23, 23)).T.tocsr()
    order = A.data - 1
    return A.indptr, vals[:, order], order


def test_gat_layer_vs_dense_attention():
    """softmax_j(LeakyReLU(a0.x_i + a1.x_j)) x_j as dense masked attention in
    torch float64, heads with distinct vectors."""
    indptr, indices, _ = _graph(21, long_row=False)
    rng = np.random.default_rng(21)
    H, D = 2, 4
    Xp = rng.standard_normal((H, 23, D))
    a0 = rng.standard_normal((H, 1, D))
    a1 = rng.standard_normal((H, 1, D))
    got = O.gat_layer(indptr, indices, Xp, a0, a1)          # [N, H, D]
    X = torch.from_numpy(Xp)
    s0 = (X * torch.from_numpy(a0)).sum(-1)                 # [H, N]
    s1 = (X * torch.from_numpy(a1)).sum(-1)
    logits = torch.nn.functional.leaky_relu(s0[:, :, None] + s1[:, None, :], 0.2)
    mask = torch.zeros(23, 23, dtype=torch.bool)
    rows = np.repeat(np.arange(23), np.diff(indptr))
    mask[rows, indices] = True
    p = torch.softmax(logits.masked_fill(~mask, float("-inf")), dim=-1).nan_to_num(0.0)
    want = torch.bmm(p, X).permute(1, 0, 2).numpy()
    assert np.allclose(got, want, atol=1e-12)
    bad = O.gat_layer(indptr, indices, Xp, a1, a0)          # planted: a0 / a1 swapped
    assert not np.allclose(bad, want)


def test_gcn_layer_vs_dense():
    indptr, indices, _ = _graph(23)
    rng = np.random.default_rng(23)
    vals = rng.standard_normal(indices.size)
    X = rng.standard_normal((23, 6))
    W = rng.standard_normal((6, 3))
    A = np.zeros((23, 23))
    for e, i, j in _edges(indptr, indices):
        A[i, j] = vals[e]
    assert np.allclose(O.gcn_layer(indptr, indices, vals, X, W), A @ (X @ W), atol=1e-12)
    assert not np.allclose(O.gcn_layer(indptr, indices, vals, X, W), A.T @ (X @ W))


def test_rel_fro():
    ref = np.array([3.0, 4.0])
    assert O.rel_fro(ref, ref) == 0.0
    assert O.rel_fro(np.array([3.0, 4.5]), ref) == pytest.approx(0.1)
    assert O.rel_fro(np.array([0.5]), np.array([0.0])) == 0.5               # zero reference: absolute norm
