"""Striped node order and the owner-split SpMM (attgnn_spmm_compact_striped)
against the FP64 oracle: every owner count, the widths where the parts are
split by owner and where they are not, split windows, non-symmetric graphs,
host and device builds, bitwise run-to-run and across part sizes, and the
2-layer GCN step against the unstriped graph."""
import numpy as np
import pytest
import torch

from att_amd import graphs as G
from att_amd import ops
from att_amd.layers import GCNConv, GCNGraph
from oracle import dense as O

pytestmark = pytest.mark.gpu


def _graph(n, deg, seed, symmetric=True):
    g = G.chung_lu(n, deg, 2.1, seed=seed)
    if symmetric:
        return g
    rng = np.random.default_rng(seed)
    keep = rng.random(g.nnz) < 0.6
    rows = g.row_ids()[keep].astype(np.int64)
    keys = np.unique(np.concatenate([rows * n + g.indices[keep], np.arange(n) * (n + 1)]))
    ip = np.zeros(n + 1, np.int64)
    ip[1:] = np.cumsum(np.bincount(keys // n, minlength=n))
    return G.CSR(ip, (keys % n).astype(np.int32), n)


def _oracle(g, vals, X, dtype, transpose=False):
    """FP64 A X (or A^T X) over the tile values rounded as the GPU stores them."""
    v = torch.from_numpy(vals.astype(np.float32)).to(dtype).double().numpy()
    ip, ix = g.indptr, g.indices
    if transpose:
        import scipy.sparse as sp
        t = sp.csr_matrix((v, ix, ip), shape=(g.n, g.n)).T.tocsr()
        ip, ix, v = t.indptr, t.indices, t.data
    return O.spmm(ip, ix, v, X.double().cpu().numpy())


# owner groups x feature slices per part: 4x2, 4x2, 8x1, 2x4, 8x1, 4x1, 4x2 (two owners per task),
# 2x4 (four owners per task); 48 features do not fill whole lanes: the usual mapping
@pytest.mark.parametrize("owners,dim,dtype", [(4, 256, torch.float16), (4, 256, torch.bfloat16),
                                              (8, 128, torch.float16), (2, 512, torch.float16),
                                              (8, 64, torch.float16), (4, 64, torch.float16),
                                              (8, 256, torch.float16), (8, 512, torch.bfloat16),
                                              (32, 256, torch.float16), (4, 48, torch.bfloat16)])
def test_striped_spmm_vs_fp64(cuda, owners, dim, dtype):
    g = _graph(5000, 40, seed=owners + dim)
    vals = G.gcn_norm_values(g, G.gcn_norm_degree(g))
    graph = GCNGraph(g, cuda, tile_dtype=dtype, symmetric=True, values=vals, build="device", stripe=owners,
                     chunk_blocks=4)
    assert graph.num_nodes == G.stripe_rows(g.n, owners)
    torch.manual_seed(dim)
    X = torch.empty(g.n, dim).uniform_(-1, 1).to(dtype)
    Xs = graph.to_striped(X.to(cuda))
    Yp = graph.spmm(Xs, out_dtype=torch.float32)
    assert Yp.shape == (graph.num_nodes, dim)
    Y = graph.from_striped(Yp)
    ref = _oracle(g, vals, X, dtype)
    assert O.rel_fro(Y.double().cpu().numpy(), ref) < 2e-6
    # padding positions hold zero rows
    pad = torch.ones(graph.num_nodes, dtype=torch.bool, device=cuda)
    pad[graph.stripe_index] = False
    assert int(pad.sum()) == graph.num_nodes - g.n and not Yp[pad].any()
    # bitwise run to run
    assert torch.equal(Yp, graph.spmm(Xs, out_dtype=torch.float32))
    # every owner-group count (each task walking owners/groups segments), and
    # the usual task mapping, give the same sums up to rounding
    _, col, _, _, _ = graph.fwd["arrays"]
    rp, win, at = graph.fwd["tasks"]
    masks, base, values = graph.fwd["compact"]
    slices = dim // 128 if dim >= 128 else 1   # fp16 / bf16: 8 features per lane from 128 up
    groups = 1
    while groups <= min(owners, 8 // slices):
        Yg = ops.spmm_compact_striped(rp, col, masks, base, values, win, at, graph.fwd["oseg"], owners, Xs, 16,
                                      out_dtype=torch.float32, groups=groups)
        assert O.rel_fro(graph.from_striped(Yg).double().cpu().numpy(), ref) < 2e-6, groups
        assert torch.equal(Yg, ops.spmm_compact_striped(rp, col, masks, base, values, win, at, graph.fwd["oseg"],
                                                        owners, Xs, 16, out_dtype=torch.float32, groups=groups))
        groups *= 2
    Yu = ops.spmm_compact_striped(rp, col, masks, base, values, win, at, graph.fwd["oseg"], owners, Xs, 16,
                                  out_dtype=torch.float32, split=False)
    assert O.rel_fro(graph.from_striped(Yu).double().cpu().numpy(), ref) < 2e-6


@pytest.mark.parametrize("owners,dim", [(2, 512), (4, 256), (8, 128), (8, 256), (4, 64), (16, 256)])
def test_striped_bitwise_across_part_sizes_and_builds(cuda, owners, dim):
    """The SpMM runs over the canonical task table: part sizes 1, 2, 32 and
    unbounded give the same bits; the host build equals the device build."""
    g = _graph(3000, 60, seed=7 + owners + dim)
    vals = G.gcn_norm_values(g, G.gcn_norm_degree(g))
    X = torch.randn(g.n, dim).half().to(cuda)
    outs = []
    slices = dim // 128 if dim >= 128 else 1
    for ps, build in ((1, "device"), (2, "host"), (32, "device"), (1 << 20, "host")):
        graph = GCNGraph(g, cuda, tile_dtype=torch.float16, symmetric=True, values=vals, build=build,
                         part_size=ps, stripe=owners, chunk_blocks=3)
        _, col, _, _, _ = graph.fwd["arrays"]
        rp, win, at = graph.fwd["tasks"]
        masks, base, values = graph.fwd["compact"]
        Xs = graph.to_striped(X)
        outs.append(torch.stack([ops.spmm_compact_striped(rp, col, masks, base, values, win, at, graph.fwd["oseg"],
                                                          owners, Xs, 16, out_dtype=torch.float16, groups=gr)
                                 for gr in (1, min(owners, 8 // slices))]))
    for y in outs[1:]:
        assert torch.equal(y, outs[0])


def test_striped_transpose_nonsymmetric(cuda):
    g = _graph(4000, 30, seed=3, symmetric=False)
    vals = G.gcn_norm_values(g, G.gcn_norm_degree(g))
    for build in ("device", "host"):
        graph = GCNGraph(g, cuda, tile_dtype=torch.float16, values=vals, build=build, stripe=4)
        assert not graph.symmetric
        X = torch.randn(g.n, 256).half()
        Xs = graph.to_striped(X.to(cuda))
        for tr in (False, True):
            Y = graph.from_striped(graph.spmm(Xs, transpose=tr, out_dtype=torch.float32))
            ref = _oracle(g, vals, X, torch.float16, transpose=tr)
            assert O.rel_fro(Y.double().cpu().numpy(), ref) < 2e-6, (build, tr)


def test_striped_empty_and_single_window(cuda):
    """A graph of a few nodes (one window, most of it padding) and one with
    an isolated tail: every real row still matches the oracle."""
    for n in (5, 70):
        g = _graph(n, 3, seed=n)
        vals = G.gcn_norm_values(g, G.gcn_norm_degree(g))
        graph = GCNGraph(g, cuda, tile_dtype=torch.float16, symmetric=True, values=vals, build="device",
                         stripe=4)
        X = torch.randn(g.n, 256).half()
        Y = graph.from_striped(graph.spmm(graph.to_striped(X.to(cuda)), out_dtype=torch.float32))
        assert O.rel_fro(Y.double().cpu().numpy(), _oracle(g, vals, X, torch.float16)) < 2e-6


def test_striped_gcn_step_equals_unstriped(cuda):
    """Two GCN layers (256 wide, then 48 run at 64) forward and backward on
    the striped graph give the unstriped graph's loss and gradients up to
    rounding: node order only relabels the rows."""
    g = _graph(6000, 50, seed=11)
    vals = G.gcn_norm_values(g, G.gcn_norm_degree(g))
    X = torch.randn(g.n, 96).half().to(cuda)
    labels = torch.randint(0, 41, (g.n,)).to(cuda)
    res = []
    for stripe in (None, 4):
        graph = GCNGraph(g, cuda, tile_dtype=torch.float16, symmetric=True, values=vals, build="device",
                         stripe=stripe)
        torch.manual_seed(0)
        c1, c2 = GCNConv(96, 256).to(cuda), GCNConv(256, 41).to(cuda)
        h = torch.relu(c1(graph.to_striped(X), graph))
        out = graph.from_striped(c2(h, graph)).float().log_softmax(1)
        loss = -out.gather(1, labels[:, None]).mean()
        loss.backward()
        res.append((loss.item(), c1.weights.grad.clone(), c2.weights.grad.clone()))
    (l0, a0, b0), (l1, a1, b1) = res
    assert abs(l0 - l1) < 1e-3 * abs(l0)
    assert ((a0 - a1).norm() / a0.norm()).item() < 1e-2
    assert ((b0 - b1).norm() / b0.norm()).item() < 1e-2
