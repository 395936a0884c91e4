"""The reference's .npz format and dataset semantics, on the CPU: round trip,
COO -> CSR (sorted, repeated edges merged, as scipy), the FP16 rsqrt degrees,
and tile values equal to the reference formula Half(float(d_i) * float(d_j))
through the host build."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from att_amd import datasets as D
from att_amd import ops


def test_npz_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    n = 50
    src, dst = rng.integers(0, n, 300), rng.integers(0, n, 300)
    feats = rng.standard_normal((n, 7)).astype(np.float32)
    labels = rng.integers(0, 5, n)
    masks = [rng.random(n) < 0.5 for _ in range(3)]
    path = tmp_path / "g.npz"
    D.save_npz(path, n, src, dst, feats, labels, *masks)
    g = D.load_npz(path)
    assert int(g["num_nodes"]) == n and int(g["in_size"]) == 7 and int(g["out_size"]) == int(labels.max()) + 1
    assert np.array_equal(g["src_li"], src) and np.array_equal(g["dst_li"], dst)
    assert np.array_equal(g["features"], feats) and np.array_equal(g["labels"], labels)
    for k, m in zip(("train_mask", "val_mask", "test_mask"), masks):
        assert np.array_equal(g[k], m)


def test_npz_refuses_pickles_and_missing_keys(tmp_path):
    bad = tmp_path / "bad.npz"
    np.savez(bad, num_nodes=np.int64(3), src_li=np.array([0, 1], dtype=object), dst_li=np.array([1, 2]))
    with pytest.raises(ValueError):
        D.load_npz(bad)
    miss = tmp_path / "miss.npz"
    np.savez(miss, num_nodes=np.int64(3), src_li=np.array([0, 1]))
    with pytest.raises(ValueError):
        D.load_npz(miss)


def test_csr_from_edges_matches_scipy():
    rng = np.random.default_rng(1)
    n = 200
    src, dst = rng.integers(0, n, 3000), rng.integers(0, n, 3000)   # repeats included
    g = D.csr_from_edges(src, dst, n)
    ref = sp.coo_matrix((np.ones(src.size), (src, dst)), shape=(n, n)).tocsr()
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(g.indptr, ref.indptr) and np.array_equal(g.indices, ref.indices)
    with pytest.raises(ValueError):
        D.csr_from_edges([0, 5], [1, 2], 5)


def test_reference_degrees_and_tile_values():
    rng = np.random.default_rng(2)
    n = 300
    src, dst = rng.integers(0, n, 2500), rng.integers(0, n, 2500)
    g = D.csr_from_edges(np.concatenate([src, dst]), np.concatenate([dst, src]), n)
    d16 = D.reference_degrees(g)
    counts = np.diff(g.indptr)
    want = torch.rsqrt(torch.tensor(counts.tolist(), dtype=torch.float32)).to(torch.float16).numpy()
    assert np.array_equal(d16.view(np.uint16), want.view(np.uint16))
    assert np.isinf(d16[counts == 0]).all()
    vals = D.reference_tile_values(g, d16)
    rp = g.indptr.astype(np.int32)
    a = ops.blocks_build(rp, g.indices, "edge_f32", vals, torch.float16, 8, 8, 32, emit_empty=False)
    b = ops.blocks_build(rp, g.indices, "deg_f16", d16, torch.float16, 8, 8, 32, emit_empty=False)
    for k in ("rp", "col", "val", "window", "atomic"):
        assert torch.equal(a[k].view(torch.int16) if k == "val" else a[k],
                           b[k].view(torch.int16) if k == "val" else b[k]), k
