"""Graph datasets in the reference's ``.npz`` format and the dataset objects its
training scripts pass around as ``inputInfo``.

File format (writer: benchmark-v3/test.py:44-57; readers:
end2end_ori/gcn_final/mgat_csr/mdataset_fp16.py:27-53 and the GAT twin):
``num_nodes`` (scalar), ``src_li`` / ``dst_li`` (edge endpoints, row = src),
optionally ``features`` [N, F], ``labels`` [N], ``train_mask`` / ``val_mask`` /
``test_mask`` [N] bool, ``in_size``, ``out_size``.  Files are read with
``numpy.load(allow_pickle=False)``.

``MGAT_dataset_csr`` mirrors the reference's dataset classes
(end2end_ori/gcn_final/mgat_csr/mdataset_fp16.py:21-87,
end2end_ori/gat_final/mgat_csr/mdataset_fp16.py:21-92):
  * the CSR of the COO edge list (sorted columns, repeated edges merged, as
    scipy's ``coo_matrix(...).tocsr()`` gives the reference's structure);
  * ``num_nodes = n + 8 - n % 8`` and ``num_edges = len(src_li)``;
  * GCN: tile values ``Half(float(d_i) * float(d_j))`` with
    ``d = rsqrt(row count).half()`` of the CSR as loaded (no self-loops added,
    mdataset_fp16.py:54-56, gcn/Block_gcn/example.cpp:105);
  * GAT: edge-id maps of A and of A^T;
  * ``x`` ~ N(0, 1) in FP16 and ``y`` uniform in [0, num_classes) from torch's
    global generator when the file holds no features / labels.  Like the
    reference, ``featuredim`` and ``classes`` do not set the random sizes:
    128 features, 16 classes (GCN) or 10 classes (GAT).
The window blocks are this library's (16-row windows, 32-column blocks, built
on the GPU by default); the reference's call uses 8-row windows and
``wide = 8`` (mdataset_fp16.py:63), which ``window=8, wide=8`` reproduces.
"""
import os

import numpy as np
import torch

from . import graphs as G
from .layers import GATGraph, GCNGraph

NPZ_KEYS = ("num_nodes", "src_li", "dst_li", "features", "labels", "train_mask", "val_mask", "test_mask",
            "in_size", "out_size")


def save_npz(path, num_nodes, src, dst, features=None, labels=None, train_mask=None, val_mask=None,
             test_mask=None):
    """Write a graph in the reference's .npz layout (benchmark-v3/test.py:44-57)."""
    arrays = {"num_nodes": np.int64(num_nodes), "src_li": np.asarray(src, np.int64),
              "dst_li": np.asarray(dst, np.int64)}
    if features is not None:
        arrays["features"] = np.asarray(features, np.float32)
        arrays["in_size"] = np.int64(arrays["features"].shape[1])
    if labels is not None:
        arrays["labels"] = np.asarray(labels, np.int64)
        arrays["out_size"] = np.int64(int(arrays["labels"].max()) + 1 if arrays["labels"].size else 0)
    for name, m in (("train_mask", train_mask), ("val_mask", val_mask), ("test_mask", test_mask)):
        if m is not None:
            arrays[name] = np.asarray(m, bool)
    np.savez(path, **arrays)


def load_npz(path):
    """The arrays of a reference-format .npz (pickled objects are refused)."""
    with np.load(path, allow_pickle=False) as f:
        out = {k: f[k] for k in f.files if k in NPZ_KEYS}
    for k in ("num_nodes", "src_li", "dst_li"):
        if k not in out:
            raise ValueError(f"{path}: missing '{k}'")
    if out["src_li"].shape != out["dst_li"].shape or out["src_li"].ndim != 1:
        raise ValueError(f"{path}: src_li and dst_li must be 1-D of one length")
    return out


def csr_from_edges(src, dst, n):
    """CSR (rows = src) with ascending columns and repeated edges merged."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    if src.size and (src.min() < 0 or dst.min() < 0 or src.max() >= n or dst.max() >= n):
        raise ValueError("edge endpoints outside [0, num_nodes)")
    return G._from_keys(np.unique(src * n + dst), n)


def reference_degrees(csr):
    """d = rsqrt(row count) in FP32, rounded to FP16 (mdataset_fp16.py:54-56);
    rows without entries give inf, as in the reference."""
    return torch.rsqrt(torch.from_numpy(csr.degrees().astype(np.float32))).to(torch.float16).numpy()


def reference_tile_values(csr, d16):
    """Per-entry FP32 products float(d_i) * float(d_j) of FP16 degrees; rounded
    to FP16 in the tiles they equal Half(float(d[j]) * float(d[col]))
    (gcn/Block_gcn/example.cpp:105)."""
    d = np.asarray(d16, np.float16).astype(np.float32)
    return (d[csr.row_ids()] * d[csr.indices]).astype(np.float32)


class MGAT_dataset_csr:
    """Dataset object of the reference's training scripts.

    ``data``: a dataset name (read from ``<root>/<data>.npz``), a path to an
    .npz, or a dict of the .npz arrays.  ``kind``: "gcn" (window blocks of the
    normalised adjacency; the object is also a :class:`GCNGraph` through
    ``self.graph``) or "gat" (edge-id maps; :class:`GATGraph`).  The layers
    take the dataset itself as ``inputInfo``.
    """

    def __init__(self, data, featuredim=None, classes=None, root=".", kind="gcn", device=None, build="device",
                 window=16, wide=32, part_size=None):
        if kind not in ("gcn", "gat"):
            raise ValueError(f"kind must be 'gcn' or 'gat', got {kind!r}")
        if isinstance(data, dict):
            self.graph_arrays = data
        else:
            path = data if str(data).endswith(".npz") else os.path.join(root, f"{data}.npz")
            self.graph_arrays = load_npz(path)
        self.kind = kind
        self.device = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.featuredim, self.classes = featuredim, classes   # unused, as in the reference
        self.num_features = 128
        self.num_classes = 16 if kind == "gcn" else 10
        self.init_edges(build, window, wide, part_size)
        self.init_embedding()
        self.init_labels()
        self.init_masks()

    # -- mdataset_fp16.py:39-64
    def init_edges(self, build, window, wide, part_size):
        g = self.graph_arrays
        src, dst = g["src_li"], g["dst_li"]
        self.num_nodes_ori = int(g["num_nodes"])
        self.num_nodes = self.num_nodes_ori + 8 - self.num_nodes_ori % 8   # the reference's padded count
        self.num_edges = int(src.size)
        self.edge_index = np.stack([src, dst])
        self.csr = csr_from_edges(src, dst, self.num_nodes_ori)
        if self.kind == "gcn":
            d16 = reference_degrees(self.csr)
            self.graph = GCNGraph(self.csr, self.device, window=window, wide=wide, part_size=part_size,
                                  tile_dtype=torch.float16, values=reference_tile_values(self.csr, d16),
                                  build=build)
        else:
            self.graph = GATGraph(self.csr, self.device, window=window, wide=wide, part_size=part_size, build=build)

    # -- mdataset_fp16.py:67-81
    def init_embedding(self):
        f = self.graph_arrays.get("features")
        if f is not None:
            self.x = torch.from_numpy(np.asarray(f, np.float32)).to(torch.float16)
            self.num_features = int(self.x.shape[1])
        else:
            self.x = torch.randn(self.num_nodes_ori, self.num_features).to(dtype=torch.float16)

    def init_labels(self):
        lab = self.graph_arrays.get("labels")
        if lab is not None:
            self.y = torch.from_numpy(np.asarray(lab, np.int64))
            self.num_classes = int(self.graph_arrays.get("out_size", int(self.y.max()) + 1))
        else:
            self.y = torch.randint(low=0, high=self.num_classes, size=(self.num_nodes_ori,))

    def init_masks(self):
        for name in ("train_mask", "val_mask", "test_mask"):
            m = self.graph_arrays.get(name)
            setattr(self, name, None if m is None else torch.from_numpy(np.asarray(m, bool)))

    def to(self, device):
        """Features, labels and masks to `device` (the structure already lives
        on the device it was built on)."""
        self.x = self.x.to(device)
        self.y = self.y.to(device)
        for name in ("train_mask", "val_mask", "test_mask"):
            m = getattr(self, name)
            if m is not None:
                setattr(self, name, m.to(device))
        return self

    # the layers call these on inputInfo
    def __getattr__(self, name):
        graph = self.__dict__.get("graph")
        if graph is None:
            raise AttributeError(name)
        return getattr(graph, name)


def synthetic_npz(path, n, mean_degree, num_features=128, num_classes=16, seed=0, gamma=2.1, train_frac=0.6,
                  val_frac=0.2, target_nnz=None):
    """A Chung-Lu graph (symmetric, self-loops; ~target_nnz entries if given)
    written as a reference-format .npz with N(0,1) features, random labels and
    random split masks."""
    g = G.chung_lu(n, mean_degree, gamma, seed=seed, target_nnz=target_nnz)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    feats = rng.standard_normal((n, num_features)).astype(np.float32)
    labels = rng.integers(0, num_classes, n)
    u = rng.random(n)
    save_npz(path, n, g.row_ids(), g.indices, feats, labels, u < train_frac,
             (u >= train_frac) & (u < train_frac + val_frac), u >= train_frac + val_frac)
    return g
