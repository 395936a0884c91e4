"""Training harness of the reference's end-to-end experiments, on this library.

Mirrors (same class / function names, arguments and loop):
  Net      (GCN)  end2end_ori/gcn_final/mgat_csr/gat_mgnn.py:17-40
  GATNet   (GAT)  end2end_ori/gat_final/mgat_csr/gat_mgnn.py:17-37 (there also named Net)
  train           gat_mgnn.py:55-67 -- Adam(lr 1e-2, weight decay 5e-4), NLL loss
                  over all nodes, one optimiser step per epoch
  test1           gat_mgnn.py:42-52 -- accuracy on test_mask
  test            end2end_ori/gcn_final/mgat_csr/test_mgat_csr.py:12-32 -- builds
                  the dataset and model, 10 warm-up epochs, then `epoches` timed
                  epochs; returns seconds
The plain FP32 PyTorch baselines (torch.sparse SpMM for GCN, gather /
scatter message passing for GAT) run the same architecture, loop and timing.

    python -m att_amd.train --model gcn --npz graph.npz --layers 3 --hidden 64 --epochs 300
    python -m att_amd.train --model gat --synthetic 169343,14 --epochs 300 --plain
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import datasets as D
from .layers import GATConv_multi, GCNConv


class Net(nn.Module):
    """GCN: conv1 -> ReLU -> dropout -> (hidden convs) -> conv2 -> log_softmax."""

    def __init__(self, in_feats, hidden_feats, out_feats, num_layers, dropout):
        super().__init__()
        self.conv1 = GCNConv(in_feats, hidden_feats)
        self.hidden_layers = nn.ModuleList(GCNConv(hidden_feats, hidden_feats) for _ in range(num_layers - 2))
        self.conv2 = GCNConv(hidden_feats, out_feats)
        self.num_layers, self.dropout = num_layers, dropout

    def forward(self, inputInfo):
        x = F.relu(self.conv1(inputInfo.x, inputInfo))
        x = F.dropout(x, self.dropout, training=self.training)
        for conv in self.hidden_layers:
            x = F.relu(conv(x, inputInfo))
            x = F.dropout(x, self.dropout, training=self.training)
        x = self.conv2(x, inputInfo).half()
        return F.log_softmax(x, dim=1)


class GATNet(nn.Module):
    """GAT: conv1 -> ELU -> (hidden convs, ReLU, dropout) -> conv2 -> log_softmax,
    single-head layers by default as in the reference (heads is accepted and
    passed to the layers)."""

    def __init__(self, in_feats, hidden_feats, out_feats, dropout, alpha, heads, num_layers, per_layer_heads=1):
        super().__init__()
        h = per_layer_heads
        self.dropout = dropout
        self.conv1 = GATConv_multi(in_feats, hidden_feats, dropout, alpha, h)
        self.hidden_layers = nn.ModuleList(GATConv_multi(hidden_feats * h, hidden_feats, dropout, alpha, h)
                                           for _ in range(num_layers - 2))
        self.conv2 = GATConv_multi(hidden_feats * h, out_feats, dropout, alpha, 1)

    def forward(self, inputInfo):
        x = F.elu(self.conv1(inputInfo.x, inputInfo).flatten(1))
        for conv in self.hidden_layers:
            x = F.relu(conv(x, inputInfo).flatten(1))
            x = F.dropout(x, self.dropout, training=self.training)
        x = self.conv2(x, inputInfo).flatten(1)
        return F.log_softmax(x, dim=1)


def test1(model, inputInfo):
    """Accuracy of the arg-max prediction on the test mask."""
    model.eval()
    with torch.no_grad():
        logits = model(inputInfo)
        mask = inputInfo.test_mask
        logits, labels = logits[mask], inputInfo.y[mask]
        correct = torch.sum(torch.argmax(logits, dim=1) == labels)
        return correct.item() * 1.0 / max(len(labels), 1)


def nll_loss(logp, y):
    """Mean negative log-likelihood of the labels (F.nll_loss's value), as a
    gather and a mean."""
    return -logp.float().gather(1, y[:, None]).mean()


def train(model, inputInfo, epoches):
    """`epoches` full-graph training steps; returns the last loss (a tensor)."""
    optimizer = torch.optim.Adam(model.parameters(), lr=1e-2, weight_decay=5e-4)
    loss = None
    for _ in range(epoches):
        model.train()
        logits = model(inputInfo)
        loss = nll_loss(logits, inputInfo.y)
        optimizer.zero_grad()
        loss.backward()
        optimizer.step()
    return loss


def timed_train(inputInfo, model, epoches, warmup=10):
    """`warmup` epochs, then (wall seconds of `epoches` epochs, last loss)."""
    train(model, inputInfo, warmup)
    torch.cuda.synchronize()
    start = time.time()
    loss = train(model, inputInfo, epoches)
    torch.cuda.synchronize()
    return round(time.time() - start, 4), float(loss)


def test(data, epoches, heads, layers, featuredim, hidden, classes, kind="gcn", root="."):
    """The reference's driver: dataset, model, 10 warm-up epochs, then the
    seconds of `epoches` epochs (rounded to 0.1 ms)."""
    device = torch.device("cuda", 0)
    inputInfo = D.MGAT_dataset_csr(data, featuredim, classes, root=root, kind=kind, device=device).to(device)
    model = build_model(kind, inputInfo, layers, hidden).to(device)
    return timed_train(inputInfo, model, epoches)[0]


# ----------------------------------------------------------------------------
# plain FP32 PyTorch baselines of the same architectures
# ----------------------------------------------------------------------------

class PlainInfo:
    """FP32 inputs for the plain models: CSR A_hat (GCN) or COO rows / cols (GAT)."""

    def __init__(self, ds):
        dev = ds.x.device
        csr = ds.csr
        self.x = ds.x.float()
        self.y = ds.y
        self.test_mask = ds.test_mask
        if ds.kind == "gcn":
            vals = D.reference_tile_values(csr, D.reference_degrees(csr))
            self.A = torch.sparse_csr_tensor(torch.from_numpy(csr.indptr), torch.from_numpy(csr.indices.astype(np.int64)),
                                             torch.from_numpy(vals), (csr.n, csr.n)).to(dev)
        else:
            self.rows = torch.from_numpy(csr.row_ids().astype(np.int64)).to(dev)
            self.cols = torch.from_numpy(csr.indices.astype(np.int64)).to(dev)
        self.n = csr.n


class PlainGCNConv(nn.Module):
    def __init__(self, i, o):
        super().__init__()
        self.weights = nn.Parameter(torch.empty(i, o))
        nn.init.uniform_(self.weights, -1.0 / np.sqrt(o), 1.0 / np.sqrt(o))

    def forward(self, x, info):
        return torch.sparse.mm(info.A, x @ self.weights)


class PlainNet(Net):
    def __init__(self, in_feats, hidden_feats, out_feats, num_layers, dropout):
        nn.Module.__init__(self)
        self.conv1 = PlainGCNConv(in_feats, hidden_feats)
        self.hidden_layers = nn.ModuleList(PlainGCNConv(hidden_feats, hidden_feats) for _ in range(num_layers - 2))
        self.conv2 = PlainGCNConv(hidden_feats, out_feats)
        self.num_layers, self.dropout = num_layers, dropout

    def forward(self, info):
        x = F.relu(self.conv1(info.x, info))
        x = F.dropout(x, self.dropout, training=self.training)
        for conv in self.hidden_layers:
            x = F.relu(conv(x, info))
            x = F.dropout(x, self.dropout, training=self.training)
        return F.log_softmax(self.conv2(x, info), dim=1)


class PlainGATConv(nn.Module):
    def __init__(self, i, o, dropout, alpha, heads=1):
        super().__init__()
        self.fc = nn.Linear(i, o * heads, bias=False)
        self.a0 = nn.Parameter(torch.randn(heads, o) * 0.1)
        self.a1 = nn.Parameter(torch.randn(heads, o) * 0.1)
        self.o, self.heads, self.alpha, self.dropout = o, heads, alpha, dropout

    def forward(self, x, info):
        from .benchmarks.plain import gat_fp32
        return gat_fp32(x, self.fc.weight, self.a0, self.a1, info.rows, info.cols, info.n, self.heads, self.o,
                        slope=self.alpha)


class PlainGATNet(GATNet):
    def __init__(self, in_feats, hidden_feats, out_feats, dropout, alpha, heads, num_layers, per_layer_heads=1):
        nn.Module.__init__(self)
        h = per_layer_heads
        self.dropout = dropout
        self.conv1 = PlainGATConv(in_feats, hidden_feats, dropout, alpha, h)
        self.hidden_layers = nn.ModuleList(PlainGATConv(hidden_feats * h, hidden_feats, dropout, alpha, h)
                                           for _ in range(num_layers - 2))
        self.conv2 = PlainGATConv(hidden_feats * h, out_feats, dropout, alpha, 1)


def build_model(kind, ds, layers, hidden, plain=False):
    if kind == "gcn":
        cls = PlainNet if plain else Net
        return cls(ds.num_features, hidden, ds.num_classes, layers, 0.5)
    cls = PlainGATNet if plain else GATNet
    return cls(ds.num_features, hidden, ds.num_classes, 0.5, 0.2, 1, layers)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--model", choices=["gcn", "gat"], default="gcn")
    ap.add_argument("--npz", help="graph in the reference's .npz format")
    ap.add_argument("--synthetic", default="169343,14",
                    help="nodes,mean_degree[,nnz] of a Chung-Lu graph written to a temporary .npz (no --npz)")
    ap.add_argument("--layers", type=int, default=3)
    ap.add_argument("--hidden", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--plain", action="store_true", help="also time the plain FP32 PyTorch model")
    ap.add_argument("--plain-epochs", type=int, default=0,
                    help="epochs of the plain model (default: --epochs); its time is reported per epoch too")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    torch.manual_seed(args.seed)
    tmp = None
    path = args.npz
    if path is None:
        parts = args.synthetic.split(",")
        n, deg = int(parts[0]), float(parts[1])
        nnz = int(parts[2]) if len(parts) > 2 else None
        tmp = tempfile.TemporaryDirectory()
        path = os.path.join(tmp.name, "synthetic.npz")
        t = time.perf_counter()
        D.synthetic_npz(path, n, deg, num_classes=16 if args.model == "gcn" else 10, seed=args.seed, target_nnz=nnz)
        print(f"synthetic graph written ({time.perf_counter() - t:.1f}s)", file=sys.stderr, flush=True)
    t = time.perf_counter()
    ds = D.MGAT_dataset_csr(path, kind=args.model, device=dev).to(dev)
    prep = time.perf_counter() - t
    print(f"dataset: {ds.num_nodes_ori} nodes, {ds.csr.nnz} entries ({prep:.1f}s)", file=sys.stderr, flush=True)
    model = build_model(args.model, ds, args.layers, args.hidden).to(dev)
    secs, loss = timed_train(ds, model, args.epochs, args.warmup)
    print(f"trained {args.epochs} epochs in {secs:.2f}s", file=sys.stderr, flush=True)
    rec = {"model": args.model, "graph": os.path.basename(path), "nodes": ds.num_nodes_ori, "nnz": ds.csr.nnz,
           "layers": args.layers, "hidden": args.hidden, "epochs": args.epochs, "seconds": secs,
           "ms_per_epoch": round(secs / args.epochs * 1e3, 3), "final_loss": round(loss, 4),
           "dataset_build_s": round(prep, 3)}
    if ds.test_mask is not None:
        rec["test_accuracy"] = round(test1(model, ds), 4)
    if args.plain:
        torch.manual_seed(args.seed)
        info = PlainInfo(ds)
        pm = build_model(args.model, ds, args.layers, args.hidden, plain=True).to(dev)
        pe = args.plain_epochs or args.epochs
        psecs, ploss = timed_train(info, pm, pe, min(args.warmup, 3))
        rec.update(plain_fp32_epochs=pe, plain_fp32_seconds=psecs, plain_fp32_ms_per_epoch=round(psecs / pe * 1e3, 3),
                   plain_fp32_final_loss=round(ploss, 4), speedup=round(psecs / pe / (secs / args.epochs), 2))
    print(json.dumps(rec), flush=True)
    if tmp is not None:
        tmp.cleanup()


if __name__ == "__main__":
    main()
