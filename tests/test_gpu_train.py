"""The training harness on the GPU: reference-format .npz -> dataset object ->
GCN / GAT nets -> training loop; the forward pass agrees with the plain FP32
model carrying the same weights, training lowers the loss, and the reference's
driver signature returns seconds."""
import numpy as np
import pytest
import torch

from att_amd import datasets as D
from att_amd import train as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def npz(tmp_path_factory):
    path = tmp_path_factory.mktemp("g") / "cl3k.npz"
    D.synthetic_npz(str(path), 3000, 8, num_features=32, num_classes=6, seed=3)
    return str(path)


@pytest.mark.parametrize("kind", ["gcn", "gat"])
def test_dataset_and_training(kind, npz, cuda):
    torch.manual_seed(0)
    ds = D.MGAT_dataset_csr(npz, kind=kind, device=cuda).to(cuda)
    assert ds.num_nodes_ori == 3000 and ds.num_nodes == 3000 + 8 - 3000 % 8
    assert ds.num_features == 32 and ds.num_classes == 6 and ds.x.dtype == torch.float16
    model = T.build_model(kind, ds, 3, 32).to(cuda)
    first = float(T.train(model, ds, 1))
    last = float(T.train(model, ds, 30))
    assert np.isfinite(first) and np.isfinite(last) and last < first, (first, last)
    acc = T.test1(model, ds)
    assert 0.0 <= acc <= 1.0


def test_gcn_matches_plain_fp32(npz, cuda):
    torch.manual_seed(1)
    ds = D.MGAT_dataset_csr(npz, kind="gcn", device=cuda).to(cuda)
    ours = T.build_model("gcn", ds, 3, 32).to(cuda).eval()
    plain = T.build_model("gcn", ds, 3, 32, plain=True).to(cuda).eval()
    with torch.no_grad():
        for a, b in zip([ours.conv1, *ours.hidden_layers, ours.conv2],
                        [plain.conv1, *plain.hidden_layers, plain.conv2]):
            b.weights.copy_(a.weights)
        got = ours(ds).float()
        want = plain(T.PlainInfo(ds))
    err = ((got - want).norm() / want.norm()).item()
    assert err < 2e-2, err


def test_gat_matches_plain_fp32(npz, cuda):
    torch.manual_seed(2)
    ds = D.MGAT_dataset_csr(npz, kind="gat", device=cuda).to(cuda)
    ours = T.build_model("gat", ds, 2, 16).to(cuda).eval()
    plain = T.build_model("gat", ds, 2, 16, plain=True).to(cuda).eval()
    with torch.no_grad():
        for a, b in zip([ours.conv1, ours.conv2], [plain.conv1, plain.conv2]):
            b.fc.weight.copy_(a.fc.weight)
            b.a0.copy_(a.a0.view_as(b.a0))
            b.a1.copy_(a.a1.view_as(b.a1))
        got = ours(ds).float()
        want = plain(T.PlainInfo(ds))
    err = ((got - want).norm() / want.norm()).item()
    assert err < 3e-2, err


def test_reference_driver_returns_seconds(npz, cuda):
    secs = T.test(npz, 3, 1, 3, 64, 32, 10, kind="gcn")
    assert isinstance(secs, float) and secs > 0
