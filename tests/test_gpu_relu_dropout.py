"""Fused ReLU + dropout against a numpy restatement of its keep rule (the
lowbias32 hash of the element index and the seed), and against F.relu /
dropout semantics: values, keep rate, backward, autograd."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from att_amd import ops
from att_amd.layers import relu_dropout

pytestmark = pytest.mark.gpu


def lowbias32(x):
    """C. Wellons' 32-bit integer hash on uint32 numpy arrays."""
    x = x.astype(np.uint32)
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def keep_oracle(seed, n, p):
    key = lowbias32(np.array([(seed & 0xFFFFFFFF) ^ (seed >> 32)], np.uint64).astype(np.uint32))[0]
    h = lowbias32((np.uint32(key) + np.arange(n, dtype=np.uint64).astype(np.uint32)).astype(np.uint32))
    return (h >> np.uint32(8)) >= np.uint32(int(p * 16777216.0))


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32])
@pytest.mark.parametrize("p", [0.5, 0.1, 0.0, 1.0])
def test_forward_matches_keep_rule(cuda, dtype, p):
    torch.manual_seed(0)
    n = 4096 * 24
    x = torch.randn(n, device=cuda).to(dtype)
    x[5::1009] = float("nan")                     # NaN stays NaN where kept, as relu and dropout leave it
    seed = 0x123456789ABCDEF
    y, mask = ops.relu_dropout(x, p, seed)
    keep = torch.from_numpy(keep_oracle(seed, n, p)).to(cuda)
    passed = keep & (x > 0)
    # the kernel's scale: 1 / (1 - p) formed in float32
    scale = np.float32(0.0) if p >= 1.0 else np.float32(1.0) / (np.float32(1.0) - np.float32(p))
    kept_nan = keep & torch.isnan(x)
    want = torch.where(passed | kept_nan, (x.float() * torch.tensor(scale, device=cuda)).to(dtype),
                       torch.zeros_like(x))
    assert torch.equal(torch.isnan(y), kept_nan)
    assert torch.equal(torch.nan_to_num(y), torch.nan_to_num(want))
    bits = ((mask[:, None].int() >> torch.arange(8, device=cuda)) & 1).bool().reshape(-1)
    assert torch.equal(bits, passed)
    if 0.0 < p < 1.0:
        rate = keep.float().mean().item()
        assert abs(rate - (1.0 - p)) < 5 * np.sqrt(p * (1 - p) / n)
    y2, mask2 = ops.relu_dropout(x, p, seed)     # the same bits (NaN included) run to run
    assert torch.equal(y.view(torch.uint8), y2.view(torch.uint8)) and torch.equal(mask, mask2)
    if 0.0 < p < 1.0:
        _, mask3 = ops.relu_dropout(x, p, seed + 1)
        assert not torch.equal(mask, mask3)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_backward_and_autograd(cuda, dtype):
    torch.manual_seed(1)
    x = torch.randn(1000, 256, device=cuda).to(dtype).requires_grad_(True)
    w = torch.randn(1000, 256, device=cuda).to(dtype)
    g = torch.Generator().manual_seed(5)
    y = relu_dropout(x, 0.5, True, generator=g)
    (y.float() * w.float()).sum().backward()
    passed = y != 0
    want = torch.where(passed, (w.float() * 2.0).to(dtype), torch.zeros_like(w))
    assert torch.equal(x.grad, want)
    # the same generator state draws the same seed: the same mask
    y_again = relu_dropout(x.detach(), 0.5, True, generator=torch.Generator().manual_seed(5))
    assert torch.equal(y.detach(), y_again)
    # out of training: F.relu
    assert torch.equal(relu_dropout(x.detach(), 0.5, False), F.relu(x.detach()))
