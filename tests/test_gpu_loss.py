"""Fused NLL of log_softmax (the GCN training loss) against the PyTorch chain
of the reference's model: F.log_softmax over the class columns, the label's
term gathered, the mean over the selected rows, autograd's gradient."""
import pytest
import torch
import torch.nn.functional as F

from att_amd import ops
from att_amd.layers import nll_log_softmax_loss

pytestmark = pytest.mark.gpu


def _reference(logits, labels, classes, rows):
    x = logits.detach().float().requires_grad_(True)
    lp = F.log_softmax(x[:, :classes], dim=1)
    term = -lp.gather(1, labels[:, None]).squeeze(1)
    loss = (term if rows is None else term.index_select(0, rows)).mean()
    loss.backward()
    return loss.detach(), x.grad


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32])
@pytest.mark.parametrize("use_rows", [False, True])
@pytest.mark.parametrize("n,ld,classes", [(5000, 48, 41), (40000, 16, 7), (3000, 80, 70)])
def test_loss_and_grad_match_torch(cuda, dtype, use_rows, n, ld, classes):
    """Class counts held in registers (7: one per lane, 41: three) and read
    from memory per pass (70 > 64); 40,000 rows take more row groups than
    the grid has workgroups (grid-stride loop)."""
    torch.manual_seed(3)
    logits = (torch.randn(n, ld, device=cuda) * 3).to(dtype)
    labels = torch.randint(0, classes, (n,), device=cuda)
    rows = torch.randperm(n, device=cuda)[:(3 * n) // 5] if use_rows else None   # unsorted on purpose
    loss, grad = ops.nll_log_softmax(logits, labels, classes, rows)
    ref_loss, ref_grad = _reference(logits, labels, classes, rows)
    assert abs(loss.item() - ref_loss.item()) <= 2e-6 * abs(ref_loss.item())
    # both gradients are (softmax - one-hot) / n formed in FP32 -- here as
    # exp(z - lse), in autograd through log_softmax's backward -- so the
    # softmax terms agree to a few 2^-24 (absolute, they are at most 1), then
    # each is rounded once to the logits' type
    n_sel = n if rows is None else rows.numel()
    g, r = grad.float(), ref_grad.to(dtype).float()
    ulp = {torch.float16: 2.0 ** -10, torch.bfloat16: 2.0 ** -7, torch.float32: 2.0 ** -23}[dtype]
    tiny = {torch.float16: 2.0 ** -24, torch.bfloat16: 0.0, torch.float32: 0.0}[dtype]
    assert torch.all((g - r).abs() <= ulp * r.abs() + 8 * 2.0 ** -24 / n_sel + tiny)
    assert not grad[:, classes:].any()                       # padding classes untouched (zero)
    if rows is not None:
        other = torch.ones(n, dtype=torch.bool, device=cuda)
        other[rows] = False
        assert not grad[other].any()                         # rows outside the mean untouched
    # bitwise run to run
    loss2, grad2 = ops.nll_log_softmax(logits, labels, classes, rows)
    assert torch.equal(loss, loss2) and torch.equal(grad, grad2)


def test_autograd_function_and_bad_labels(cuda):
    """The autograd function: its loss is the combined call's, its gradient
    the reference's scaled by the upstream gradient inside the kernel (the
    scaled FP32 value rounded once -- these gradients sit in FP16's
    subnormal range, where scaling after the rounding would differ); the
    loss-only and gradient-only calls equal the combined one."""
    torch.manual_seed(4)
    n, ld, classes = 700, 48, 41
    logits = torch.randn(n, ld, device=cuda, dtype=torch.float16).requires_grad_(True)
    labels = torch.randint(0, classes, (n,), device=cuda)
    rows = torch.arange(0, n, 3, device=cuda)
    loss_ref, grad = ops.nll_log_softmax(logits.detach(), labels, classes, rows)
    lo, gn = ops.nll_log_softmax(logits.detach(), labels, classes, rows, grad=False)
    ln, go = ops.nll_log_softmax(logits.detach(), labels, classes, rows, loss=False)
    assert gn is None and ln is None and torch.equal(lo, loss_ref) and torch.equal(go, grad)
    loss = nll_log_softmax_loss(logits, labels, classes, rows)
    assert torch.equal(loss, loss_ref)
    _, ref = _reference(logits, labels, classes, rows)
    for scale in (1.0, 2.0, 0.3):
        logits.grad = None
        (scale * nll_log_softmax_loss(logits, labels, classes, rows)).backward()
        g, r = logits.grad.float(), (scale * ref).half().float()
        tol = 2.0 ** -10 * r.abs() + 8 * scale * 2.0 ** -24 / rows.numel() + 2.0 ** -24
        assert torch.all((g - r).abs() <= tol), scale
        if scale == 1.0:
            assert torch.equal(logits.grad, grad)
    bad = labels.clone()
    bad[rows[5]] = classes                                   # a label outside the class columns
    lb, gb = ops.nll_log_softmax(logits.detach(), bad, classes, rows)
    assert torch.isnan(lb) and torch.isnan(gb[rows[5], :classes]).all()
    assert not torch.isnan(gb[rows[6]]).any()


@pytest.mark.parametrize("stripe", [None, 8])
def test_gcn_step_with_fused_ops_equals_torch_ops(cuda, stripe):
    """The benchmark's 2-layer GCN step (GCNConv, fused ReLU + dropout, fused
    loss over the real rows) against the same step with F.relu, the same
    dropout mask applied by torch, F.log_softmax and the gathered NLL: the
    loss and both weight gradients agree to FP16 / FP32 rounding."""
    import torch.nn.functional as F
    from att_amd import graphs as G
    from att_amd.layers import GCNConv, GCNGraph
    g = G.chung_lu(3000, 20, 2.1, seed=11)
    vals = G.gcn_norm_values(g, G.gcn_norm_degree(g))
    graph = GCNGraph(g, cuda, tile_dtype=torch.float16, symmetric=True, values=vals, build="device", stripe=stripe)
    gen = torch.Generator().manual_seed(2)
    X = graph.to_striped(torch.randn(g.n, 96, generator=gen).half().to(cuda))
    labels = graph.to_striped(torch.randint(0, 41, (g.n,), generator=gen).to(cuda))
    real = graph.stripe_index.sort().values if stripe else torch.arange(g.n, device=cuda)
    torch.manual_seed(7)
    c1, c2 = GCNConv(96, 64).to(cuda), GCNConv(64, 48).to(cuda)
    seed = 0x5EED

    # fused: the ops the benchmark step uses
    y1 = c1(X, graph)
    h = ReluDropoutSeeded.apply(y1, seed)
    loss_f = nll_log_softmax_loss(c2(h, graph), labels, 41, real)
    loss_f.backward()
    gf = [c1.weights.grad.clone(), c2.weights.grad.clone()]
    for c in (c1, c2):
        c.weights.grad = None

    # torch ops with the fused op's mask
    y1 = c1(X, graph)
    _, mask = ops.relu_dropout(y1.detach(), 0.5, seed)
    keep = ((mask[:, None].int() >> torch.arange(8, device=cuda)) & 1).bool().reshape(y1.shape)
    keep = keep | (y1.detach() <= 0)     # the mask records kept positive entries; relu zeroes the rest anyway
    h = F.relu(y1) * keep * 2.0
    lp = F.log_softmax(c2(h, graph)[:, :41].float(), dim=1)
    loss_t = -lp.gather(1, labels[:, None]).squeeze(1).index_select(0, real).mean()
    loss_t.backward()
    gt = [c1.weights.grad, c2.weights.grad]
    assert abs(loss_f.item() - loss_t.item()) <= 1e-5 * abs(loss_t.item())
    for a, b in zip(gf, gt):
        assert ((a - b).norm() / b.norm()).item() < 2e-3


class ReluDropoutSeeded(torch.autograd.Function):
    """layers.ReluDropout with a fixed seed (the step above fixes its mask)."""

    @staticmethod
    def forward(ctx, x, seed):
        y, mask = ops.relu_dropout(x.contiguous(), 0.5, seed)
        ctx.save_for_backward(mask)
        return y

    @staticmethod
    def backward(ctx, dy):
        (mask,) = ctx.saved_tensors
        return ops.relu_dropout_bwd(dy.contiguous(), 0.5, mask), None
