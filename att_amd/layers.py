"""GCN / GAT / GATv2 layers on the window-block kernels, with hand-written
backward passes behind torch.autograd.Function.

Reference layers (end2end_*/**/mgat_csr/mgat_conv.py):
  MGCNFunction / GCNConv          end2end_ori/gcn_final/mgat_csr/mgat_conv.py:27-108
  GATConv_multi                   end2end_head/gat_final/mgat_csr/mgat_conv.py:154-199
  GATv2Conv_multi                 end2end_head/gatv2_final/mgat_csr/mgat_conv.py:151-198
                                  (logits as end2end_ori/gatv2_final/mgat_csr/mgat_conv.py:174-196)
Graph containers mirror the reference datasets' attribute names
(end2end_ori/gcn_final/mgat_csr/mdataset_fp16.py:39-64,
 end2end_head/gat_final/mgat_csr/mdataset_fp16.py:35-62).

Differences from the reference, all deliberate:
  * gradients are exact (the reference's attention-vector backward pairs a0
    with the column sums and adds cross terms, mgat_conv.py:42-67);
  * the edge softmax subtracts the row max (same result, no overflow) and is
    one fused kernel; attention probabilities are recomputed in the backward
    pass from the per-row log-sum-exp instead of being kept per edge;
  * dropout on attention uses a counter-based hash so the backward pass
    regenerates the mask;
  * the SpMM backward uses the transposed structure, so non-symmetric graphs
    are handled (the reference reuses A, valid only for symmetric A).
"""
import math

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import graphs as G
from . import ops


def default_part_size(total_blocks, lo=4, hi=512):
    """Blocks per part: enough parts to keep ~2k waves busy, long parts otherwise."""
    return int(min(max(total_blocks // 8192, lo), hi))


def _blocks_estimate(csr, wide):
    """Upper bound of the block count (every entry its own condensed column)."""
    return max(csr.nnz // wide, 1)


def transpose_csr(csr):
    """CSR of A^T plus, for each of its entries, the position of the entry in A."""
    rows = csr.row_ids().astype(np.int64)
    cols = csr.indices.astype(np.int64)
    order = np.lexsort((rows, cols))          # by column, then row
    t_indptr = np.zeros(csr.n + 1, np.int64)
    t_indptr[1:] = np.cumsum(np.bincount(cols, minlength=csr.n))
    return G.CSR(t_indptr, rows[order].astype(np.int32), csr.n), order.astype(np.int64)


class GCNGraph:
    """Device-resident window-block structure of A_hat = D^-1/2 (A+I) D^-1/2.

    ``csr`` already holds the self-loops; values are d_i d_j with
    d = 1/sqrt(row count) (mdataset_fp16.py:54-56), rounded once to
    ``tile_dtype``.  Rows are padded to a multiple of ``window`` so every row
    lies in a window that has a part (no zero-fill before the SpMM).

    ``stripe`` (2 to 32 owners, a power of two): striped node order (graphs.stripe_pos).
    The windows keep the graph's own rows; the column ids become striped
    positions and :meth:`spmm` takes and returns node rows in striped order
    ([num_nodes, D], num_nodes = graphs.stripe_rows(n, stripe); convert with
    :meth:`to_striped` / :meth:`from_striped`), so the SpMM can split every
    part among the XCDs by owner (8 owners at 256 FP16 features, two per XCD
    slot: each XCD's L2 caches a quarter of the gathered rows of its feature
    slice).  Compact tiles only.
    """

    def __init__(self, csr, device, window=16, wide=32, part_size=None, tile_dtype=torch.float16,
                 symmetric=None, values=None, compact="auto", build="host", chunk_blocks=None, stripe=None):
        if build not in ("host", "device"):
            raise ValueError(f"build must be 'host' or 'device', got {build!r}")
        self.build = build
        self.window, self.wide = window, wide
        self.chunk_blocks = chunk_blocks
        self.stripe = stripe
        if stripe and wide != 32:
            raise ValueError("striped node order runs over compact tiles (wide 32)")
        # compact tiles win once a gathered row is >= 256 bytes; "auto" keeps
        # both forms and picks per call
        self.compact = True if stripe else (compact if wide == 32 else False)
        self.num_nodes_ori = csr.n
        padded = csr.padded(64 * stripe if stripe else window)
        self.num_nodes = padded.n
        self.num_edges = csr.nnz
        if values is None:
            d = G.gcn_norm_degree(padded)
            d[~np.isfinite(d)] = 0.0
            values = G.gcn_norm_values(padded, d)
        self.part_size = part_size or default_part_size(_blocks_estimate(padded, wide))
        self.tile_dtype = tile_dtype
        if build == "device":
            # structure, symmetry test and A^T all on the device
            rp_d = torch.from_numpy(padded.rowptr32()).to(device)
            col_d = torch.from_numpy(padded.indices).to(device)
            vals_d = torch.from_numpy(np.asarray(values, np.float32)).to(device)
            self.fwd = self._build_device(rp_d, col_d, vals_d, device)
            t_rowptr = t_col = t_perm = None
            if symmetric is None or not symmetric:
                t_rowptr, t_col, t_perm = ops.csr_transpose_device(rp_d, col_d, padded.n)
                if symmetric is None:
                    symmetric = bool(torch.equal(t_rowptr, rp_d) and torch.equal(t_col, col_d))
            self.symmetric = symmetric
            self.bwd = self.fwd if symmetric else \
                self._build_device(t_rowptr, t_col, vals_d.index_select(0, t_perm), device)
        else:
            self.fwd = self._build(padded, values, device)
            if symmetric is None:
                symmetric = bool(np.array_equal(np.sort(csr.transpose_perm()), np.arange(csr.nnz)))
            self.symmetric = symmetric
            if symmetric:
                self.bwd = self.fwd
            else:
                t, order = transpose_csr(padded)
                self.bwd = self._build(t, np.asarray(values)[order], device)
        if stripe:
            self.stripe_index = G.stripe_pos(torch.arange(csr.n, dtype=torch.int64, device=device), stripe,
                                             self.num_nodes)
            for s in (self.fwd, self.bwd):
                if "oseg" not in s:
                    rp, win, _ = s["tasks"]
                    s["oseg"] = ops.stripe_segments(rp, win, s["arrays"][1], stripe, self.num_nodes)
        # reference attribute names
        (self.row_pointers, self.column_index, self.degrees, self.t_window_rowTensor,
         self.t_atomicTensor) = self.fwd["arrays"]

    def _build_device(self, rp_d, col_d, vals_d, device):
        if self.stripe:   # column ids -> striped positions, rows re-sorted
            rows = torch.repeat_interleave(torch.arange(rp_d.numel() - 1, device=device),
                                           (rp_d[1:] - rp_d[:-1]).long())
            scol = G.stripe_pos(col_d.long(), self.stripe, rp_d.numel() - 1)
            order = torch.sort(rows * (rp_d.numel() - 1) + scol).indices
            col_d, vals_d = scol.index_select(0, order).int(), vals_d.index_select(0, order)
            del rows, scol, order
        b = ops.blocks_build_device(rp_d, col_d, "edge_f32", vals_d, self.tile_dtype, self.window, self.wide,
                                    self.part_size, emit_empty=True, compact=bool(self.compact))
        return self._finish(b, device)

    def _build(self, csr, values, device):
        values = np.asarray(values, np.float32)
        if self.stripe:   # column ids -> striped positions, rows re-sorted
            scol = G.stripe_pos(csr.indices.astype(np.int64), self.stripe, csr.n)
            order = np.lexsort((scol, csr.row_ids()))
            csr, values = G.CSR(csr.indptr, scol[order].astype(np.int32), csr.n), values[order]
        b = ops.blocks_build(csr.rowptr32(), csr.indices, "edge_f32", values,
                             self.tile_dtype, self.window, self.wide, self.part_size, emit_empty=True,
                             compact=self.compact)
        return self._finish(b, device)

    def to_striped(self, X):
        """Node rows [n, ...] in the order :meth:`spmm` takes: [num_nodes, ...]
        with node j at row stripe_pos(j) and zero padding rows (X itself
        without striped order)."""
        if not self.stripe:
            return X
        out = X.new_zeros((self.num_nodes,) + tuple(X.shape[1:]))
        out.index_copy_(0, self.stripe_index, X)
        return out

    def from_striped(self, Y):
        """The graph's node rows [n, ...] of rows in :meth:`spmm`'s order."""
        if not self.stripe:
            return Y[:self.num_nodes_ori]
        return Y.index_select(0, self.stripe_index)

    def _finish(self, b, device):
        keys = ("rp", "col", "val", "window", "atomic")
        if self.compact is True:   # compact only: the dense tiles are not kept
            b["val"] = torch.zeros(1, dtype=self.tile_dtype)
        arrays = [b[k].to(device) for k in keys]
        self.n_unique = b["n_unique"]
        out = {"arrays": arrays, "parts": int(b["window"].numel()), "n_cols": int(b["col"].numel())}
        # the SpMM runs over the canonical task table (chunks of the windows
        # fixed by the structure's size): results do not depend on part_size
        out["tasks"] = ops.canonical_tasks(arrays[0], arrays[3], out["n_cols"], self.wide, self.chunk_blocks)
        if self.compact:
            out["compact"] = [b[k].to(device) for k in ("masks", "block_base", "values")]
        return out

    def spmm(self, X, transpose=False, out_dtype=None):
        s = self.bwd if transpose else self.fwd
        _, col, val, _, _ = s["arrays"]
        rp, win, at = s["tasks"]
        if self.stripe:
            if X.shape[0] != self.num_nodes:
                raise ValueError(f"X must hold the {self.num_nodes} rows of the striped order (to_striped)")
            masks, base, values = s["compact"]
            return ops.spmm_compact_striped(rp, col, masks, base, values, win, at, s["oseg"], self.stripe,
                                            X.contiguous(), self.window, out_dtype=out_dtype or X.dtype,
                                            rows_covered=True)
        if X.shape[0] < self.num_nodes_ori:
            raise ValueError("X has fewer rows than the graph")
        use_compact = self.compact is True or (self.compact == "auto" and X.shape[1] * X.element_size() >= 256)
        if use_compact:
            masks, base, values = s["compact"]
            return ops.spmm_compact(rp, col, masks, base, values, win, at, X.contiguous(), self.num_nodes_ori,
                                    self.window, out_dtype=out_dtype or X.dtype, rows_covered=True)
        return ops.spmm_block(rp, col, val, win, at, X.contiguous(), self.num_nodes_ori, self.window, self.wide,
                              out_dtype=out_dtype or X.dtype, rows_covered=True)


class ReluDropout(torch.autograd.Function):
    """F.dropout(F.relu(x), p) as one kernel each way (ops.relu_dropout): the
    forward keeps a one-bit mask of the elements that passed."""

    @staticmethod
    def forward(ctx, x, p, seed):
        y, mask = ops.relu_dropout(x.contiguous(), p, seed)
        ctx.save_for_backward(mask)
        ctx.p = p
        return y

    @staticmethod
    def backward(ctx, dy):
        (mask,) = ctx.saved_tensors
        return ops.relu_dropout_bwd(dy.contiguous(), ctx.p, mask), None, None


def relu_dropout(x, p=0.5, training=True, generator=None):
    """F.dropout(F.relu(x), p, training) (the reference GCN between its
    layers, end2end_ori/gcn_final/mgat_csr/gat_mgnn.py:17-40) fused.  Each
    call draws a 62-bit seed from ``generator`` (a CPU torch.Generator; None:
    the default one), so the masks follow torch.manual_seed; out of training,
    or with p = 0, it is F.relu."""
    if not training or p == 0.0:
        return F.relu(x)
    seed = int(torch.randint(0, 1 << 62, (1,), generator=generator))
    return ReluDropout.apply(x, float(p), seed)


class NllLogSoftmax(torch.autograd.Function):
    """Mean NLL of log_softmax over the selected rows (ops.nll_log_softmax):
    the loss forward, the gradient scaled by the upstream one backward, one
    kernel each."""

    @staticmethod
    def forward(ctx, logits, labels, classes, rows):
        logits = logits.contiguous()
        loss, _ = ops.nll_log_softmax(logits, labels, classes, rows, grad=False)
        ctx.save_for_backward(logits, labels, rows)
        ctx.classes = classes
        return loss

    @staticmethod
    def backward(ctx, g):
        logits, labels, rows = ctx.saved_tensors
        _, grad = ops.nll_log_softmax(logits, labels, ctx.classes, rows, loss=False,
                                      grad_scale=g.detach().float().reshape(1))
        return grad, None, None, None


def nll_log_softmax_loss(logits, labels, classes, rows=None):
    """The reference GCN's loss, F.nll_loss(F.log_softmax(logits[:, :classes]),
    labels) (end2end_ori/gcn_final/mgat_csr/gat_mgnn.py:39,66), averaged over
    ``rows`` (int64 indices; None: all rows), as one fused kernel forward and
    none but a scaling backward."""
    return NllLogSoftmax.apply(logits, labels, int(classes), rows)


def weight_grad(X, dY, min_rows_per_split=2048, max_splits=64):
    """X^T dY (fp32, or fp64 for fp64 inputs) for a long row dimension N: a plain GEMM gets a handful of
    output tiles that each walk all N rows, so the rows are split into S chunks
    run as one batched product (fp16 / bf16 operands, FP32 accumulation) whose
    S partial products are summed in FP32 in a fixed order."""
    acc = torch.promote_types(X.dtype, torch.float32)
    n = X.shape[0]
    S = max(1, min(max_splits, n // min_rows_per_split))
    if S == 1:
        return torch.mm(X.t(), dY).to(acc)
    m = (n // S) * S
    part = torch.bmm(X[:m].view(S, m // S, X.shape[1]).transpose(1, 2),
                     dY[:m].contiguous().view(S, m // S, dY.shape[1]))
    out = part.to(acc).sum(0)
    if m < n:
        out += torch.mm(X[m:].t(), dY[m:]).to(acc)
    return out


class LinearSplitK(torch.autograd.Function):
    """Y = X W^T with W cast to X's dtype (nn.Linear without bias); the weight
    gradient uses :func:`weight_grad` and is returned in W's dtype."""

    @staticmethod
    def forward(ctx, X, W):
        Wc = W.to(X.dtype)
        ctx.save_for_backward(X, Wc)
        ctx.w_dtype = W.dtype
        return F.linear(X, Wc)

    @staticmethod
    def backward(ctx, dY):
        X, Wc = ctx.saved_tensors
        dY = dY.to(X.dtype)
        dX = dY.reshape(-1, dY.shape[-1]) @ Wc if ctx.needs_input_grad[0] else None
        dW = weight_grad(X.reshape(-1, X.shape[-1]), dY.reshape(-1, dY.shape[-1])).t().to(ctx.w_dtype) \
            if ctx.needs_input_grad[1] else None
        return (dX.view(X.shape) if dX is not None else None), dW


class MGCNFunction(torch.autograd.Function):
    """Y = A_hat . (X W); backward dXW = A_hat^T dY, dX = dXW W^T, dW = X^T dXW
    (end2end_ori/gcn_final/mgat_csr/mgat_conv.py:27-74)."""

    @staticmethod
    def forward(ctx, X, weights, inputInfo):
        ctx.save_for_backward(X, weights)
        ctx.inputInfo = inputInfo
        X_prime = torch.mm(X, weights)
        return inputInfo.spmm(X_prime)

    @staticmethod
    def backward(ctx, d_output):
        X, weights = ctx.saved_tensors
        g = ctx.inputInfo
        d_input_prime = g.spmm(d_output.to(X.dtype).contiguous(), transpose=True)
        # dX only when the input needs it (the first layer's features do not)
        d_input = torch.mm(d_input_prime, weights.t()) if ctx.needs_input_grad[0] else None
        d_weights = weight_grad(X, d_input_prime).to(weights.dtype) if ctx.needs_input_grad[1] else None
        return d_input, d_weights, None


def spmm_width(d, elt_bytes, line=128):
    """Feature width to run the SpMM at: rows that neither divide nor fill whole
    128-byte lines straddle lines on every other gather, so they are padded up
    to the next multiple of a line (e.g. 48 fp16 features -> 64)."""
    row = d * elt_bytes
    if row % line == 0 or line % row == 0:
        return d
    per_line = line // elt_bytes
    return -(-d // per_line) * per_line


class GCNConv(nn.Module):
    """end2end_ori/gcn_final/mgat_csr/mgat_conv.py:82-108."""

    def __init__(self, input_dim, output_dim, dtype=torch.float16):
        super().__init__()
        self.weights = nn.Parameter(torch.empty(input_dim, output_dim))
        self.dtype = dtype
        self.reset_parameters()

    def reset_parameters(self):
        stdv = 1.0 / math.sqrt(self.weights.size(1))
        self.weights.data.uniform_(-stdv, stdv)

    def forward(self, X, inputInfo):
        W = self.weights.to(self.dtype)
        out = W.shape[1]
        wide = spmm_width(out, W.element_size())
        if wide != out:   # zero columns so gathered rows are whole cache lines
            return MGCNFunction.apply(X, F.pad(W, (0, wide - out)), inputInfo)[:, :out]
        return MGCNFunction.apply(X, W, inputInfo)


# ----------------------------------------------------------------------------
# attention layers
# ----------------------------------------------------------------------------

class GATGraph:
    """Device-resident structures for the attention layers: CSR rows (fused
    edge softmax), window blocks with edge-id maps of A (aggregation, SDDMM)
    and of A^T (aggregation backward), and the transpose permutation (column
    sums).  Attribute names follow the reference dataset
    (end2end_head/gat_final/mgat_csr/mdataset_fp16.py:53-65)."""

    def __init__(self, csr, device, window=16, wide=32, part_size=None, build="host", chunk_blocks=None):
        if build not in ("host", "device"):
            raise ValueError(f"build must be 'host' or 'device', got {build!r}")
        self.build = build
        self.window, self.wide = window, wide
        self.num_nodes_ori = csr.n
        padded = csr.padded(window)
        self.num_nodes = padded.n
        self.num_edges = csr.nnz
        # at least 32 blocks per part: the heads multiply the partial-tile
        # traffic of split windows
        self.part_size = part_size or max(default_part_size(_blocks_estimate(padded, wide)), 32)
        self.rowptr = torch.from_numpy(csr.rowptr32()).to(device)
        self.col = torch.from_numpy(csr.indices).to(device)
        if build == "device":
            # one device build of A with its transposed map: on a symmetric
            # pattern the structure of A^T is A's, with map_T as its edge ids
            b = ops.blocks_build_device(torch.from_numpy(padded.rowptr32()).to(device), self.col, "map", None,
                                        torch.int32, window, wide, self.part_size, emit_empty=True,
                                        want_trans=True, compact=(wide == 32))
            perm = ops.map_transpose_perm(b["val"], b["map_t"], csr.nnz)
            symmetric = csr.nnz == 0 or int(perm.min().item()) >= 0
        else:
            b = ops.blocks_build(padded.rowptr32(), padded.indices, "map", None, torch.int32, window, wide,
                                 self.part_size, emit_empty=True, compact=(wide == 32))
            symmetric = False
        self.n_unique = b["n_unique"]
        self.fwd = [b[k].to(device) for k in ("rp", "col", "val", "window", "atomic")]
        # compact edge list of A (block offsets, edge ids, packed positions) for the SDDMM
        self.edges = [b[k].to(device) for k in ("block_base", "values", "positions")] if wide == 32 else None
        # compact edge-id maps (row masks, block offsets, edge ids) of A and A^T for the aggregations
        self.cmap_fwd = [b[k].to(device) for k in ("masks", "block_base", "values")] if wide == 32 else None
        self.cmap_bwd = None
        if symmetric:
            self.bwd = [b["rp"], b["col"], b["map_t"], b["window"], b["atomic"]]
            self.t_rowptr = self.rowptr
            self.t_col = self.col
            self.t_perm = perm
        elif build == "device":
            # A^T on the device, its edge-id map re-expressed in A's edge order
            t_rowptr, t_col, t_perm = ops.csr_transpose_device(torch.from_numpy(padded.rowptr32()).to(device),
                                                               self.col, padded.n)
            bt = ops.blocks_build_device(t_rowptr, t_col, "map", None, torch.int32, window, wide, self.part_size,
                                         emit_empty=True)
            self.bwd = [bt["rp"], bt["col"], ops.map_remap(bt["val"], t_perm), bt["window"], bt["atomic"]]
            self.t_rowptr = t_rowptr[: csr.n + 1].contiguous()
            self.t_col = t_col
            self.t_perm = t_perm
        self.chunk_blocks = chunk_blocks
        if symmetric or build == "device":
            self.row_pointers, self.column_index, self.degrees, self.window_t, self.automic = self.fwd
            self.degrees_trans = self.bwd[2]
            self._compact_bwd()
            self._use_tasks()
            return
        t, order = transpose_csr(padded)
        bt = ops.blocks_build(t.rowptr32(), t.indices, "map", None, torch.int32, window, wide, self.part_size,
                              emit_empty=True)
        mt = bt["val"].numpy()
        mt = np.where(mt >= 0, order[np.maximum(mt, 0)], -1).astype(np.int32)
        self.bwd = [bt["rp"].to(device), bt["col"].to(device), torch.from_numpy(mt).to(device),
                    bt["window"].to(device), bt["atomic"].to(device)]
        self.t_rowptr = torch.from_numpy(t.rowptr32()[: csr.n + 1].copy()).to(device)
        self.t_perm = torch.from_numpy(order[: int(t.indptr[csr.n])]).to(device)
        self.t_col = torch.from_numpy(t.indices[: int(t.indptr[csr.n])].copy()).to(device)
        # reference names
        self.row_pointers, self.column_index, self.degrees, self.window_t, self.automic = self.fwd
        self.degrees_trans = self.bwd[2]
        self._compact_bwd()
        self._use_tasks()

    def _use_tasks(self):
        """The kernels run over the canonical task tables of A and A^T (chunks of
        at least 32 blocks -- the heads multiply the partial-tile traffic of
        split windows -- fixed by the structure's size): results do not depend
        on part_size.  The reference-named attributes keep the structure."""
        for name in ("fwd", "bwd"):
            rp, col, emap, win, _ = getattr(self, name)
            chunk = self.chunk_blocks or max(ops.canonical_chunk(col.numel(), self.wide), 32)
            t_rp, t_win, t_at = ops.canonical_tasks(rp, win, col.numel(), self.wide, chunk)
            setattr(self, name, [t_rp, col, emap, t_win, t_at])

    def _compact_bwd(self):
        if self.wide == 32:
            c = ops.compact_tiles_device(self.bwd[2], self.bwd[1].numel(), self.window, positions=False)
            self.cmap_bwd = [c["masks"], c["block_base"], c["values"]]

    def aggregate(self, p, X, transpose=False, m_ori=None, out_dtype=torch.float32, compact=True):
        """Y[i] = sum_j P[i, j] X[j] over A (``transpose``: over A^T) with P's
        entries p[h, e] in A's CSR edge order; X and Y [N, H, D].  Runs over
        the compact edge-id lists when the graph has them (bitwise equal to
        the dense edge-id map tiles, which ``compact=False`` selects)."""
        rp, col, emap, win, at = self.bwd if transpose else self.fwd
        cm = self.cmap_bwd if transpose else self.cmap_fwd
        m = self.num_nodes_ori if m_ori is None else m_ori
        if compact and cm is not None:
            return ops.spmm_edge_map_compact(rp, col, *cm, p, win, at, X, "mhd", self.window, m_ori=m,
                                             x_layout="mhd", out_dtype=out_dtype, rows_covered=True)
        return ops.spmm_edge_map(rp, col, emap, p, win, at, X, "mhd", self.window, self.wide, m_ori=m,
                                 x_layout="mhd", out_dtype=out_dtype, rows_covered=True)

    def segments(self):
        """Row segments of A and of A^T for the per-edge attention kernels
        (built on first use, kept on the device)."""
        if getattr(self, "_segments", None) is None:
            rp = self.rowptr.cpu().numpy()
            trp = self.t_rowptr.cpu().numpy()
            dev = self.rowptr.device
            self._segments = (ops.DeviceSegments(G.csr_segments(rp), dev),
                              ops.DeviceSegments(G.csr_segments(trp), dev))
        return self._segments

    def long_row_lists(self):
        """Rows of A and of A^T with more than 256 entries (int32, on the
        device), which the edge-softmax kernels give a workgroup each; built on
        first use from a host copy of the row pointers."""
        if getattr(self, "_long_rows", None) is None:
            lr = ops.long_rows(self.rowptr)
            tl = lr if self.t_rowptr is self.rowptr else ops.long_rows(self.t_rowptr)
            self._long_rows = (lr, tl)
        return self._long_rows

    # the A^T aggregation and the SDDMM of the attention backward in one kernel
    fuse_backward = True

    def aggregate_sddmm(self, p, dh, Xt, m_ori=None, out_dtype=torch.float32):
        """(aggregate(p, dh, transpose=True), sddmm(dh, Xt)): dXt[j] = sum_i
        p[(i,j)] dh[i] over A^T, and dp[h, e=(i,j)] = <dh[i,h], Xt[j,h]> in
        A's edge order (fp32).  With compact edge lists and fp16 / bf16
        features of 32, 64 or 128 per head, one kernel over A^T's structure
        computes both from the same gathered rows of dh (bitwise equal to the
        two separate products); otherwise the two products run separately."""
        n, heads, dim = Xt.shape
        m = self.num_nodes_ori if m_ori is None else m_ori
        if (self.fuse_backward and self.cmap_bwd is not None and dh.dtype in (torch.float16, torch.bfloat16)
                and Xt.dtype == dh.dtype and dh.shape == Xt.shape and dim in (32, 64, 128)):
            rp, col, _, win, at = self.bwd
            return ops.spmm_edge_map_compact_dot(rp, col, *self.cmap_bwd, p, win, at, dh.contiguous(),
                                                 Xt.contiguous(), "mhd", self.window, m_ori=m, x_layout="mhd",
                                                 out_dtype=out_dtype, rows_covered=True)
        return self.aggregate(p, dh, transpose=True, m_ori=m, out_dtype=out_dtype), self.sddmm(dh, Xt)

    def sddmm(self, A, B, out_dtype=torch.float32):
        """out[h, e=(i,j)] = <A[i,h], B[j,h]> over A's edges ([N, H, D] inputs, [H, nnz] out)."""
        rp, col, emap, win, _ = self.fwd
        if self.edges is not None:
            base, eids, pos = self.edges
            return ops.sddmm_edges(rp, col, base, eids, pos, win, A.contiguous(), B.contiguous(), self.num_edges,
                                   self.window, layout="mhd", out_dtype=out_dtype, covered=True)
        return ops.sddmm_block(rp, col, emap, win, A.contiguous(), B.contiguous(), self.num_edges, self.window,
                               self.wide, layout="mhd", out_dtype=out_dtype)


class GATAttention(torch.autograd.Function):
    """h[i,h,:] = sum_j softmax_j(LeakyReLU(a0_h.Xs[i,h] + a1_h.Xt[j,h])) Xt[j,h,:]
    with dropout on the probabilities.  Xs/Xt [N, H, D]; a0/a1 [H, D].

    GAT: Xs = Xt = X W, slope 0.2.  Reference GATv2: Xs = LReLU(X W0),
    Xt = LReLU(X W1), a0 = a1 = a, slope 1 (no LeakyReLU on the logits).
    fused=True: the softmax writes only its log-sum-exp and both aggregations
    rebuild the probabilities in the SpMM (attgnn_spmm_attention); bitwise
    equal to the default fused=False, which stores p [H, nnz] between them.
    Rebuilding costs time (4 heads x 64 on a 2.48M-entry graph: aggregation
    282 -> 381 us forward, 357 -> 403 us backward) and saves the H * nnz
    probabilities, so it is the choice when those do not fit.

    minmax=True: the logits of each head are first mapped to
    (z - min_h) / (max_h - min_h), min and max over all edges of the head, as
    the reference's multi-head GATv2 does before its exponential
    (end2end_head/gatv2_final/mgat_csr/mgat_conv.py:184-188).  The softmax is
    shift-invariant, so this is a per-head temperature 1 / (max_h - min_h):
    s0 and s1 are scaled by it (LeakyReLU commutes with a positive scale) and
    the gradient through the temperature reaches the two edges that set min_h
    and max_h.
    """

    @staticmethod
    def forward(ctx, Xs, Xt, a0, a1, graph, slope, drop_p, seed, fused=False, minmax=False):
        g = graph
        # node-major [N, H] scores: one load per neighbour for all heads; one
        # pass over X when both terms read the same features (GAT)
        if Xs is Xt:
            s0, s1 = ops.node_scores_pair(Xs, a0.to(Xs.dtype).contiguous(), a1.to(Xs.dtype).contiguous(), "mhd",
                                          out_layout="nh")
        else:
            s0 = ops.node_scores(Xs, a0.to(Xs.dtype).contiguous(), "mhd", out_layout="nh")
            s1 = ops.node_scores(Xt, a1.to(Xt.dtype).contiguous(), "mhd", out_layout="nh")
        raw = None
        if minmax:
            rng, where = ops.edge_score_range(g.rowptr, g.col, s0, s1, s_layout="nh",
                                              long_rows=g.long_row_lists()[0])
            raw = (s0, s1, rng, where)
            scale = rng[:, 2]
            s0, s1 = s0 * scale, s1 * scale
        long_rows, _ = g.long_row_lists()
        rp, col, emap, win, at = g.fwd
        if fused:
            # the softmax only produces the row log-sum-exp; the aggregation
            # rebuilds each probability where it multiplies (never stored)
            p, lse = ops.edge_softmax_fwd(g.rowptr, g.col, s0, s1, slope, drop_p, seed, s_layout="nh",
                                          long_rows=long_rows, lse_only=True)
            h = ops.spmm_attention(rp, col, emap, win, at, Xt, "mhd", g.window, g.wide, s0, s1, row_offset=lse,
                                   slope=slope, drop_p=drop_p, seed=seed, m_ori=g.num_nodes_ori, x_layout="mhd",
                                   out_dtype=Xt.dtype, rows_covered=True)
            p = lse.new_empty(0)
        else:
            p, lse = ops.edge_softmax_fwd(g.rowptr, g.col, s0, s1, slope, drop_p, seed, p_dtype=Xt.dtype,
                                          s_layout="nh", long_rows=long_rows)
            h = g.aggregate(p, Xt, out_dtype=Xt.dtype)
        ctx.save_for_backward(Xs, Xt, a0, a1, s0, s1, lse, p, *(raw or ()))
        ctx.graph, ctx.slope, ctx.drop_p, ctx.seed, ctx.fused = g, slope, drop_p, seed, fused
        ctx.shared = Xs is Xt
        ctx.minmax = minmax
        return h

    @staticmethod
    def backward(ctx, dh):
        Xs, Xt, a0, a1, s0, s1, lse, p = ctx.saved_tensors[:8]
        g = ctx.graph
        n, heads, dim = Xt.shape
        dh = dh.to(Xt.dtype).contiguous()
        # aggregation: dXt[j] = sum_i p[(i,j)] dh[i]  (A^T SpMM; with fused
        # probabilities, row j's term is s1[j], column i's are s0[i] and lse[i])
        rp, col, emap_t, win, at = g.bwd
        if ctx.fused:
            dXt = ops.spmm_attention(rp, col, emap_t, win, at, dh, "mhd", g.window, g.wide, s1, s0, col_offset=lse,
                                     slope=ctx.slope, drop_p=ctx.drop_p, seed=ctx.seed, m_ori=n, x_layout="mhd",
                                     rows_covered=True)
            # dp[(i,j)] = <dh[i], Xt[j]>  (SDDMM over A)
            dp = g.sddmm(dh, Xt)
        else:
            # with the SDDMM dp[(i,j)] = <dh[i], Xt[j]> from the same gathered rows
            dXt, dp = g.aggregate_sddmm(p, dh, Xt, m_ori=n)
        # row and column sums of dL/dz (columns through the transpose permutation)
        long_rows, t_long_rows = g.long_row_lists()
        ds0, ds1 = ops.edge_softmax_bwd_sums(g.rowptr, g.col, s0, s1, lse, dp, ctx.slope, ctx.drop_p, ctx.seed,
                                             g.t_rowptr, g.t_perm, s_layout="nh", long_rows=long_rows,
                                             t_long_rows=t_long_rows)
        if ctx.minmax:
            s0r, s1r, rng, where = ctx.saved_tensors[8:]
            ds0, ds1 = ops.minmax_score_grads(s0r, s1r, rng, where, ds0, ds1, s_layout="nh")
        # s0 = Xs.a0, s1 = Xt.a1: node-term gradients fused with the aggregation
        # gradient (one pass over X; GAT's shared X W gets the summed gradient)
        shared = ctx.shared and Xs.data_ptr() == Xt.data_ptr()
        dXs, dXt, da0, da1 = ops.attention_node_bwd(Xs.contiguous(), Xt.contiguous(), a0, a1, ds0, ds1,
                                                    base=dXt, combine=shared)
        return (dXs, dXt, da0.to(a0.dtype).reshape(a0.shape), da1.to(a1.dtype).reshape(a1.shape),
                None, None, None, None, None, None)


class GATv2Attention(torch.autograd.Function):
    """GATv2 as its paper defines it: h[i,h,:] = sum_j softmax_j(a_h .
    LeakyReLU(X0[i,h] + X1[j,h])) X1[j,h,:], dropout on the probabilities.
    X0 (target side) and X1 (source side, the messages) [N, H, D]; a [H, D]."""

    @staticmethod
    def forward(ctx, X0, X1, a, graph, slope, drop_p, seed):
        g = graph
        segs, _ = g.segments()
        z, lse, p = ops.gatv2_fwd(g.rowptr, g.col, segs, X0, X1, a, slope, drop_p, seed, p_dtype=X1.dtype)
        h = g.aggregate(p, X1, out_dtype=X1.dtype)
        ctx.save_for_backward(X0, X1, a, z, lse, p)
        ctx.graph, ctx.slope, ctx.drop_p, ctx.seed = g, slope, drop_p, seed
        return h

    @staticmethod
    def backward(ctx, dh):
        X0, X1, a, z, lse, p = ctx.saved_tensors
        g = ctx.graph
        n = X1.shape[0]
        dh = dh.to(X1.dtype).contiguous()
        # messages: dX1[j] = sum_i p[(i,j)] dh[i]  (A^T edge-map SpMM), and
        # dL/dp[(i,j)] = <dh[i], X1[j]> from the same gathered rows
        base, dp = g.aggregate_sddmm(p, dh, X1, m_ori=n)
        segs, t_segs = g.segments()
        _, dX0, dX1, da = ops.gatv2_bwd(g.rowptr, g.col, segs, g.t_col, g.t_perm, t_segs, X0, X1, a, ctx.slope, z,
                                        lse, dp, ctx.drop_p, ctx.seed, base=base)
        return dX0, dX1, da.to(a.dtype).reshape(a.shape), None, None, None, None


class GATConv_multi(nn.Module):
    """end2end_head/gat_final/mgat_csr/mgat_conv.py:154-199 (forward returns [N, H, D])."""

    def __init__(self, input_dim, output_dim, dropout, alpha, head=1, dtype=torch.bfloat16):
        super().__init__()
        self.alpha, self.head, self.output_dim, self.dtype = alpha, head, output_dim, dtype
        gain = nn.init.calculate_gain("relu")
        self.fc = nn.Linear(input_dim, output_dim * head, bias=False)
        nn.init.xavier_normal_(self.fc.weight.data, gain=gain)
        self.a0 = nn.Parameter(torch.zeros(head, 1, output_dim))
        nn.init.xavier_normal_(self.a0.data, gain=gain)
        self.a1 = nn.Parameter(torch.zeros(head, 1, output_dim))
        nn.init.xavier_normal_(self.a1.data, gain=gain)
        self.dropout = dropout
        self._step = 0
        self.fused = False    # True: probabilities rebuilt in the aggregation, not stored (see GATAttention)

    def forward(self, X, inputInfo, seed=None):
        Xp = LinearSplitK.apply(X.to(self.dtype), self.fc.weight).view(-1, self.head, self.output_dim)
        drop = self.dropout if self.training else 0.0
        if seed is None:
            seed = self._step
            self._step += 1
        return GATAttention.apply(Xp, Xp, self.a0.view(self.head, -1), self.a1.view(self.head, -1), inputInfo,
                                  self.alpha, drop, seed, self.fused, False)


class GATv2Conv_multi(nn.Module):
    """The reference's GATv2 layers: X0 = LReLU(X W0), X1 = LReLU(X W1),
    logit z = a.X0[i] + a.X1[j], h = sum_j p X1[j] with p the row softmax.

    Two reference files define a class of this name.  Constructed with a head
    count (``GATv2Conv_multi(in, out, dropout, alpha, head)``) it is the
    multi-head one, end2end_head/gatv2_final/mgat_csr/mgat_conv.py:151-198,
    which maps each head's logits to (z - min) / (max - min) over all its edges
    before the exponential (:184-186; ``minmax=True``).  Without a head count
    it is the single-head one, end2end_ori/gatv2_final/mgat_csr/mgat_conv.py:
    150-202, whose logits go to the exponential as they are (``minmax=False``).
    ``minmax`` overrides either default.

    ``true_gatv2=True``: the attention of the GATv2 paper instead,
    X0 = X W0, X1 = X W1, logit = a . LReLU(X0[i] + X1[j]), h = sum p X1
    (fused per-edge kernels; no counterpart in the reference)."""

    def __init__(self, input_dim, output_dim, dropout, alpha, head=None, dtype=torch.bfloat16, true_gatv2=False,
                 minmax=None):
        super().__init__()
        self.minmax = (head is not None) if minmax is None else bool(minmax)
        head = 1 if head is None else head
        self.true_gatv2 = true_gatv2
        self.alpha, self.head, self.output_dim, self.dtype = alpha, head, output_dim, dtype
        gain = nn.init.calculate_gain("relu")
        self.fc0 = nn.Linear(input_dim, output_dim * head, bias=False)
        nn.init.xavier_normal_(self.fc0.weight.data, gain=gain)
        self.fc1 = nn.Linear(input_dim, output_dim * head, bias=False)
        nn.init.xavier_normal_(self.fc1.weight.data, gain=gain)
        self.a0 = nn.Parameter(torch.zeros(head, 1, output_dim))
        nn.init.xavier_normal_(self.a0.data, gain=gain)
        self.dropout = dropout
        self._step = 0
        self.fused = False    # reference logits; True: probabilities rebuilt in the aggregation

    def forward(self, X, inputInfo, seed=None):
        Xd = X.to(self.dtype)
        drop = self.dropout if self.training else 0.0
        if seed is None:
            seed = self._step
            self._step += 1
        a = self.a0.view(self.head, -1)
        if self.true_gatv2:
            X0 = LinearSplitK.apply(Xd, self.fc0.weight).view(-1, self.head, self.output_dim)
            X1 = LinearSplitK.apply(Xd, self.fc1.weight).view(-1, self.head, self.output_dim)
            return GATv2Attention.apply(X0, X1, a, inputInfo, self.alpha, drop, seed)
        X0 = F.leaky_relu(LinearSplitK.apply(Xd, self.fc0.weight), self.alpha).view(-1, self.head, self.output_dim)
        X1 = F.leaky_relu(LinearSplitK.apply(Xd, self.fc1.weight), self.alpha).view(-1, self.head, self.output_dim)
        return GATAttention.apply(X0, X1, a, a, inputInfo, 1.0, drop, seed, self.fused, self.minmax)
