"""A sparse matrix in window-block form on the device, in one arithmetic mode.

Modes (all accumulate in FP32; Y is FP32):
  "f32"    fp32 tiles and X, fp32 MFMA (16x16x4)
  "bf16x3" fp32 tiles and X, each split into bf16 hi + lo; A.X ~ AhXh + AhXl + AlXh
           (replaces the reference's TF32 path, gcn/SpMM/src/spmmKernel.cu:959-1075,
           on hardware without TF32)
  "bf16" / "fp16"  tiles and X rounded to the 16-bit type
  "fp8"    OCP e4m3: A scaled per row (r_i = max_j |A_ij| / 448), X per feature
           column (c_f = max_i |X_if| / 448); Y = diag(r) (A' X') diag(c), the
           scales applied at the store.
"""
import numpy as np
import torch

from . import ops

FP8_MAX = 448.0
_TILE = {"f32": torch.float32, "bf16x3": torch.float32, "bf16": torch.bfloat16, "fp16": torch.float16,
         "fp8": torch.float32}
_XT = {"f32": torch.float32, "bf16x3": torch.float32, "bf16": torch.bfloat16, "fp16": torch.float16,
       "fp8": torch.float8_e4m3fn}
_KMODE = {"f32": "f32", "bf16x3": "bf16x3", "bf16": "native", "fp16": "native", "fp8": "native"}


class BlockMatrix:
    def __init__(self, csr, values, device, mode="bf16", window=16, wide=32, part_size=None, build="host",
                 chunk_blocks=None):
        if mode not in _TILE:
            raise ValueError(f"unknown mode {mode!r}")
        from .layers import _blocks_estimate, default_part_size
        self.mode, self.window, self.wide, self.n_rows = mode, window, wide, csr.n
        values = np.asarray(values, np.float32)
        self.part_size = part_size or default_part_size(_blocks_estimate(csr, wide))
        self.row_scale = None
        if mode == "fp8":
            amax = np.zeros(csr.n, np.float32)
            np.maximum.at(amax, csr.row_ids(), np.abs(values))
            r = np.where(amax > 0, amax / FP8_MAX, 1.0).astype(np.float32)
            values = values / r[csr.row_ids()]
            self.row_scale = torch.from_numpy(r).to(device)
        padded = csr.padded(window)
        if build == "device":      # bit-identical to the host build
            b = ops.blocks_build_device(torch.from_numpy(padded.rowptr32()).to(device),
                                        torch.from_numpy(padded.indices).to(device), "edge_f32",
                                        torch.from_numpy(values).to(device), _TILE[mode], window, wide,
                                        self.part_size, emit_empty=True)
        else:
            b = ops.blocks_build(padded.rowptr32(), padded.indices, "edge_f32", values, _TILE[mode], window, wide,
                                 self.part_size, emit_empty=True)
        self.n_unique = b["n_unique"]
        arrs = [b[k].to(device) for k in ("rp", "col", "val", "window", "atomic")]
        if mode == "fp8":
            arrs[2] = arrs[2].to(torch.float8_e4m3fn)
        # the product runs over the canonical task table (results independent of part_size)
        t_rp, t_win, t_at = ops.canonical_tasks(arrs[0], arrs[3], arrs[1].numel(), wide, chunk_blocks)
        arrs = [t_rp, arrs[1], arrs[2], t_win, t_at]
        self.arrays = arrs
        self.parts = int(b["window"].numel())

    def quantize_x(self, X):
        """X in the mode's storage type (and the per-column scales for fp8)."""
        if self.mode != "fp8":
            return X.to(_XT[self.mode]).contiguous(), None
        c = X.abs().amax(dim=0).float() / FP8_MAX
        c = torch.where(c > 0, c, torch.ones_like(c))
        return (X.float() / c).to(torch.float8_e4m3fn).contiguous(), c.contiguous()

    def matmul(self, X, Xq=None, col_scale=None, out_dtype=torch.float32):
        if Xq is None:
            Xq, col_scale = self.quantize_x(X)
        rp, col, val, win, at = self.arrays
        return ops.spmm_block(rp, col, val, win, at, Xq, self.n_rows, self.window, self.wide,
                              mode=_KMODE[self.mode], row_scale=self.row_scale, col_scale=col_scale,
                              out_dtype=out_dtype, rows_covered=True)
