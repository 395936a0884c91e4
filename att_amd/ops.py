"""Torch-facing wrappers over the C ABI (libattgnn.so).

Each wrapper validates device / dtype / contiguity / shape on the host,
allocates outputs and scratch through torch's caching allocator, and launches
on torch's current HIP stream.  Nothing here computes on the CPU except the
window-block preprocessing, which is host code in the reference as well.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

_TYPE_OF = {
    torch.float32: _lib.F32,
    torch.float16: _lib.F16,
    torch.bfloat16: _lib.BF16,
    torch.float8_e4m3fn: _lib.FP8E4M3,
    torch.int32: _lib.I32,
}
_DTYPE_OF = {v: k for k, v in _TYPE_OF.items()}
_MODES = {"native": _lib.MODE_NATIVE, "bf16x3": _lib.MODE_BF16X3, "f32": _lib.MODE_F32}


def _mode(mode, dtype):
    """Arithmetic mode: MFMA in the element type for 16/8-bit data, bf16x3 for fp32."""
    if mode is None:
        mode = "bf16x3" if dtype == torch.float32 else "native"
    return _MODES[mode]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


# torch's raw accessor of a device's current stream (what its own kernels
# launch on); torch.cuda.current_stream builds a Stream object and resolves
# the device on every call, which costs more than many of the launches here
_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream(device):
    """The current stream of ``device`` as a hipStream_t."""
    if _RAW_STREAM is not None and device.index is not None:
        return ctypes.c_void_p(_RAW_STREAM(device.index))
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _need_gpu(name, *tensors):
    dev = None
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name}: expected torch.Tensor, got {type(t)}")
        if t.device.type != "cuda":
            raise ValueError(f"{name}: tensors must be on the GPU (got {t.device})")
        if not t.is_contiguous():
            raise ValueError(f"{name}: tensors must be contiguous")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError(f"{name}: tensors on different devices")
    return dev


def _need_i32(name, *tensors):
    for t in tensors:
        if t.dtype != torch.int32:
            raise TypeError(f"{name}: index arrays must be int32 (got {t.dtype})")


def check_part_spans(rp, wide, name="structure"):
    """Raise unless every part's condensed-column span rp[2p+2] - rp[2p] is a
    multiple of ``wide``: tiles and edge-id maps are laid out in blocks of
    ``wide`` columns, so a structure built with another block width would be
    read at the wrong slots (and, for a span that is not a multiple, past its
    end).  One device-to-host copy of ``rp`` per structure: the result is kept
    on the tensor with its version counter and reused until ``rp`` is written
    in place by a torch op (a raw-pointer refill does not move the counter, so
    such buffers are checked once).  Skipped while a stream is being captured
    into a graph (no host copy is possible there; captures replay structures
    checked when eager)."""
    if rp.numel() < 3:
        return
    try:
        version = rp._version
    except RuntimeError:          # inference tensors carry no version counter
        version = None
    tag = (int(wide), version, rp.numel())
    if version is not None and getattr(rp, "_attgnn_spans_ok", None) == tag:
        return
    if rp.is_cuda and torch.cuda.is_current_stream_capturing():
        return
    r = rp.detach().to("cpu", torch.int64)
    bad = torch.remainder(r[2::2] - r[:-2:2], int(wide)).nonzero()
    if bad.numel():
        p = int(bad[0, 0])
        raise ValueError(f"{name}: part {p} spans {int(r[2 * p + 2] - r[2 * p])} condensed columns, not a multiple "
                         f"of wide={wide}; was the structure built with another block width?")
    if version is not None:
        try:
            rp._attgnn_spans_ok = tag
        except (AttributeError, RuntimeError):
            pass


# --------------------------------------------------------------------------
# preprocessing (host)
# --------------------------------------------------------------------------

def blocks_build(row, col, src, src_values, out_dtype, window, wide, part_size,
                 emit_empty, want_trans=False, nthreads=0, compact=False):
    """Window-block structure of a CSR matrix (host, OpenMP).

    ``src``: "deg_f16" | "deg_f32" | "edge_f32" | "map".  Returns a dict with
    CPU int32 tensors ``rp``, ``col``, ``window``, ``atomic``, the tile values
    ``val`` (``out_dtype``; int32 edge ids for "map"), ``map_t`` (want_trans)
    and ``n_unique`` (sum over windows of distinct columns).
    """
    srcs = {"deg_f16": _lib.SRC_DEG_F16, "deg_f32": _lib.SRC_DEG_F32,
            "edge_f32": _lib.SRC_EDGE_F32, "map": _lib.SRC_MAP}
    if src not in srcs:
        raise ValueError(f"unknown src {src!r}")
    row = torch.as_tensor(row).to(torch.int32).contiguous().cpu()
    col = torch.as_tensor(col).to(torch.int32).contiguous().cpu()
    if row.dim() != 1 or col.dim() != 1:
        raise ValueError("row and col must be 1-D")
    keep = None
    if src == "map":
        sp = ctypes.c_void_p(0)
        out_dtype = torch.int32
    else:
        want = {"deg_f16": torch.float16, "deg_f32": torch.float32, "edge_f32": torch.float32}[src]
        keep = torch.as_tensor(src_values).to(want).contiguous().cpu()
        expect = row.numel() - 1 if src.startswith("deg") else col.numel()
        if keep.numel() < expect:
            raise ValueError(f"{src}: need {expect} values, got {keep.numel()}")
        sp = ctypes.c_void_p(keep.data_ptr())
    handle = ctypes.c_void_p()
    rc = lib().attgnn_blocks_build(_ptr(row), row.numel(), _ptr(col), col.numel(), srcs[src], sp,
                                   _TYPE_OF[out_dtype], window, wide, part_size, int(bool(emit_empty)),
                                   int(bool(want_trans)), nthreads, ctypes.byref(handle))
    check(rc, "attgnn_blocks_build")
    try:
        n = [ctypes.c_int64() for _ in range(5)]
        has_t = ctypes.c_int()
        check(lib().attgnn_blocks_sizes(handle, *[ctypes.byref(x) for x in n], ctypes.byref(has_t)),
              "attgnn_blocks_sizes")
        n_rp, n_col, n_val, n_parts, n_unique = [x.value for x in n]
        out = {
            "rp": torch.empty(n_rp, dtype=torch.int32),
            "col": torch.empty(n_col, dtype=torch.int32),
            "val": torch.empty(n_val, dtype=out_dtype),
            "window": torch.empty(n_parts, dtype=torch.int32),
            "atomic": torch.empty(n_parts, dtype=torch.int32),
            "map_t": torch.empty(n_val, dtype=torch.int32) if has_t.value else None,
        }
        check(lib().attgnn_blocks_export(handle, _ptr(out["rp"]), _ptr(out["col"]), _ptr(out["val"]),
                                         _ptr(out["map_t"]), _ptr(out["window"]), _ptr(out["atomic"])),
              "attgnn_blocks_export")
        out["n_unique"] = n_unique
        if compact:
            nb, nv = ctypes.c_int64(), ctypes.c_int64()
            check(lib().attgnn_blocks_compact_sizes(handle, ctypes.byref(nb), ctypes.byref(nv)),
                  "attgnn_blocks_compact_sizes")
            out["masks"] = torch.empty(max(n_col, 8), dtype=torch.int16)
            out["block_base"] = torch.empty(nb.value + 1, dtype=torch.int64)
            out["values"] = torch.empty(max(nv.value, 1), dtype=out_dtype)
            out["positions"] = torch.empty(max(nv.value, 1), dtype=torch.int16)
            check(lib().attgnn_blocks_compact_export(handle, _ptr(out["masks"]), _ptr(out["block_base"]),
                                                     _ptr(out["values"]), _ptr(out["positions"])),
                  "attgnn_blocks_compact_export")
    finally:
        lib().attgnn_blocks_free(handle)
    return out


_SRCS = {"deg_f16": _lib.SRC_DEG_F16, "deg_f32": _lib.SRC_DEG_F32, "edge_f32": _lib.SRC_EDGE_F32,
         "map": _lib.SRC_MAP}


class DevicePlan:
    """The planning phase of a device build (attgnn_blocks_plan_device) of one
    CSR: its workspace and the structure sizes read back once.  Every fill of
    the same (row, col, window, wide, part_size, emit_empty) reuses it -- the
    fill only reads the plan -- so a graph needing several tile types (values
    and an edge-id map) pays the read-back once.

    ``sync=False``: the sizes are copied to pinned host memory behind an event
    instead of read at once; :meth:`ready` polls it and :meth:`wait` (called
    by :meth:`fill`) blocks until they arrive, so the host can do other work
    while the plan runs."""

    def __init__(self, row, col, window, wide, part_size, emit_empty, sync=True):
        dev = _need_gpu("blocks_plan_device", row, col)
        _need_i32("blocks_plan_device", row, col)
        if row.dim() != 1 or col.dim() != 1 or row.numel() < 1:
            raise ValueError("row and col must be 1-D, row non-empty")
        self.row, self.col, self.dev = row, col, dev
        self.window, self.wide, self.part_size, self.emit_empty = window, wide, part_size, bool(emit_empty)
        rows1, nnz = row.numel(), col.numel()
        L = lib()
        ws_n = L.attgnn_blocks_device_workspace_bytes(rows1, nnz, window)
        if ws_n < 0:
            raise ValueError("invalid sizes for the device preprocessing")
        self.ws = torch.empty(max(ws_n, 1), dtype=torch.uint8, device=dev)
        self.ws_n = ws_n
        totals = torch.empty(8, dtype=torch.int64, device=dev)
        check(L.attgnn_blocks_plan_device(_ptr(row), rows1, _ptr(col), nnz, window, wide, part_size,
                                          int(self.emit_empty), _ptr(self.ws), ws_n, _ptr(totals), _stream(dev)),
              "attgnn_blocks_plan_device")
        self._pending = None
        if sync:
            self._sizes([int(x) for x in totals[:6].cpu()])
        else:
            host = torch.empty(8, dtype=torch.int64, pin_memory=True)
            host.copy_(totals, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(torch.cuda.current_stream(dev))
            self._pending = (host, ev, totals)

    def _sizes(self, t):
        status, self.n_rp, self.n_col, self.n_val, self.n_parts, self.n_unique = t
        check(status, "attgnn_blocks_plan_device")

    def ready(self):
        """True once the sizes are on the host (always for sync=True)."""
        return self._pending is None or self._pending[1].query()

    def wait(self):
        """Block until the sizes are on the host and take them."""
        if self._pending is not None:
            host, ev, _ = self._pending
            ev.synchronize()
            self._pending = None
            self._sizes([int(x) for x in host[:6].tolist()])

    def fill(self, src, src_values, out_dtype, want_trans=False, compact=False, n_values=None, positions=True,
             tiles_only=False):
        """One structure of this plan (keys as :func:`blocks_build`).
        ``n_values``: an upper bound of the compact value count when the
        caller knows one (an edge-id map holds each CSR entry at most once:
        nnz), which saves its read-back; see :func:`compact_tiles_device`.
        ``positions=False`` skips the compact form's per-entry positions.
        ``tiles_only``: no part table and no condensed columns (they equal
        those of every other fill of this plan); the result lacks rp, col,
        window and atomic."""
        self.wait()
        if src not in _SRCS:
            raise ValueError(f"unknown src {src!r}")
        if want_trans and src != "map":
            raise ValueError("want_trans needs src='map'")
        dev = self.dev
        if src == "map":
            out_dtype, sp, keep = torch.int32, ctypes.c_void_p(0), None
        else:
            _need_gpu("blocks_fill_device", src_values)
            want = {"deg_f16": torch.float16, "deg_f32": torch.float32, "edge_f32": torch.float32}[src]
            if src_values.dtype != want:
                raise TypeError(f"{src}: values must be {want}, got {src_values.dtype}")
            expect = self.row.numel() - 1 if src.startswith("deg") else self.col.numel()
            if src_values.numel() < expect:
                raise ValueError(f"{src}: need {expect} values, got {src_values.numel()}")
            if out_dtype not in (torch.float32, torch.float16, torch.bfloat16):
                raise TypeError(f"tile type {out_dtype} is not supported")
            keep = src_values
            sp = _ptr(keep)
        out = {"val": torch.empty(self.n_val, dtype=out_dtype, device=dev),
               "map_t": torch.empty(self.n_val, dtype=torch.int32, device=dev) if want_trans else None,
               "n_unique": self.n_unique}
        if not tiles_only:
            out.update(rp=torch.empty(self.n_rp, dtype=torch.int32, device=dev),
                       col=torch.empty(self.n_col, dtype=torch.int32, device=dev),
                       window=torch.empty(self.n_parts, dtype=torch.int32, device=dev),
                       atomic=torch.empty(self.n_parts, dtype=torch.int32, device=dev))
        check(lib().attgnn_blocks_fill_device(_ptr(self.row), self.row.numel(), _ptr(self.col), self.col.numel(),
                                              _SRCS[src], sp, _TYPE_OF[out_dtype], self.window, self.wide,
                                              self.part_size, _ptr(self.ws), self.ws_n, _ptr(out.get("rp")),
                                              _ptr(out.get("col")), _ptr(out["val"]), _ptr(out["map_t"]),
                                              _ptr(out.get("window")), _ptr(out.get("atomic")), _stream(dev)),
              "attgnn_blocks_fill_device")
        if compact:
            if self.wide != 32 or self.window > 16:
                raise ValueError("compact tiles need wide = 32 and window <= 16")
            out.update(compact_tiles_device(out["val"], self.n_col, self.window, positions=positions,
                                            n_values=n_values))
        del keep
        return out


def blocks_build_device(row, col, src, src_values, out_dtype, window, wide, part_size, emit_empty,
                        want_trans=False, compact=False):
    """Window-block structure built on the GPU from a device-resident CSR.

    Same arguments and same result keys as :func:`blocks_build`, bit-identical
    values, but every array is a device tensor.  ``row`` / ``col`` are int32
    device tensors; ``src_values`` (degrees or edge values) a device tensor.
    Reads the structure's sizes back once (and the compact value count once).
    """
    if src not in _SRCS:
        raise ValueError(f"unknown src {src!r}")
    if want_trans and src != "map":
        raise ValueError("want_trans needs src='map'")
    if src != "map":
        _need_gpu("blocks_build_device", src_values)
    return DevicePlan(row, col, window, wide, part_size, emit_empty).fill(src, src_values, out_dtype, want_trans,
                                                                          compact)


def compact_tiles_device(val, n_col, window, workspace=None, positions=True, n_values=None):
    """Compact form of n_col condensed columns of window x 32 tiles on the
    device (value tiles, or an int32 edge-id map): masks, block_base, values
    and (optionally) positions, as attgnn_blocks_compact_export lays them out.
    Reads the value count back once, unless the caller gives it: ``n_values``
    must then be at least the count (the fill writes that many values; an
    upper bound such as the CSR's nnz for an edge-id map only leaves unused
    entries at the end)."""
    dev = _need_gpu("compact_tiles_device", val)
    L = lib()
    st = _stream(dev)
    nb = n_col // 32
    masks = torch.empty(max(n_col, 8), dtype=torch.int16, device=dev)
    base = torch.empty(nb + 1, dtype=torch.int64, device=dev)
    ws_n = L.attgnn_blocks_compact_workspace_bytes(n_col)
    ws = workspace if workspace is not None and workspace.numel() >= ws_n else \
        torch.empty(ws_n, dtype=torch.uint8, device=dev)
    check(L.attgnn_blocks_compact_count_device(_ptr(val), _TYPE_OF[val.dtype], n_col, window, _ptr(masks),
                                               _ptr(base), _ptr(ws), ws.numel(), st),
          "attgnn_blocks_compact_count_device")
    nv = int(base[nb].item()) if n_values is None else int(n_values)
    values = torch.empty(max(nv, 1), dtype=val.dtype, device=dev)
    pos = torch.empty(max(nv, 1), dtype=torch.int16, device=dev) if positions else None
    check(L.attgnn_blocks_compact_fill_device(_ptr(val), _TYPE_OF[val.dtype], n_col, window, _ptr(masks), _ptr(base),
                                              _ptr(values), _ptr(pos), st),
          "attgnn_blocks_compact_fill_device")
    out = dict(masks=masks, block_base=base, values=values)
    if positions:
        out["positions"] = pos
    return out


def map_transpose_perm(emap, emap_t, nnz):
    """int64 [nnz]: CSR position of the transpose of every entry, from an edge-id
    map and its transposed map (entries no slot covers stay -1)."""
    dev = _need_gpu("map_transpose_perm", emap, emap_t)
    _need_i32("map_transpose_perm", emap, emap_t)
    if emap.numel() != emap_t.numel():
        raise ValueError("map and transposed map differ in size")
    perm = torch.full((nnz,), -1, dtype=torch.int64, device=dev)
    check(lib().attgnn_map_transpose_perm_device(_ptr(emap), _ptr(emap_t), emap.numel(), nnz, _ptr(perm),
                                                 _stream(dev)), "attgnn_map_transpose_perm_device")
    return perm


def csr_transpose_device(rowptr, col, n_cols):
    """CSR of A^T on the device: (t_rowptr int32 [n_cols+1], t_col int32 [nnz]
    (source rows, ascending per row of A^T), t_perm int64 [nnz] (A positions));
    raises on a malformed CSR."""
    dev = _need_gpu("csr_transpose_device", rowptr, col)
    _need_i32("csr_transpose_device", rowptr, col)
    n_rows, nnz = rowptr.numel() - 1, col.numel()
    L = lib()
    ws_n = L.attgnn_csr_transpose_workspace_bytes(n_rows, n_cols, nnz)
    if ws_n < 0:
        raise ValueError("sizes out of range for the device transpose")
    ws = torch.empty(max(ws_n, 1), dtype=torch.uint8, device=dev)
    t_rowptr = torch.empty(n_cols + 1, dtype=torch.int32, device=dev)
    t_col = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)[:nnz]
    t_perm = torch.empty(max(nnz, 1), dtype=torch.int64, device=dev)[:nnz]
    status = torch.empty(1, dtype=torch.int64, device=dev)
    check(L.attgnn_csr_transpose_device(_ptr(rowptr), _ptr(col), n_rows, n_cols, nnz, _ptr(t_rowptr), _ptr(t_col),
                                        _ptr(t_perm), _ptr(status), _ptr(ws), ws_n, _stream(dev)),
          "attgnn_csr_transpose_device")
    check(int(status.item()), "attgnn_csr_transpose_device")
    return t_rowptr, t_col, t_perm


def map_remap(emap, remap):
    """map entries re-expressed through `remap` (int64); -1 stays -1."""
    dev = _need_gpu("map_remap", emap, remap)
    _need_i32("map_remap", emap)
    if remap.dtype != torch.int64:
        raise TypeError("remap must be int64")
    out = torch.empty_like(emap)
    check(lib().attgnn_map_remap_device(_ptr(emap), emap.numel(), _ptr(remap), remap.numel(), _ptr(out),
                                        _stream(dev)), "attgnn_map_remap_device")
    return out


# --------------------------------------------------------------------------
# canonical SpMM task table
# --------------------------------------------------------------------------

def canonical_chunk(n_cols, wide, lo=128, hi=16384):
    """Blocks of ``wide`` columns per SpMM task, from the structure's size
    alone (never from the partSize it was built with): n_cols / 8192 columns,
    clamped to [128, 16384] and rounded down to a multiple of 32 -- e.g. 4
    blocks of 32 for small graphs, 375 for 98.4M condensed columns."""
    cols = min(max(int(n_cols) // 8192, lo), hi) // 32 * 32
    return max(cols // int(wide), 1)


def canonical_tasks(rp, window_of_part, n_cols, wide, chunk_blocks=None):
    """(task_rp, task_window, task_atomic): each window's condensed columns cut
    into chunks of ``chunk_blocks`` * ``wide`` columns from the window start
    (attgnn_blocks_tasks_device), in the structure's part format with the slots
    past the total marked window -1.  An SpMM over these tasks groups every
    window's sum by the chunk size alone, so its result is bitwise the same
    whichever partSize the structure was built with.  Kept on ``rp`` (with the
    version counters of ``rp`` and ``window_of_part``) and reused until either
    is written in place by a torch op; a buffer refilled through a raw pointer
    (another extension, a kernel of this library) does not move the counter
    and must be passed as a new tensor or with ``rp._attgnn_tasks`` cleared."""
    dev = _need_gpu("canonical_tasks", rp, window_of_part)
    _need_i32("canonical_tasks", rp, window_of_part)
    parts = window_of_part.numel()
    if rp.numel() != 2 * parts + 1:
        raise ValueError("rp and window_of_part sizes disagree")
    chunk = int(chunk_blocks or canonical_chunk(n_cols, wide))
    try:
        version = (rp._version, window_of_part._version)
    except RuntimeError:
        version = None
    tag = (version, rp.numel(), window_of_part.data_ptr(), int(n_cols), int(wide), chunk)
    cached = getattr(rp, "_attgnn_tasks", None)
    if version is not None and cached is not None and cached[0] == tag:
        return cached[1]
    L = lib()
    tmax = L.attgnn_blocks_tasks_max(parts, int(n_cols), int(wide), chunk)
    if tmax < 0:
        raise ValueError("bad structure sizes")
    i32 = dict(dtype=torch.int32, device=dev)
    t_rp = torch.empty(2 * tmax + 1, **i32)
    t_win = torch.empty(max(tmax, 0), **i32)
    t_at = torch.empty(max(tmax, 0), **i32)
    ws = torch.empty(L.attgnn_blocks_tasks_workspace_bytes(parts), dtype=torch.uint8, device=dev)
    n_tasks = torch.empty(1, dtype=torch.int64, device=dev)
    check(L.attgnn_blocks_tasks_device(_ptr(rp), _ptr(window_of_part), parts, int(n_cols), int(wide), chunk,
                                       _ptr(t_rp), _ptr(t_win), _ptr(t_at), tmax, _ptr(ws), ws.numel(),
                                       _ptr(n_tasks), _stream(dev)),
          "attgnn_blocks_tasks_device")
    out = (t_rp, t_win, t_at)
    if version is not None:
        try:
            rp._attgnn_tasks = (tag, out)
        except (AttributeError, RuntimeError):
            pass
    return out


# --------------------------------------------------------------------------
# SpMM
# --------------------------------------------------------------------------

# k-loop schedules (ATTGNN_SPMM_PIPE_*); results are bitwise identical across them
_PIPES = {None: 0, "default": 0, "one": 1 << 4, "pair": 2 << 4, "none": 3 << 4}


def _spmm_flags(rows_covered, pipe, xcd):
    """SpMM flags: rows covered, k-loop schedule (``pipe``) and the per-XCD
    feature-lane mapping (``xcd=False`` turns it off; results are the same)."""
    return ((_lib.SPMM_ROWS_COVERED if rows_covered else 0) | _PIPES[pipe]
            | (0 if xcd else _lib.SPMM_NO_XCD_SLICES))


def spmm_block(rp, col, val, window_of_part, atomic, X, m_ori, window, wide, mode=None,
               row_scale=None, col_scale=None, out=None, out_dtype=torch.float32, rows_covered=False,
               workspace=None, pipe=None, xcd=True):
    """Y[m_ori, D] = A . X over a window-block structure (all on the GPU).

    FP32 accumulation; Y is written in ``out_dtype`` (fp32 / fp16 / bf16).
    ``rows_covered``: the structure has a part for every window up to m_ori,
    so Y needs no zero-fill.
    """
    dev = _need_gpu("spmm_block", rp, col, val, window_of_part, atomic, X, row_scale, col_scale, out)
    _need_i32("spmm_block", rp, col, window_of_part, atomic)
    if X.dim() != 2:
        raise ValueError("X must be 2-D [rows, dim]")
    if rp.numel() != 2 * window_of_part.numel() + 1 or atomic.numel() != window_of_part.numel():
        raise ValueError("rp / window / atomic sizes disagree")
    check_part_spans(rp, wide, "spmm_block")
    parts = window_of_part.numel()
    n_x, dim = X.shape
    y = out if out is not None else torch.empty((m_ori, dim), dtype=out_dtype, device=dev)
    if y.shape != (m_ori, dim) or y.dtype not in (torch.float32, torch.float16, torch.bfloat16):
        raise ValueError("out must be [m_ori, dim] of float32 / float16 / bfloat16")
    ws_n = lib().attgnn_spmm_workspace_elems(parts, window, dim) if parts else 0
    if workspace is not None and workspace.numel() >= ws_n and workspace.dtype == torch.float32:
        ws = workspace
    else:
        ws = torch.empty(max(ws_n, 1), dtype=torch.float32, device=dev)
    flags = _spmm_flags(rows_covered, pipe, xcd)
    rc = lib().attgnn_spmm_block(_ptr(rp), _ptr(col), _ptr(val), _TYPE_OF[val.dtype], _ptr(window_of_part),
                                 _ptr(atomic), parts, _ptr(X), _TYPE_OF[X.dtype], n_x, _ptr(y), _TYPE_OF[y.dtype],
                                 m_ori, dim, window, wide, _mode(mode, X.dtype), _ptr(row_scale), _ptr(col_scale), _ptr(ws),
                                 ws_n, flags, _stream(dev))
    check(rc, "attgnn_spmm_block")
    return y


def spmm_compact(rp, col, masks, block_base, values, window_of_part, atomic, X, m_ori, window, mode=None,
                 row_scale=None, col_scale=None, out_dtype=torch.float32, rows_covered=False, workspace=None,
                 pipe=None, xcd=True):
    """attgnn_spmm_compact: spmm_block over compact tiles (wide = 32)."""
    dev = _need_gpu("spmm_compact", rp, col, masks, block_base, values, window_of_part, atomic, X, row_scale,
                    col_scale)
    _need_i32("spmm_compact", rp, col, window_of_part, atomic)
    parts = window_of_part.numel()
    n_x, dim = X.shape
    y = torch.empty((m_ori, dim), dtype=out_dtype, device=dev)
    ws_n = lib().attgnn_spmm_workspace_elems(parts, window, dim) if parts else 0
    ws = workspace if workspace is not None and workspace.numel() >= ws_n else \
        torch.empty(max(ws_n, 1), dtype=torch.float32, device=dev)
    flags = _spmm_flags(rows_covered, pipe, xcd)
    rc = lib().attgnn_spmm_compact(_ptr(rp), _ptr(col), _ptr(masks), _ptr(block_base), _ptr(values),
                                   _TYPE_OF[values.dtype], _ptr(window_of_part), _ptr(atomic), parts, _ptr(X),
                                   _TYPE_OF[X.dtype], n_x, _ptr(y), _TYPE_OF[y.dtype], m_ori, dim, window,
                                   _mode(mode, X.dtype), _ptr(row_scale), _ptr(col_scale), _ptr(ws), ws_n, flags,
                                   _stream(dev))
    check(rc, "attgnn_spmm_compact")
    return y


def stripe_segments(rp, window_of_part, col, owners, m):
    """attgnn_stripe_segments_device: int32 [parts, owners + 1], the first
    position of each owner's segment in every part's real columns of a
    structure whose column ids are striped positions (graphs.stripe_pos) of an
    m-node graph.  Kept on ``rp`` (with the version counters of ``rp`` and
    ``window_of_part``) like :func:`canonical_tasks`."""
    dev = _need_gpu("stripe_segments", rp, window_of_part, col)
    _need_i32("stripe_segments", rp, window_of_part, col)
    parts = window_of_part.numel()
    try:
        version = (rp._version, window_of_part._version, col._version)
    except RuntimeError:
        version = None
    tag = (version, rp.numel(), window_of_part.data_ptr(), col.data_ptr(), int(owners), int(m))
    cached = getattr(rp, "_attgnn_oseg", None)
    if version is not None and cached is not None and cached[0] == tag:
        return cached[1]
    seg = torch.empty((parts, owners + 1), dtype=torch.int32, device=dev)
    check(lib().attgnn_stripe_segments_device(_ptr(rp), _ptr(window_of_part), _ptr(col), parts, int(owners), int(m),
                                              _ptr(seg), _stream(dev)),
          "attgnn_stripe_segments_device")
    if version is not None:
        rp._attgnn_oseg = (tag, seg)
    return seg


def _drop_check(name, t):
    if t.dtype not in (torch.float32, torch.float16, torch.bfloat16):
        raise TypeError(f"{name}: fp32 / fp16 / bf16 tensors only, got {t.dtype}")
    if not t.is_contiguous() or t.numel() % 8 or t.data_ptr() % 16:
        raise ValueError(f"{name}: needs a contiguous tensor of a multiple of 8 elements, 16-byte aligned")


def relu_dropout(x, p, seed):
    """attgnn_relu_dropout_fwd: (y, mask) with y = x / (1 - p) where x > 0
    and the element is kept, else 0; the keep decision of element i is a hash
    of (seed, i); mask holds one bit per element (uint8 [numel / 8])."""
    dev = _need_gpu("relu_dropout", x)
    _drop_check("relu_dropout", x)
    if not 0.0 <= p <= 1.0:
        raise ValueError(f"p must be in [0, 1], got {p}")
    y = torch.empty_like(x)
    mask = torch.empty(x.numel() // 8, dtype=torch.uint8, device=dev)
    check(lib().attgnn_relu_dropout_fwd(_ptr(x), _TYPE_OF[x.dtype], x.numel(), float(p),
                                        int(seed) & 0xFFFFFFFFFFFFFFFF, _ptr(y), _ptr(mask), _stream(dev)),
          "attgnn_relu_dropout_fwd")
    return y, mask


def relu_dropout_bwd(dy, p, mask):
    """attgnn_relu_dropout_bwd: dx = dy / (1 - p) where the mask bit is set, else 0."""
    dev = _need_gpu("relu_dropout_bwd", dy, mask)
    _drop_check("relu_dropout_bwd", dy)
    if mask.dtype != torch.uint8 or mask.numel() != dy.numel() // 8:
        raise ValueError("mask must be uint8 [numel / 8]")
    dx = torch.empty_like(dy)
    check(lib().attgnn_relu_dropout_bwd(_ptr(dy), _TYPE_OF[dy.dtype], dy.numel(), float(p), _ptr(mask), _ptr(dx),
                                        _stream(dev)), "attgnn_relu_dropout_bwd")
    return dx


def nll_log_softmax(logits, labels, classes, rows=None, loss=True, grad=True, grad_scale=None):
    """attgnn_nll_log_softmax: (loss, grad) for logits [n, ld] (fp32 / fp16 /
    bf16, the first ``classes`` columns being the classes), int64 labels [n]
    and the int64 row indices ``rows`` the mean runs over (None: every row).
    loss is the FP32 0-d mean of -log_softmax(logits[r, :classes])[labels[r]];
    grad its gradient w.r.t. logits times ``grad_scale`` (a 0-d fp32 device
    tensor, None: 1), in their shape and type, zero outside the selected rows
    and the class columns.  ``loss=False`` / ``grad=False`` skip one (None in
    its place).  Bad labels or row indices give NaN."""
    want_loss, want_grad = loss, grad
    if not (want_loss or want_grad):
        raise ValueError("nothing to compute")
    dev = _need_gpu("nll_log_softmax", logits, labels, *([rows] if rows is not None else []))
    if logits.dim() != 2 or not logits.is_contiguous():
        raise ValueError("logits must be a contiguous 2-D tensor")
    if logits.dtype not in (torch.float32, torch.float16, torch.bfloat16):
        raise TypeError(f"logits must be fp32 / fp16 / bf16, got {logits.dtype}")
    n, ld = logits.shape
    if labels.dtype != torch.int64 or labels.shape != (n,) or not labels.is_contiguous():
        raise ValueError(f"labels must be a contiguous int64 [{n}] tensor")
    if rows is not None and (rows.dtype != torch.int64 or rows.dim() != 1 or not rows.is_contiguous()):
        raise ValueError("rows must be a contiguous 1-D int64 tensor")
    if not 0 < classes <= ld:
        raise ValueError(f"classes must be in [1, {ld}], got {classes}")
    n_sel = n if rows is None else rows.numel()
    if n_sel == 0:
        raise ValueError("the mean needs at least one row")
    if grad_scale is not None and (grad_scale.dtype != torch.float32 or grad_scale.numel() != 1
                                   or grad_scale.device != dev):
        raise ValueError("grad_scale must be a one-element fp32 tensor on the logits' device")
    grad = torch.zeros_like(logits) if want_grad else None
    loss = torch.empty((), dtype=torch.float32, device=dev) if want_loss else None
    ws_n = lib().attgnn_nll_log_softmax_workspace_elems(n_sel) if want_loss else 0
    ws = torch.empty(ws_n, dtype=torch.float32, device=dev) if want_loss else None
    gs = grad_scale.contiguous() if grad_scale is not None else None
    check(lib().attgnn_nll_log_softmax(_ptr(logits), _TYPE_OF[logits.dtype], n, ld, int(classes), _ptr(labels),
                                       _ptr(rows), n_sel, _ptr(loss), _ptr(grad), _ptr(gs), _ptr(ws), ws_n,
                                       _stream(dev)),
          "attgnn_nll_log_softmax")
    return loss, grad


def spmm_compact_striped(rp, col, masks, block_base, values, window_of_part, atomic, owner_seg, owners, X, window,
                         mode=None, out_dtype=torch.float32, rows_covered=False, workspace=None, pipe=None,
                         split=True, groups=None):
    """attgnn_spmm_compact_striped: Y' = A X' in striped node order.

    The structure's windows run over the graph's rows in its own order and its
    column ids are striped positions (graphs.stripe_pos with ``owners``); X' is
    [m, D] with node j's features at row stripe_pos(j), and Y' [m, D] receives
    row j at stripe_pos(j).  Where the feature-slice count times ``owners`` is
    8 (FP16 / BF16 at D = 256 with 4 owner groups) the parts are split by owner so
    each XCD gathers one owner's rows of one slice; ``split=False`` keeps the
    usual task mapping (results then differ in rounding only).  ``groups``:
    the owner groups per part (at most min(owners, 8 / slices); default: the
    fewest that leave each XCD about 16 MB of X) -- each task walks the
    segments of owners/groups owners in turn."""
    dev = _need_gpu("spmm_compact_striped", rp, col, masks, block_base, values, window_of_part, atomic, owner_seg, X)
    _need_i32("spmm_compact_striped", rp, col, window_of_part, atomic, owner_seg)
    parts = window_of_part.numel()
    m, dim = X.shape
    if owner_seg.shape != (parts, owners + 1):
        raise ValueError(f"owner_seg must be [{parts}, {owners + 1}], got {tuple(owner_seg.shape)}")
    y = torch.empty((m, dim), dtype=out_dtype, device=dev)
    ws_n = lib().attgnn_spmm_striped_workspace_elems(parts, window, dim, owners) if parts else 0
    ws = workspace if workspace is not None and workspace.numel() >= ws_n else \
        torch.empty(max(ws_n, 1), dtype=torch.float32, device=dev)
    flags = _spmm_flags(rows_covered, pipe, split)
    if groups:
        if groups & (groups - 1) or not 1 <= groups <= 8:
            raise ValueError(f"groups must be 1, 2, 4 or 8, got {groups}")
        flags |= (groups.bit_length()) << 12
    rc = lib().attgnn_spmm_compact_striped(_ptr(rp), _ptr(col), _ptr(masks), _ptr(block_base), _ptr(values),
                                           _TYPE_OF[values.dtype], _ptr(window_of_part), _ptr(atomic), parts,
                                           _ptr(owner_seg), int(owners), _ptr(X), _TYPE_OF[X.dtype], m, _ptr(y),
                                           _TYPE_OF[y.dtype], m, dim, window, _mode(mode, X.dtype), _ptr(ws), ws_n,
                                           flags, _stream(dev))
    check(rc, "attgnn_spmm_compact_striped")
    return y


def spmm_edge_map(rp, col, emap, edge_val, window_of_part, atomic, X, out_layout, window, wide,
                  m_ori=None, mode=None, x_layout="hmd", out_dtype=torch.float32, rows_covered=False, pipe=None,
                  workspace=None, xcd=True):
    """Edge-map SpMM.  X [H, N, D] ("hmd") or [N, H, D] ("mhd"); edge_val [H, nnz]
    (or [nnz]); Y [H, M, D] ("hmd") or [M, H, D] ("mhd") in ``out_dtype`` (FP32
    accumulation).  ``rows_covered``: every window up to M has a part, so Y
    needs no zero-fill."""
    dev = _need_gpu("spmm_edge_map", rp, col, emap, edge_val, window_of_part, atomic, X)
    _need_i32("spmm_edge_map", rp, col, emap, window_of_part, atomic)
    return _edge_map(None, rp, col, emap, edge_val, window_of_part, atomic, X, out_layout, window, wide, m_ori, mode,
                     x_layout, out_dtype, rows_covered, pipe, workspace, xcd, dev)


def spmm_edge_map_compact(rp, col, masks, block_base, ids, edge_val, window_of_part, atomic, X, out_layout, window,
                          m_ori=None, mode=None, x_layout="hmd", out_dtype=torch.float32, rows_covered=False,
                          pipe=None, workspace=None, xcd=True):
    """:func:`spmm_edge_map` over the compact form of the map (wide = 32):
    ``masks``, ``block_base`` and the packed edge ``ids`` of a map built with
    ``compact=True``.  Bitwise equal to :func:`spmm_edge_map`."""
    dev = _need_gpu("spmm_edge_map_compact", rp, col, masks, block_base, ids, edge_val, window_of_part, atomic, X)
    _need_i32("spmm_edge_map_compact", rp, col, ids, window_of_part, atomic)
    if masks.dtype not in (torch.int16, torch.uint16) or block_base.dtype != torch.int64:
        raise TypeError("masks must be 16-bit and block_base int64")
    return _edge_map((masks, block_base), rp, col, ids, edge_val, window_of_part, atomic, X, out_layout, window, 32,
                     m_ori, mode, x_layout, out_dtype, rows_covered, pipe, workspace, xcd, dev)


def _edge_map(compact, rp, col, emap, edge_val, window_of_part, atomic, X, out_layout, window, wide, m_ori, mode,
              x_layout, out_dtype, rows_covered, pipe, workspace, xcd, dev, dot_x=None):
    if X.dim() == 2:
        X = X.unsqueeze(0 if x_layout == "hmd" else 1)
    if x_layout == "hmd":
        heads, n_x, dim = X.shape
    else:
        n_x, heads, dim = X.shape
    ev = edge_val.reshape(heads, -1) if edge_val.dim() != 2 else edge_val
    if ev.shape[0] != heads:
        raise ValueError(f"edge values have {ev.shape[0]} heads, X has {heads}")
    nnz = ev.shape[1]
    m = n_x if m_ori is None else m_ori
    layout = {"hmd": _lib.LAYOUT_HMD, "mhd": _lib.LAYOUT_MHD}[out_layout]
    shape = (heads, m, dim) if layout == _lib.LAYOUT_HMD else (m, heads, dim)
    if out_dtype not in (torch.float32, torch.float16, torch.bfloat16):
        raise TypeError(f"output type {out_dtype} is not supported")
    y = torch.empty(shape, dtype=out_dtype, device=dev)
    parts = window_of_part.numel()
    ws_n = heads * lib().attgnn_spmm_workspace_elems(parts, window, dim) if parts else 0
    ws = workspace if workspace is not None and workspace.numel() >= ws_n else \
        torch.empty(max(ws_n, 1), dtype=torch.float32, device=dev)
    if ev.dtype != X.dtype:
        ev = ev.to(X.dtype)
    ev = ev.contiguous()
    xl = {"hmd": _lib.LAYOUT_HMD, "mhd": _lib.LAYOUT_MHD}[x_layout]
    flags = _spmm_flags(rows_covered, pipe, xcd)
    tail = (_ptr(ev), _TYPE_OF[ev.dtype], nnz, _ptr(window_of_part), _ptr(atomic), parts, _ptr(X), _TYPE_OF[X.dtype],
            xl, heads, n_x, _ptr(y), _TYPE_OF[out_dtype], m, dim, layout, window)
    if dot_x is not None:
        if dot_x.dim() == 2:
            dot_x = dot_x.unsqueeze(0 if x_layout == "hmd" else 1)
        dh, dot_rows, dd = dot_x.shape if x_layout == "hmd" else (dot_x.shape[1], dot_x.shape[0], dot_x.shape[2])
        if (dh, dd) != (heads, dim) or dot_x.dtype != X.dtype or not dot_x.is_contiguous() or \
                dot_x.device != X.device:
            raise ValueError("dot_x must be a contiguous tensor of X's heads, width, type, layout and device")
        dot = torch.empty((heads, nnz), dtype=torch.float32, device=dev)
        rc = lib().attgnn_spmm_edge_map_compact_dot(_ptr(rp), _ptr(col), _ptr(compact[0]), _ptr(compact[1]),
                                                    _ptr(emap), *tail, _ptr(dot_x), dot_rows, _ptr(dot), _ptr(ws),
                                                    ws_n, flags, _stream(dev))
        check(rc, "attgnn_spmm_edge_map_compact_dot")
        return y, dot
    if compact is None:
        check_part_spans(rp, wide, "spmm_edge_map")
        rc = lib().attgnn_spmm_edge_map(_ptr(rp), _ptr(col), _ptr(emap), *tail, wide, _mode(mode, X.dtype), _ptr(ws),
                                        ws_n, flags, _stream(dev))
        check(rc, "attgnn_spmm_edge_map")
    else:
        rc = lib().attgnn_spmm_edge_map_compact(_ptr(rp), _ptr(col), _ptr(compact[0]), _ptr(compact[1]), _ptr(emap),
                                                *tail, _mode(mode, X.dtype), _ptr(ws), ws_n, flags, _stream(dev))
        check(rc, "attgnn_spmm_edge_map_compact")
    return y


def spmm_edge_map_compact_dot(rp, col, masks, block_base, ids, edge_val, window_of_part, atomic, X, dot_x,
                              out_layout, window, m_ori=None, x_layout="mhd", out_dtype=torch.float32,
                              rows_covered=False, workspace=None):
    """:func:`spmm_edge_map_compact` (native mode) that also returns, for every
    edge e at window row r and column c of the structure, dot[h, e] =
    <dot_x[h, r, :], X[h, c, :]> in fp32 (dot_x: X's heads, width, type and
    layout, any row count -- rows past it read as zero; fp16 / bf16, 32, 64 or
    128 features).  Over A^T's structure with A's edge
    ids and X = dL/dh, dot_x = Xt this is the GAT backward's aggregation
    gradient and the SDDMM <dh[i], Xt[j]> in one pass.  Returns (Y, dot);
    bitwise equal to the two separate products."""
    dev = _need_gpu("spmm_edge_map_compact_dot", rp, col, masks, block_base, ids, edge_val, window_of_part, atomic,
                    X, dot_x)
    _need_i32("spmm_edge_map_compact_dot", rp, col, ids, window_of_part, atomic)
    if masks.dtype not in (torch.int16, torch.uint16) or block_base.dtype != torch.int64:
        raise TypeError("masks must be 16-bit and block_base int64")
    if X.dtype not in (torch.float16, torch.bfloat16):
        raise TypeError("spmm_edge_map_compact_dot: X must be fp16 or bf16")
    return _edge_map((masks, block_base), rp, col, ids, edge_val, window_of_part, atomic, X, out_layout, window, 32,
                     m_ori, None, x_layout, out_dtype, rows_covered, None, workspace, True, dev, dot_x=dot_x)


def _node_terms(t, heads, layout, name):
    """(pointer, head stride, node stride, nodes) of fp32 node terms, [N, H] ("nh") or [H, N] ("hn")."""
    if t is None:
        return None, 0, 0, None
    if t.dtype != torch.float32 or t.dim() != 2 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous fp32 2-D tensor")
    if layout == "nh" and t.shape[1] == heads:
        return _ptr(t), 1, heads, t.shape[0]
    if layout == "hn" and t.shape[0] == heads:
        return _ptr(t), t.shape[1], 1, t.shape[1]
    raise ValueError(f"{name}: shape {tuple(t.shape)} is not {layout!r} with {heads} heads")


def spmm_attention(rp, col, emap, window_of_part, atomic, X, out_layout, window, wide, row_logit, col_logit,
                   row_offset=None, col_offset=None, slope=0.2, drop_p=0.0, seed=0, m_ori=None, mode=None,
                   x_layout="hmd", out_dtype=torch.float32, rows_covered=False, pipe=None, logit_layout="nh",
                   offset_layout="hn", xcd=True):
    """Attention aggregation with the probabilities rebuilt in the kernel:
    entry (r, c) = exp(LeakyReLU(row_logit[r] + col_logit[c]) - row_offset[r] - col_offset[c]) with the edge
    softmax's dropout; equal bit for bit to :func:`spmm_edge_map` over the p that
    :func:`edge_softmax_fwd` writes.  Logit terms are fp32 [N, H] (``logit_layout="nh"``, the node scores'
    node-major form) or [H, N]; offsets (the softmax's log-sum-exp) [H, N] (``offset_layout="hn"``) or
    [N, H].  X / Y layouts and types as :func:`spmm_edge_map`."""
    dev = _need_gpu("spmm_attention", rp, col, emap, window_of_part, atomic, X, row_logit, col_logit, row_offset,
                    col_offset)
    _need_i32("spmm_attention", rp, col, emap, window_of_part, atomic)
    check_part_spans(rp, wide, "spmm_attention")
    if X.dim() == 2:
        X = X.unsqueeze(0 if x_layout == "hmd" else 1)
    if x_layout == "hmd":
        heads, n_x, dim = X.shape
    else:
        n_x, heads, dim = X.shape
    r1, lhs, lns, n1 = _node_terms(row_logit, heads, logit_layout, "row_logit")
    c1, _, _, n2 = _node_terms(col_logit, heads, logit_layout, "col_logit")
    r2, ohs, ons, n3 = _node_terms(row_offset, heads, offset_layout, "row_offset")
    c2, ohs2, ons2, n4 = _node_terms(col_offset, heads, offset_layout, "col_offset")
    if r1 is None or c1 is None:
        raise ValueError("row_logit and col_logit are required")
    if r2 is None:
        ohs, ons, n3 = ohs2, ons2, n4
    if n1 != n2 or any(n is not None and n != n1 for n in (n3, n4)):
        raise ValueError("the node terms cover different node counts")
    m = n_x if m_ori is None else m_ori
    layout = {"hmd": _lib.LAYOUT_HMD, "mhd": _lib.LAYOUT_MHD}[out_layout]
    shape = (heads, m, dim) if layout == _lib.LAYOUT_HMD else (m, heads, dim)
    if out_dtype not in (torch.float32, torch.float16, torch.bfloat16):
        raise TypeError(f"output type {out_dtype} is not supported")
    y = torch.empty(shape, dtype=out_dtype, device=dev)
    parts = window_of_part.numel()
    ws_n = heads * lib().attgnn_spmm_workspace_elems(parts, window, dim) if parts else 0
    ws = torch.empty(max(ws_n, 1), dtype=torch.float32, device=dev)
    terms = _lib.AttnTerms(r1, c1, r2, c2, lhs, lns, ohs, ons, n1, float(slope), float(drop_p),
                           int(seed) & (2**64 - 1))
    xl = {"hmd": _lib.LAYOUT_HMD, "mhd": _lib.LAYOUT_MHD}[x_layout]
    rc = lib().attgnn_spmm_attention(_ptr(rp), _ptr(col), _ptr(emap), _ptr(window_of_part), _ptr(atomic), parts,
                                     ctypes.byref(terms), _ptr(X), _TYPE_OF[X.dtype], xl, heads, n_x, _ptr(y),
                                     _TYPE_OF[out_dtype], m, dim, layout, window, wide, _mode(mode, X.dtype),
                                     _ptr(ws), ws_n, _spmm_flags(rows_covered, pipe, xcd),
                                     _stream(dev))
    check(rc, "attgnn_spmm_attention")
    return y


def rowsum_edge_map(rp, emap, window_of_part, edge_val, m_rows, out_layout, window, wide):
    """Row sums of per-edge values; fp32 [H, M] ("hmd") or [M, H] ("mhd")."""
    dev = _need_gpu("rowsum_edge_map", rp, emap, window_of_part, edge_val)
    _need_i32("rowsum_edge_map", rp, emap, window_of_part)
    check_part_spans(rp, wide, "rowsum_edge_map")
    ev = edge_val if edge_val.dim() == 2 else edge_val.reshape(1, -1)
    heads, nnz = ev.shape
    layout = {"hmd": _lib.LAYOUT_HMD, "mhd": _lib.LAYOUT_MHD}[out_layout]
    shape = (heads, m_rows) if layout == _lib.LAYOUT_HMD else (m_rows, heads)
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    rc = lib().attgnn_rowsum_edge_map(_ptr(rp), _ptr(emap), _ptr(window_of_part), window_of_part.numel(),
                                      _ptr(ev), _TYPE_OF[ev.dtype], nnz, heads, _ptr(out), layout, m_rows,
                                      window, wide, _stream(dev))
    check(rc, "attgnn_rowsum_edge_map")
    return out


# --------------------------------------------------------------------------
# SDDMM and per-edge attention terms
# --------------------------------------------------------------------------

def _hmd_strides(X, layout):
    """(heads, rows, dim, head_stride, row_stride) of a 2-D or 3-D tensor."""
    if X.dim() == 2:
        return 1, X.shape[0], X.shape[1], 0, X.shape[1]
    if layout == "hmd":
        h, n, d = X.shape
        return h, n, d, n * d, d
    n, h, d = X.shape
    return h, n, d, d, h * d


def sddmm_block(rp, col, emap, window_of_part, A, B, nnz, window, wide, layout="hmd",
                out_dtype=None, mode=None):
    """out[h, e=(i,j)] = <A_h[i], B_h[j]> in CSR edge order ([H, nnz])."""
    dev = _need_gpu("sddmm_block", rp, col, emap, window_of_part, A, B)
    _need_i32("sddmm_block", rp, col, emap, window_of_part)
    check_part_spans(rp, wide, "sddmm_block")
    if A.dtype != B.dtype:
        raise TypeError("A and B must share a dtype")
    ha, na, da, ahs, ars = _hmd_strides(A, layout)
    hb, nb, db, bhs, brs = _hmd_strides(B, layout)
    if ha != hb or da != db:
        raise ValueError("A and B disagree in heads or dim")
    out_dtype = out_dtype or A.dtype
    out = torch.zeros((ha, nnz), dtype=out_dtype, device=dev)
    rc = lib().attgnn_sddmm_block(_ptr(rp), _ptr(col), _ptr(emap), _ptr(window_of_part),
                                  window_of_part.numel(), _ptr(A), ahs, ars, _ptr(B), bhs, brs,
                                  _TYPE_OF[A.dtype], ha, na, da, _ptr(out), _TYPE_OF[out_dtype], nnz, window, wide,
                                  _mode(mode, A.dtype), _stream(dev))
    check(rc, "attgnn_sddmm_block")
    return out


def sddmm_edges(rp, col, block_base, edge_ids, positions, window_of_part, A, B, nnz, window, layout="hmd",
                out_dtype=None, mode=None, covered=False):
    """sddmm_block over a compact edge list (wide = 32); [H, nnz].  ``covered``:
    every edge lies in a window of the structure, so the output needs no zero-fill."""
    dev = _need_gpu("sddmm_edges", rp, col, block_base, edge_ids, positions, window_of_part, A, B)
    _need_i32("sddmm_edges", rp, col, edge_ids, window_of_part)
    if A.dtype != B.dtype:
        raise TypeError("A and B must share a dtype")
    ha, na, da, ahs, ars = _hmd_strides(A, layout)
    hb, nb, db, bhs, brs = _hmd_strides(B, layout)
    if ha != hb or da != db:
        raise ValueError("A and B disagree in heads or dim")
    out_dtype = out_dtype or A.dtype
    out = (torch.empty if covered else torch.zeros)((ha, nnz), dtype=out_dtype, device=dev)
    rc = lib().attgnn_sddmm_edges(_ptr(rp), _ptr(col), _ptr(block_base), _ptr(edge_ids), _ptr(positions),
                                  _ptr(window_of_part), window_of_part.numel(), _ptr(A), ahs, ars, _ptr(B), bhs, brs,
                                  _TYPE_OF[A.dtype], ha, na, da, _ptr(out), _TYPE_OF[out_dtype], nnz, window,
                                  _mode(mode, A.dtype), _stream(dev))
    check(rc, "attgnn_sddmm_edges")
    return out


def node_scores(X, vec, layout="hmd", out_layout="hn"):
    """s[h, i] = X_h[i] . vec_h in fp32: [H, N] ("hn") or node-major [N, H]
    ("nh", the layout the edge softmax reads with one load per neighbour)."""
    dev = _need_gpu("node_scores", X, vec)
    h, n, d, hs, rs = _hmd_strides(X, layout)
    if vec.numel() != h * d:
        raise ValueError(f"vector has {vec.numel()} elements, expected {h}*{d}")
    if out_layout == "hn":
        s = torch.empty((h, n), dtype=torch.float32, device=dev)
        shs, srs = n, 1
    elif out_layout == "nh":
        s = torch.empty((n, h), dtype=torch.float32, device=dev)
        shs, srs = 1, h
    else:
        raise ValueError(f"out_layout must be 'hn' or 'nh', got {out_layout!r}")
    rc = lib().attgnn_node_scores(_ptr(X), _TYPE_OF[X.dtype], hs, rs, _ptr(vec), _TYPE_OF[vec.dtype], h, n, d,
                                  _ptr(s), shs, srs, _stream(dev))
    check(rc, "attgnn_node_scores")
    return s


def node_scores_pair(X, vec0, vec1, layout="hmd", out_layout="hn"):
    """(node_scores(X, vec0), node_scores(X, vec1)) from one pass over X,
    each bitwise equal to its single call."""
    dev = _need_gpu("node_scores_pair", X, vec0, vec1)
    h, n, d, hs, rs = _hmd_strides(X, layout)
    if vec0.numel() != h * d or vec1.numel() != h * d or vec0.dtype != vec1.dtype:
        raise ValueError(f"vectors must have {h}*{d} elements of one dtype")
    if out_layout == "hn":
        shape, shs, srs = (h, n), n, 1
    elif out_layout == "nh":
        shape, shs, srs = (n, h), 1, h
    else:
        raise ValueError(f"out_layout must be 'hn' or 'nh', got {out_layout!r}")
    s0 = torch.empty(shape, dtype=torch.float32, device=dev)
    s1 = torch.empty(shape, dtype=torch.float32, device=dev)
    rc = lib().attgnn_node_scores_pair(_ptr(X), _TYPE_OF[X.dtype], hs, rs, _ptr(vec0), _ptr(vec1),
                                       _TYPE_OF[vec0.dtype], h, n, d, _ptr(s0), _ptr(s1), shs, srs, _stream(dev))
    check(rc, "attgnn_node_scores_pair")
    return s0, s1


def attention_node_bwd(Xs, Xt, a0, a1, ds0, ds1, base=None, combine=False):
    """Fused backward of s0 = Xs.a0, s1 = Xt.a1 plus the aggregation gradient.

    Xs, Xt [N, H, D] (same dtype), a0 / a1 [H, D], ds0 / ds1 fp32 [H, N],
    base fp32 [N, H, D] or None.  Returns (dXs, dXt, da0, da1): dXs = ds0 a0,
    dXt = base + ds1 a1 (in X's dtype), da0 / da1 fp32 [H, D].  combine=True
    (Xs is Xt): dXs is None and dXt holds the sum of both gradients.
    """
    dev = _need_gpu("attention_node_bwd", Xs, Xt, ds0, ds1, base)
    if Xs.dtype != Xt.dtype or Xs.shape != Xt.shape or Xs.dim() != 3:
        raise ValueError("Xs and Xt must be [N, H, D] tensors of one dtype")
    if combine and Xs.data_ptr() != Xt.data_ptr():
        raise ValueError("combine=True needs Xs and Xt to be the same tensor")
    n, heads, dim = Xs.shape
    for t, nm in ((ds0, "ds0"), (ds1, "ds1")):
        if t.dtype != torch.float32 or tuple(t.shape) != (heads, n):
            raise ValueError(f"{nm} must be fp32 [H, N]")
    if base is not None and (base.dtype != torch.float32 or tuple(base.shape) != (n, heads, dim)):
        raise ValueError("base must be fp32 [N, H, D]")
    a0f = a0.reshape(heads, dim).float().contiguous()
    a1f = a1.reshape(heads, dim).float().contiguous()
    dXt = torch.empty_like(Xt)
    dXs = None if combine else torch.empty_like(Xs)
    da0 = torch.empty((heads, dim), dtype=torch.float32, device=dev)
    da1 = torch.empty((heads, dim), dtype=torch.float32, device=dev)
    L = lib()
    ws_n = L.attgnn_attention_node_bwd_workspace_elems(n, heads, dim)
    ws = torch.empty(max(ws_n, 1), dtype=torch.float32, device=dev)
    check(L.attgnn_attention_node_bwd(_ptr(Xs), _ptr(Xt), _TYPE_OF[Xs.dtype], n, heads, dim, _ptr(a0f), _ptr(a1f),
                                      _ptr(ds0), _ptr(ds1), _ptr(base), _ptr(dXs), _ptr(dXt), _ptr(da0), _ptr(da1),
                                      _ptr(ws), ws_n, _stream(dev)),
          "attgnn_attention_node_bwd")
    return dXs, dXt, da0, da1


def edge_from_nodes(rp, col, emap, window_of_part, nnz, heads, n_rows, window, wide, out_dtype,
                    s0=None, s1=None, g0=None, g0_strides=None):
    """out[h, e=(i,j)] = S0[h, i] + S1[h, j] over the map's edges ([H, nnz])."""
    dev = _need_gpu("edge_from_nodes", rp, col, emap, window_of_part, s0, s1, g0)
    _need_i32("edge_from_nodes", rp, col, emap, window_of_part)
    check_part_spans(rp, wide, "edge_from_nodes")
    out = torch.zeros((heads, nnz), dtype=out_dtype, device=dev)
    if g0 is not None:
        hs0, rs0 = g0_strides
        g0t = _TYPE_OF[g0.dtype]
    else:
        hs0, rs0 = (n_rows, 1)
        g0t = 0
    rc = lib().attgnn_edge_from_nodes(_ptr(rp), _ptr(col), _ptr(emap), _ptr(window_of_part),
                                      window_of_part.numel(), _ptr(s0), hs0, rs0, _ptr(s1),
                                      (s1.shape[1] if s1 is not None else 0), 1, _ptr(g0), g0t, heads, n_rows,
                                      _ptr(out), _TYPE_OF[out_dtype], nnz, window, wide, _stream(dev))
    check(rc, "attgnn_edge_from_nodes")
    return out


def dtype_code(dtype):
    return _TYPE_OF[dtype]


# --------------------------------------------------------------------------
# fused edge softmax (CSR rows)
# --------------------------------------------------------------------------

def edge_score_range(rowptr, col, s0, s1, s_layout="nh", long_rows=None):
    """Per-head range of the logits s0[i] + s1[j] over the CSR's edges:
    (range fp32 [H, 3] = min, max, 1/(max - min) or 0; where int32 [H, 4] =
    row, column of the minimum's edge, then of the maximum's; -1 without
    edges).  s0 / s1 fp32 node terms, [N, H] ("nh") or [H, N] ("hn");
    long_rows: the graph's long-row list (``long_rows(rowptr)``), whose rows
    then get a workgroup each."""
    dev = _need_gpu("edge_score_range", rowptr, col, s0, s1, long_rows)
    _need_i32("edge_score_range", rowptr, col, *([] if long_rows is None else [long_rows]))
    if s0.shape != s1.shape or s0.dtype != torch.float32 or s1.dtype != torch.float32 or s0.dim() != 2:
        raise ValueError("s0 and s1 must be fp32 [N, H] / [H, N] tensors of one shape")
    heads = s0.shape[1] if s_layout == "nh" else s0.shape[0]
    n_nodes = s0.shape[0] if s_layout == "nh" else s0.shape[1]
    n = rowptr.numel() - 1
    if n > n_nodes:
        raise ValueError(f"{n} CSR rows but node terms for {n_nodes} nodes")
    hs, rs = (1, heads) if s_layout == "nh" else (n_nodes, 1)
    ws_n = lib().attgnn_edge_score_range_workspace_elems(heads)
    ws = torch.empty(ws_n, dtype=torch.int64, device=dev)
    rng = torch.empty((heads, 3), dtype=torch.float32, device=dev)
    where = torch.empty((heads, 4), dtype=torch.int32, device=dev)
    check(lib().attgnn_edge_score_range(_ptr(rowptr), _ptr(col), n, col.numel(), heads, _ptr(s0), _ptr(s1), hs, rs,
                                        _ptr(ws), ws_n, _ptr(rng), _ptr(where), _ptr(long_rows),
                                        0 if long_rows is None else long_rows.numel(), _stream(dev)),
          "attgnn_edge_score_range")
    return rng, where


def minmax_score_grads(s0, s1, rng, where, ds0s, ds1s, s_layout="nh"):
    """Gradients of the raw node terms s0 / s1 (fp32, [N, H] "nh" or [H, N]
    "hn") under the min-max logit map, from ds0s / ds1s (fp32 [H, N]) of the
    scaled terms; rng / where from edge_score_range.  Returns (ds0, ds1)
    fp32 [H, N]."""
    dev = _need_gpu("minmax_score_grads", s0, s1, rng, where, ds0s, ds1s)
    heads, n = ds0s.shape
    if ds1s.shape != ds0s.shape or ds0s.dtype != torch.float32 or ds1s.dtype != torch.float32:
        raise ValueError("ds0s / ds1s must be fp32 [H, N] of one shape")
    want = (n, heads) if s_layout == "nh" else (heads, n)
    if tuple(s0.shape) != want or tuple(s1.shape) != want or s0.dtype != torch.float32 or s1.dtype != torch.float32:
        raise ValueError(f"s0 / s1 must be fp32 {want}")
    if tuple(rng.shape) != (heads, 3) or tuple(where.shape) != (heads, 4) or where.dtype != torch.int32:
        raise ValueError("rng [H, 3] fp32 and where [H, 4] int32 expected")
    s0, s1, rng, where = s0.contiguous(), s1.contiguous(), rng.contiguous(), where.contiguous()
    ds0s, ds1s = ds0s.contiguous(), ds1s.contiguous()
    hs, rs = (1, heads) if s_layout == "nh" else (n, 1)
    ds0, ds1 = torch.empty_like(ds0s), torch.empty_like(ds1s)
    ws_n = lib().attgnn_minmax_grads_workspace_elems(heads)
    ws = torch.empty(ws_n, dtype=torch.float32, device=dev)
    check(lib().attgnn_minmax_score_grads(_ptr(s0), _ptr(s1), n, heads, hs, rs, _ptr(rng), _ptr(where), _ptr(ds0s),
                                          _ptr(ds1s), _ptr(ds0), _ptr(ds1), _ptr(ws), ws_n, _stream(dev)),
          "attgnn_minmax_score_grads")
    return ds0, ds1


def long_rows(rowptr):
    """int32 device tensor of the rows with more than ATTGNN_LONG_ROW (256)
    entries, the list the row kernels take so a workgroup handles each of them
    (one host copy of the row pointers)."""
    rp = torch.as_tensor(rowptr).to(torch.int32).contiguous().cpu()
    n = rp.numel() - 1
    if n < 0:
        raise ValueError("rowptr must have at least one element")
    k = lib().attgnn_long_rows(_ptr(rp), n, None)
    out = torch.empty(max(k, 0), dtype=torch.int32)
    if k > 0:
        lib().attgnn_long_rows(_ptr(rp), n, _ptr(out))
    dev = rowptr.device if isinstance(rowptr, torch.Tensor) else torch.device("cpu")
    return out.to(dev)


def _rows_arg(name, rows, dev):
    """(pointer, count) of an optional long-row list."""
    if rows is None or rows.numel() == 0:
        return None, 0
    if rows.dtype != torch.int32 or rows.device != dev or not rows.is_contiguous():
        raise ValueError(f"{name} must be a contiguous int32 tensor on {dev}")
    return _ptr(rows), rows.numel()


def _score_layout(s0, s1, s_layout):
    """(heads, n, head_stride, row_stride) of [H, N] ("hn") or [N, H] ("nh") scores."""
    if s0.shape != s1.shape or s0.dim() != 2 or s0.dtype != torch.float32 or s1.dtype != torch.float32:
        raise ValueError("s0 and s1 must be fp32 tensors of one 2-D shape")
    if not (s0.is_contiguous() and s1.is_contiguous()):
        raise ValueError("s0 and s1 must be contiguous")
    if s_layout == "hn":
        heads, n = s0.shape
        return heads, n, n, 1
    if s_layout == "nh":
        n, heads = s0.shape
        return heads, n, 1, heads
    raise ValueError(f"s_layout must be 'hn' or 'nh', got {s_layout!r}")


def edge_softmax_fwd(rowptr, col, s0, s1, slope, drop_p=0.0, seed=0, p_dtype=torch.float32, s_layout="hn",
                     long_rows=None, lse_only=False):
    """p [H, nnz] and lse [H, N] of the fused LeakyReLU + softmax + dropout;
    s0 / s1 [H, N] ("hn") or [N, H] ("nh").  ``long_rows``: the list from
    :func:`long_rows` (rows above 256 entries get a workgroup each).
    ``lse_only``: returns (None, lse) without writing p (for :func:`spmm_attention`)."""
    dev = _need_gpu("edge_softmax_fwd", rowptr, col, s0, s1)
    _need_i32("edge_softmax_fwd", rowptr, col)
    heads, n, shs, srs = _score_layout(s0, s1, s_layout)
    lp, ln = _rows_arg("long_rows", long_rows, dev)
    nnz = col.numel()
    p = None if lse_only else torch.empty((heads, nnz), dtype=p_dtype, device=dev)
    lse = torch.empty((heads, n), dtype=torch.float32, device=dev)
    rc = lib().attgnn_edge_softmax_fwd(_ptr(rowptr), _ptr(col), n, nnz, heads, _ptr(s0), _ptr(s1), shs, srs,
                                       float(slope), float(drop_p), int(seed) & (2**64 - 1), _ptr(p),
                                       _TYPE_OF.get(p_dtype, _TYPE_OF[torch.float32]), _ptr(lse), lp, ln,
                                       _stream(dev))
    check(rc, "attgnn_edge_softmax_fwd")
    return p, lse


def edge_softmax_bwd(rowptr, col, s0, s1, lse, dp, slope, drop_p=0.0, seed=0, s_layout="hn", long_rows=None):
    """dz [H, nnz] (fp32) and ds0 [H, N] (row sums of dz)."""
    dev = _need_gpu("edge_softmax_bwd", rowptr, col, s0, s1, lse, dp)
    _need_i32("edge_softmax_bwd", rowptr, col)
    heads, n, shs, srs = _score_layout(s0, s1, s_layout)
    lp, ln = _rows_arg("long_rows", long_rows, dev)
    nnz = col.numel()
    dz = torch.empty((heads, nnz), dtype=torch.float32, device=dev)
    ds0 = torch.empty((heads, n), dtype=torch.float32, device=dev)
    rc = lib().attgnn_edge_softmax_bwd(_ptr(rowptr), _ptr(col), n, nnz, heads, _ptr(s0), _ptr(s1), shs, srs,
                                       _ptr(lse), _ptr(dp), _TYPE_OF[dp.dtype], float(slope), float(drop_p),
                                       int(seed) & (2**64 - 1), _ptr(dz), _ptr(ds0), lp, ln, _stream(dev))
    check(rc, "attgnn_edge_softmax_bwd")
    return dz, ds0


def edge_softmax_bwd_sums(rowptr, col, s0, s1, lse, dp, slope, drop_p, seed, t_rowptr, t_perm, s_layout="hn",
                          long_rows=None, t_long_rows=None):
    """(ds0, ds1) fp32 [H, N]: the row and column sums of the softmax backward's
    dz, without returning dz (bitwise equal to edge_softmax_bwd followed by
    gather_rowsum over the transpose)."""
    dev = _need_gpu("edge_softmax_bwd_sums", rowptr, col, s0, s1, lse, dp, t_rowptr, t_perm)
    _need_i32("edge_softmax_bwd_sums", rowptr, col, t_rowptr)
    if t_perm.dtype != torch.int64:
        raise TypeError("t_perm must be int64")
    heads, n, shs, srs = _score_layout(s0, s1, s_layout)
    nnz = col.numel()
    if rowptr.numel() != n + 1 or t_rowptr.numel() != n + 1 or t_perm.numel() != nnz:
        raise ValueError("rowptr / t_rowptr / t_perm sizes disagree with the scores and col")
    lp, ln = _rows_arg("long_rows", long_rows, dev)
    tp, tn = _rows_arg("t_long_rows", t_long_rows, dev)
    ds0 = torch.empty((heads, n), dtype=torch.float32, device=dev)
    ds1 = torch.empty((heads, n), dtype=torch.float32, device=dev)
    L = lib()
    ws_n = L.attgnn_edge_softmax_bwd_sums_workspace_elems(nnz, heads)
    ws = torch.empty(max(ws_n, 1), dtype=torch.float32, device=dev)
    check(L.attgnn_edge_softmax_bwd_sums(_ptr(rowptr), _ptr(col), n, nnz, heads, _ptr(s0), _ptr(s1), shs, srs,
                                         _ptr(lse), _ptr(dp), _TYPE_OF[dp.dtype], float(slope), float(drop_p),
                                         int(seed) & (2**64 - 1), _ptr(t_rowptr), _ptr(t_perm), lp, ln, tp, tn,
                                         _ptr(ds0), _ptr(ds1), _ptr(ws), ws_n, _stream(dev)),
          "attgnn_edge_softmax_bwd_sums")
    return ds0, ds1


class DeviceSegments:
    """Device copy of :func:`att_amd.graphs.csr_segments` plus the C struct."""

    def __init__(self, seg, device):
        self.arrays = {k: torch.from_numpy(seg[k]).to(device) for k in ("seg_ptr", "seg_row", "seg_slot", "row_seg")}
        self.n_seg, self.n_slots = seg["n_seg"], seg["n_slots"]
        self.rows = self.arrays["row_seg"].numel() - 1
        self.struct = _lib.Segments(*[_ptr(self.arrays[k]).value for k in ("seg_ptr", "seg_row", "seg_slot",
                                                                             "row_seg")],
                                    self.n_seg, self.n_slots)


def gatv2_fwd(rowptr, col, segs, X0, X1, att, slope, drop_p=0.0, seed=0, p_dtype=None):
    """GATv2 logits z [H, nnz] (fp32), row log-sum-exp [H, N] and probabilities
    p [H, nnz] (dropout applied) from X0, X1 [N, H, D] and att [H, D]; `segs`:
    :class:`DeviceSegments` of the rows."""
    dev = _need_gpu("gatv2_fwd", rowptr, col, X0, X1)
    _need_i32("gatv2_fwd", rowptr, col)
    if X0.shape != X1.shape or X0.dtype != X1.dtype or X0.dim() != 3:
        raise ValueError("X0 and X1 must be [N, H, D] of one dtype")
    n, heads, dim = X0.shape
    nnz = col.numel()
    if rowptr.numel() != n + 1 or segs.rows != n:
        raise ValueError("rowptr / segments must cover the N rows")
    a = att.reshape(heads, dim).float().contiguous()
    p_dtype = p_dtype or X0.dtype
    z = torch.empty((heads, nnz), dtype=torch.float32, device=dev)
    lse = torch.empty((heads, n), dtype=torch.float32, device=dev)
    p = torch.empty((heads, nnz), dtype=p_dtype, device=dev)
    check(lib().attgnn_gatv2_fwd(_ptr(rowptr), _ptr(col), ctypes.byref(segs.struct), n, nnz, heads, dim, _ptr(X0),
                                 _ptr(X1), _TYPE_OF[X0.dtype], _ptr(a), float(slope), float(drop_p),
                                 int(seed) & (2**64 - 1), _ptr(z), _ptr(lse), _ptr(p), _TYPE_OF[p_dtype],
                                 _stream(dev)), "attgnn_gatv2_fwd")
    return z, lse, p


def gatv2_bwd(rowptr, col, segs, t_col, t_perm, t_segs, X0, X1, att, slope, z, lse, dp, drop_p=0.0, seed=0,
              base=None):
    """Backward of :func:`gatv2_fwd`: (dz [H, nnz] fp32, dX0, dX1 in X's dtype,
    datt [H, D] fp32); dX1 includes ``base`` (fp32 [N, H, D]) when given.
    ``t_col`` / ``t_perm`` / ``t_segs``: the transposed CSR's entries (row of A,
    A edge id) and its segments."""
    dev = _need_gpu("gatv2_bwd", rowptr, col, t_col, t_perm, X0, X1, z, lse, dp, base)
    _need_i32("gatv2_bwd", rowptr, col, t_col)
    if t_perm.dtype != torch.int64:
        raise TypeError("t_perm must be int64")
    n, heads, dim = X0.shape
    nnz = col.numel()
    if t_col.numel() != nnz or t_perm.numel() != nnz or segs.rows != n or t_segs.rows != n:
        raise ValueError("transposed CSR / segments do not match A")
    if base is not None and (base.dtype != torch.float32 or tuple(base.shape) != (n, heads, dim)):
        raise ValueError("base must be fp32 [N, H, D]")
    a = att.reshape(heads, dim).float().contiguous()
    dz = torch.empty((heads, nnz), dtype=torch.float32, device=dev)
    dX0 = torch.empty_like(X0)
    dX1 = torch.empty_like(X1)
    datt = torch.empty((heads, dim), dtype=torch.float32, device=dev)
    L = lib()
    ws_n = L.attgnn_gatv2_bwd_workspace_elems(heads, dim, ctypes.byref(segs.struct), ctypes.byref(t_segs.struct))
    ws = torch.empty(max(ws_n, 1), dtype=torch.float32, device=dev)
    check(L.attgnn_gatv2_bwd(_ptr(rowptr), _ptr(col), ctypes.byref(segs.struct), _ptr(t_col), _ptr(t_perm),
                             ctypes.byref(t_segs.struct), n, nnz, heads, dim, _ptr(X0), _ptr(X1), _TYPE_OF[X0.dtype],
                             _ptr(a), float(slope), float(drop_p), int(seed) & (2**64 - 1), _ptr(z), _ptr(lse),
                             _ptr(dp), _TYPE_OF[dp.dtype], _ptr(base), _ptr(dz), _ptr(dX0), _ptr(dX1), _ptr(datt),
                             _ptr(ws), ws_n, _stream(dev)), "attgnn_gatv2_bwd")
    return dz, dX0, dX1, datt


def gather_rowsum(rowptr, perm, vals, n, long_rows=None):
    """out[h, j] = sum_{e in row j} vals[h, perm[e]] (column sums via the transpose permutation)."""
    dev = _need_gpu("gather_rowsum", rowptr, perm, vals)
    heads, nnz = vals.shape
    if perm.dtype != torch.int64:
        raise TypeError("perm must be int64")
    lp, ln = _rows_arg("long_rows", long_rows, dev)
    out = torch.empty((heads, n), dtype=torch.float32, device=dev)
    rc = lib().attgnn_gather_rowsum(_ptr(rowptr), _ptr(perm), n, nnz, heads, _ptr(vals), _ptr(out), lp, ln,
                                    _stream(dev))
    check(rc, "attgnn_gather_rowsum")
    return out
