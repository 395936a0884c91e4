"""TEST INFRASTRUCTURE ONLY -- ctypes front end of the C preprocessing oracle.

``block_balance`` returns the five (or six) arrays the reference's
preprocessing returns, computed by ``block_oracle.c`` (a sequential C
restatement of gcn/Block_gcn/example.cpp:48-348).  ``block_balance_py`` is an
independent pure-Python restatement of the same lines used to cross-check the
C code on small inputs.  ``load_reference`` imports the reference's own
example.cpp when it has been compiled under ``oracle/_ref`` (``make -C oracle
ref``); it is absent on machines without the reference tree.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libblockoracle.so")
REF_DIR = os.path.join(HERE, "_ref")

FP16, FP32, MAP = 0, 1, 2


class _Result(ctypes.Structure):
    _fields_ = [
        ("rp", ctypes.POINTER(ctypes.c_int32)), ("n_rp", ctypes.c_int64),
        ("col", ctypes.POINTER(ctypes.c_int32)), ("n_col", ctypes.c_int64),
        ("val", ctypes.c_void_p), ("n_val", ctypes.c_int64),
        ("map_t", ctypes.POINTER(ctypes.c_int32)),
        ("win", ctypes.POINTER(ctypes.c_int32)), ("atomic", ctypes.POINTER(ctypes.c_int32)),
        ("n_parts", ctypes.c_int64),
    ]


_lib = None


def build():
    """Compile the C oracle (gcc via oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.bo_block_balance.restype = ctypes.c_int
        L.bo_block_balance.argtypes = [
            ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.POINTER(_Result)]
        L.bo_free.argtypes = [ctypes.POINTER(_Result)]
        L.bo_float_to_half.restype = ctypes.c_uint16
        L.bo_float_to_half.argtypes = [ctypes.c_float]
        L.bo_half_to_float.restype = ctypes.c_float
        L.bo_half_to_float.argtypes = [ctypes.c_uint16]
        _lib = L
    return _lib


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    buf = ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint8 * (n * np.dtype(dtype).itemsize))).contents
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


def block_balance(row, col, degree, window, wide, part_size, kind, want_trans=False):
    """Run the C oracle.  ``degree`` is float16 (FP16), float32 (FP32) or None (MAP).

    Returns [rp, col, val, window_of_part, atomic] (+ map_t inserted after val
    for MAP with want_trans), all numpy arrays.
    """
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    if kind == FP16:
        deg = np.ascontiguousarray(degree, dtype=np.float16).view(np.uint16)
    elif kind == FP32:
        deg = np.ascontiguousarray(degree, dtype=np.float32)
    else:
        deg = np.zeros(1, np.int32)
    res = _Result()
    rc = lib().bo_block_balance(row.ctypes.data, row.size, col.ctypes.data, deg.ctypes.data,
                                kind, window, wide, part_size, int(want_trans), ctypes.byref(res))
    if rc != 0:
        raise ValueError("bad arguments to block_balance")
    try:
        vdt = {FP16: np.float16, FP32: np.float32, MAP: np.int32}[kind]
        out = [_arr(res.rp, res.n_rp, np.int32), _arr(res.col, res.n_col, np.int32),
               _arr(res.val, res.n_val, vdt)]
        if kind == MAP and want_trans:
            out.append(_arr(res.map_t, res.n_val, np.int32))
        out += [_arr(res.win, res.n_parts, np.int32), _arr(res.atomic, res.n_parts, np.int32)]
    finally:
        lib().bo_free(ctypes.byref(res))
    return out


def block_balance_py(row, col, degree, window, wide, part_size, kind):
    """Independent pure-Python restatement (small inputs only).

    Follows gcn/Block_gcn/example.cpp:56-175 step by step; values use numpy's
    float16/float32 arithmetic (float16(float32(a)*float32(b)) for FP16).
    """
    row = [int(x) for x in row]
    col = [int(x) for x in col]
    rows = len(row) - 1
    rp, cols, vals, wins, ats = [0], [], [], [], []
    for i in range(rows // window):
        merged = sorted(set(col[row[i * window]:row[i * window + window]]))
        v = len(merged)
        pad = -(-v // wide) * wide
        blocks = pad // wide
        if kind != FP32 and blocks == 0:
            continue
        pos = {c: k for k, c in enumerate(merged)}
        cols += merged + [-1] * (pad - v)
        if kind == MAP:
            tile = [-1] * (pad * window)
        else:
            tile = [0.0] * (pad * window)
        for p in range(window):
            j = i * window + p
            for m in range(row[j], row[j + 1]):
                c = pos[col[m]]
                o = (c // wide) * window * wide + p * wide + c % wide
                if kind == FP16:
                    tile[o] = np.float16(np.float32(np.float16(degree[j])) * np.float32(np.float16(degree[col[m]])))
                elif kind == FP32:
                    tile[o] = np.float32(np.float32(degree[j]) * np.float32(degree[col[m]]))
                else:
                    tile[o] = m
        vals += tile
        if blocks <= part_size:
            pairs, at = [(v, pad - v)], 0
        else:
            n, res = divmod(blocks, part_size)
            at = 1
            if res:
                pairs = [(part_size * wide, 0)] * n + [(res * wide - (pad - v), pad - v)]
            else:
                pairs = [(part_size * wide, 0)] * (n - 1) + [(part_size * wide - (pad - v), pad - v)]
        for a, b in pairs:
            rp.append(rp[-1] + a)
            rp.append(rp[-1] + b)
            wins.append(i)
            ats.append(at)
    vdt = {FP16: np.float16, FP32: np.float32, MAP: np.int32}[kind]
    return [np.array(rp, np.int32), np.array(cols, np.int32), np.array(vals, vdt),
            np.array(wins, np.int32), np.array(ats, np.int32)]


def load_reference():
    """Import the reference's own example.cpp built under oracle/_ref, or None."""
    if not os.path.isdir(REF_DIR):
        return None
    if REF_DIR not in sys.path:
        sys.path.insert(0, REF_DIR)
    try:
        import ref_block_gcn  # noqa: F401  (built by `make -C oracle ref`)
    except ImportError:
        return None
    return ref_block_gcn
