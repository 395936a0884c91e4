"""The product preprocessing (libattgnn.so, attgnn_blocks_build) is bit-exact
against the oracle and the reference's golden outputs; malformed input raises."""
import hashlib

import numpy as np
import pytest
import torch

from att_amd import ops
from att_amd._lib import AttGnnError
from oracle import block as B

GRID = [(w, wd, ps) for w in (8, 16) for wd in (4, 8, 16, 32) for ps in (1, 2, 32)]
NAMES = ("rp", "col", "val", "window", "atomic")


def digest(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(str(a.dtype).encode() + str(a.shape).encode() + a.tobytes()).hexdigest()


def build(row, col, kind, d, w, wd, ps, trans=False):
    if kind == "fp16":
        r = ops.blocks_build(row, col, "deg_f16", d, torch.float16, w, wd, ps, emit_empty=False)
    elif kind == "tf32":
        r = ops.blocks_build(row, col, "deg_f32", d, torch.float32, w, wd, ps, emit_empty=True)
    else:
        r = ops.blocks_build(row, col, "map", None, torch.int32, w, wd, ps, emit_empty=False, want_trans=trans)
    names = ("rp", "col", "val", "map_t", "window", "atomic") if trans else NAMES
    return [r[k].numpy() for k in names]


def test_golden_toy(golden_dir):
    g = np.load(f"{golden_dir}/toy.npz")
    for w, wd, ps in GRID:
        for dname in ("deg16_ones", "deg16"):
            for n, a in zip(NAMES, build(g["row"], g["col"], "fp16", g[dname], w, wd, ps)):
                assert a.tobytes() == g[f"fp16_{dname}_{w}_{wd}_{ps}_{n}"].tobytes()
        for n, a in zip(NAMES, build(g["row"], g["col"], "tf32", g["deg32"], w, wd, ps)):
            assert a.tobytes() == g[f"tf32_{w}_{wd}_{ps}_{n}"].tobytes()


def test_golden_chung_lu_digests(golden_dir):
    g = np.load(f"{golden_dir}/cl2k.npz")
    for w, wd, ps in GRID:
        for n, a in zip(NAMES, build(g["row"], g["col"], "fp16", g["deg16"], w, wd, ps)):
            assert digest(a) == str(g[f"sha_fp16_{w}_{wd}_{ps}_{n}"])
        for n, a in zip(NAMES, build(g["row"], g["col"], "tf32", g["deg32"], w, wd, ps)):
            assert digest(a) == str(g[f"sha_tf32_{w}_{wd}_{ps}_{n}"])


@pytest.mark.parametrize("seed", range(5))
def test_matches_oracle_random(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    rows = [np.unique(rng.integers(0, n, size=int(rng.integers(0, 30)))) for _ in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    d16 = (rng.random(n) + 0.1).astype(np.float16)
    d32 = rng.random(n).astype(np.float32)
    for w, wd, ps in GRID:
        for a, b in zip(build(rp, col, "fp16", d16, w, wd, ps), B.block_balance(rp, col, d16, w, wd, ps, B.FP16)):
            assert a.tobytes() == b.tobytes()
        for a, b in zip(build(rp, col, "tf32", d32, w, wd, ps), B.block_balance(rp, col, d32, w, wd, ps, B.FP32)):
            assert a.tobytes() == b.tobytes()
        for a, b in zip(build(rp, col, "map", None, w, wd, ps, trans=True),
                        B.block_balance(rp, col, None, w, wd, ps, B.MAP, want_trans=True)):
            assert a.tobytes() == b.tobytes()


def test_unsorted_rows_same_structure():
    rng = np.random.default_rng(3)
    n = 64
    rows = [np.unique(rng.integers(0, n, size=9)) for _ in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    shuf = np.concatenate([rng.permutation(r) for r in rows]).astype(np.int32)
    d = np.ones(n, np.float32)
    a = build(rp, col, "tf32", d, 8, 8, 2)
    b = build(rp, shuf, "tf32", d, 8, 8, 2)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()


def test_edge_values_and_bf16_tiles():
    rng = np.random.default_rng(4)
    n = 80
    rows = [np.unique(rng.integers(0, n, size=7)) for _ in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    v = rng.standard_normal(col.size).astype(np.float32)
    r32 = ops.blocks_build(rp, col, "edge_f32", v, torch.float32, 16, 32, 4, emit_empty=True)
    rbf = ops.blocks_build(rp, col, "edge_f32", v, torch.bfloat16, 16, 32, 4, emit_empty=True)
    assert torch.equal(r32["val"].to(torch.bfloat16), rbf["val"])
    assert sorted(r32["val"][r32["val"] != 0].tolist()) == sorted(v[:rp[n // 16 * 16]][v[:rp[n // 16 * 16]] != 0].tolist())


def test_malformed_input_raises():
    with pytest.raises(AttGnnError):
        ops.blocks_build([0, 2, 1], [0, 1], "deg_f32", np.ones(2), torch.float32, 8, 8, 2, emit_empty=True)
    with pytest.raises(AttGnnError):
        ops.blocks_build([0, 1, 2], [0, 5], "deg_f32", np.ones(2), torch.float32, 8, 8, 2, emit_empty=True)
    with pytest.raises(AttGnnError):
        ops.blocks_build([0, 1, 2], [0, 1], "deg_f32", np.ones(2), torch.float32, 0, 8, 2, emit_empty=True)
    with pytest.raises(AttGnnError):
        ops.blocks_build([0, 1, 3], [0, 1], "deg_f32", np.ones(2), torch.float32, 8, 8, 2, emit_empty=True)


def _expand_compact(masks, base, values, C, W):
    """Dense [C*W] tile array from compact tiles (test-side restatement of the format)."""
    masks = masks.view(np.uint16) if masks.dtype != np.uint16 else masks
    dense = np.zeros(C * W, values.dtype)
    for b in range(C // 32):
        idx = int(base[b])
        for g in range(4):
            for r in range(W):
                for c in range(8 * g, 8 * g + 8):
                    if (int(masks[b * 32 + c]) >> r) & 1:
                        dense[(b * W + r) * 32 + c] = values[idx]
                        idx += 1
        assert idx == base[b + 1]
    return dense


@pytest.mark.parametrize("window,dtype", [(16, torch.float16), (16, torch.float32), (8, torch.bfloat16)])
def test_compact_tiles_expand_to_dense(window, dtype):
    rng = np.random.default_rng(21)
    n = 300
    rows = [np.unique(rng.integers(0, n, size=int(rng.integers(0, 25)))) for _ in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    v = rng.standard_normal(col.size).astype(np.float32)
    r = ops.blocks_build(rp, col, "edge_f32", v, dtype, window, 32, 3, emit_empty=True, compact=True)
    C = r["col"].numel()
    dense = _expand_compact(r["masks"].numpy().view(np.uint16), r["block_base"].numpy(),
                            r["values"].numpy() if dtype != torch.bfloat16 else r["values"].view(torch.int16).numpy(),
                            C, window)
    want = r["val"].numpy() if dtype != torch.bfloat16 else r["val"].view(torch.int16).numpy()
    assert np.array_equal(dense, want)
    # one packed value per CSR entry of the covered rows (values are nonzero a.s.)
    assert int(r["block_base"][-1]) == rp[n // window * window]
