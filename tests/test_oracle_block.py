"""The preprocessing oracle (oracle/block_oracle.c) pinned against the
reference: golden outputs of the reference's own example.cpp, an independent
pure-Python restatement, invariants, and the reference build itself when present."""
import hashlib

import numpy as np
import pytest

from oracle import block as B

GRID = [(w, wd, ps) for w in (8, 16) for wd in (4, 8, 16, 32) for ps in (1, 2, 32)]
NAMES = ("rp", "col", "val", "window", "atomic")


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def digest(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(str(a.dtype).encode() + str(a.shape).encode() + a.tobytes()).hexdigest()


def rand_csr(rng, n, deg, sorted_rows=True):
    rows = []
    for _ in range(n):
        r = np.unique(rng.integers(0, n, size=int(rng.integers(0, 2 * deg + 1))))
        if not sorted_rows:
            rng.shuffle(r)
        rows.append(r)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate(rows).astype(np.int32) if n else np.zeros(0, np.int32)
    return rp, col


def test_half_conversion_exhaustive():
    L = B.lib()
    bits = np.arange(65536, dtype=np.uint16)
    want = bits.view(np.float16).astype(np.float32)
    got = np.array([L.bo_half_to_float(int(b)) for b in bits], np.float32)
    fin = np.isfinite(want)
    assert np.array_equal(got[fin].view(np.uint32), want[fin].view(np.uint32))
    assert np.all(np.isnan(got[np.isnan(want)]))


def test_float_to_half_rne_against_numpy():
    L = B.lib()
    rng = np.random.default_rng(0)
    f = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-9, 6, 20000),
        np.array([65504, 65519.99, 65520, 65536, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8, 0.0, -0.0, 1e-45,
                  np.inf, -np.inf], np.float32),
        # exact ties between adjacent halves
        (np.arange(1, 2000, dtype=np.float32) + 0.5) * np.float32(2.0 ** -10),
        (np.arange(1, 2000, dtype=np.float32) + 0.5) * np.float32(2.0 ** -24),
    ]).astype(np.float32)
    got = np.array([L.bo_float_to_half(float(x)) for x in f], np.uint16)
    want = f.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


def test_toy_golden_all_parameters(golden_dir):
    g = np.load(f"{golden_dir}/toy.npz")
    for w, wd, ps in GRID:
        for dname in ("deg16_ones", "deg16"):
            out = B.block_balance(g["row"], g["col"], g[dname], w, wd, ps, B.FP16)
            for n, a in zip(NAMES, out):
                assert same(a, g[f"fp16_{dname}_{w}_{wd}_{ps}_{n}"]), (dname, w, wd, ps, n)
        out = B.block_balance(g["row"], g["col"], g["deg32"], w, wd, ps, B.FP32)
        for n, a in zip(NAMES, out):
            assert same(a, g[f"tf32_{w}_{wd}_{ps}_{n}"]), ("tf32", w, wd, ps, n)


def test_chung_lu_golden_digests(golden_dir):
    g = np.load(f"{golden_dir}/cl2k.npz")
    for w, wd, ps in GRID:
        o16 = B.block_balance(g["row"], g["col"], g["deg16"], w, wd, ps, B.FP16)
        o32 = B.block_balance(g["row"], g["col"], g["deg32"], w, wd, ps, B.FP32)
        for n, a in zip(NAMES, o16):
            assert digest(a) == str(g[f"sha_fp16_{w}_{wd}_{ps}_{n}"]), ("fp16", w, wd, ps, n)
        for n, a in zip(NAMES, o32):
            assert digest(a) == str(g[f"sha_tf32_{w}_{wd}_{ps}_{n}"]), ("tf32", w, wd, ps, n)


@pytest.mark.parametrize("seed", range(6))
def test_c_oracle_matches_python_restatement(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 90))
    rp, col = rand_csr(rng, n, int(rng.integers(0, 10)))
    d16 = (rng.random(n) * 2 + 0.01).astype(np.float16)
    d32 = rng.random(n).astype(np.float32)
    for w, wd, ps in [(8, 8, 32), (16, 32, 2), (8, 4, 1), (16, 16, 1), (8, 16, 2)]:
        for kind, d in ((B.FP16, d16), (B.FP32, d32), (B.MAP, None)):
            c = B.block_balance(rp, col, d, w, wd, ps, kind)
            p = B.block_balance_py(rp, col, d, w, wd, ps, kind)
            for a, b in zip(c, p):
                assert same(a, b), (seed, w, wd, ps, kind)


@pytest.mark.parametrize("seed", range(4))
def test_invariants(seed):
    """Tile expansion reproduces the CSR; columns ascend then pad with -1;
    parts respect partSize; atomic iff the window is split; fp16 drops empty
    windows while tf32 emits (0,0)."""
    rng = np.random.default_rng(100 + seed)
    n = 150
    rp, col = rand_csr(rng, n, 6)
    d32 = (rng.random(n) + 0.5).astype(np.float32)
    for w, wd, ps in [(8, 8, 2), (16, 32, 1), (16, 8, 32)]:
        r, c, val, win, at = B.block_balance(rp, col, d32, w, wd, ps, B.FP32)
        dense = np.zeros((n, n), np.float32)
        for p in range(len(win)):
            lo, hi = r[2 * p], r[2 * p + 2]
            assert r[2 * p + 1] - lo <= ps * wd
            assert (hi - lo) % wd == 0 and hi - lo <= ps * wd
            cols = c[lo:r[2 * p + 1]]
            assert np.all(np.diff(cols) > 0) and np.all(c[r[2 * p + 1]:hi] == -1)
            for cc in range(lo, hi):
                if c[cc] < 0:
                    continue
                for rr in range(w):
                    v = val[(cc // wd) * w * wd + rr * wd + cc % wd]
                    if v != 0:
                        dense[win[p] * w + rr, c[cc]] = v
        ref = np.zeros((n, n), np.float32)
        for i in range(n // w * w):
            for m in range(rp[i], rp[i + 1]):
                ref[i, col[m]] = d32[i] * d32[col[m]]
        assert np.array_equal(dense, ref)
        nwin = n // w
        counts = np.bincount(win, minlength=nwin)
        assert counts.size == nwin and np.all(counts >= 1)          # tf32 keeps empty windows
        for i in range(nwin):
            assert np.all(at[win == i] == (1 if counts[i] > 1 else 0))
        r16 = B.block_balance(rp, col, d32.astype(np.float16), w, wd, ps, B.FP16)
        empty = [i for i in range(nwin) if rp[(i + 1) * w] == rp[i * w]]
        assert not set(empty) & set(r16[3].tolist())                 # fp16 drops them


def test_map_and_transposed_map():
    rng = np.random.default_rng(5)
    n = 120
    rp, col = rand_csr(rng, n, 5)
    # symmetrise so every entry has a transpose
    A = np.zeros((n, n), bool)
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = True
    A |= A.T
    rp = np.concatenate([[0], np.cumsum(A.sum(1))]).astype(np.int32)
    col = np.concatenate([np.nonzero(A[i])[0] for i in range(n)]).astype(np.int32)
    r, c, m, mt, win, at = B.block_balance(rp, col, None, 8, 8, 2, B.MAP, want_trans=True)
    seen = m[m >= 0]
    assert np.array_equal(np.sort(seen), np.arange(n // 8 * 8 and rp[n // 8 * 8]))
    rows = np.repeat(np.arange(n), np.diff(rp))
    for s in np.nonzero(m >= 0)[0]:
        e, et = m[s], mt[s]
        assert rows[et] == col[e] and col[et] == rows[e]


def test_matches_reference_build_when_present():
    ref = B.load_reference()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import torch
    rng = np.random.default_rng(11)
    for _ in range(5):
        n = int(rng.integers(1, 300))
        rp, col = rand_csr(rng, n, int(rng.integers(0, 15)))
        d16 = (rng.random(n) * 2 + 0.01).astype(np.float16)
        d32 = rng.random(n).astype(np.float32)
        for w, wd, ps in GRID:
            o = B.block_balance(rp, col, d16, w, wd, ps, B.FP16)
            r = [t.numpy() for t in ref.blockProcess_fp16_balance(torch.from_numpy(rp), torch.from_numpy(col),
                                                                  torch.from_numpy(d16), w, wd, ps)]
            assert all(same(a, b) for a, b in zip(o, r))
            o = B.block_balance(rp, col, d32, w, wd, ps, B.FP32)
            r = [t.numpy() for t in ref.blockProcess_tf32_balance(torch.from_numpy(rp), torch.from_numpy(col),
                                                                  torch.from_numpy(d32), w, wd, ps)]
            assert all(same(a, b) for a, b in zip(o, r))
