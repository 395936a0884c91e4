"""Shared helpers for the GPU parity tests (CPU-side graph construction)."""
import numpy as np

TOY_ROW = np.array([0, 6, 7, 8, 10, 10, 11, 13, 15, 18, 19, 19, 19, 20, 20, 20, 20, 24,
                    25, 26, 28, 28, 29, 31, 33, 36, 37, 37, 37, 38, 38, 38, 38], np.int32)
TOY_COL = np.array([1, 2, 3, 11, 16, 21, 1, 1, 2, 24, 25, 16, 22, 0, 27, 0, 4, 25, 9, 27, 1, 3, 11, 21, 1, 1, 2,
                    24, 25, 16, 22, 0, 27, 0, 4, 25, 9, 27], np.int32)


def rand_csr(rng, n, max_deg, hub_rows=0):
    rows = []
    for i in range(n):
        k = int(rng.integers(0, max_deg + 1))
        if i < hub_rows:
            k = n // 2
        rows.append(np.unique(rng.integers(0, n, size=k)))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = (np.concatenate(rows) if rows else np.zeros(0)).astype(np.int32)
    return rp, col


def csr_rows(rp):
    return np.repeat(np.arange(len(rp) - 1), np.diff(rp))
