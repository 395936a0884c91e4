"""Seeded synthetic graphs and GCN normalisation (numpy, host).

The reference loads real graphs from ``.npz`` files (src_li, dst_li,
num_nodes; end2end_ori/gcn_final/mgat_csr/mdataset_fp16.py:27-53) that are not
shipped; the workloads here are generated instead, deterministically from a
seed with numpy's PCG64:

* ``chung_lu``     power-law graph: expected degrees w_i ~ (i + i0)^(-1/(gamma-1)),
                   endpoints drawn with probability ~ w, self pairs dropped,
                   symmetrised, deduplicated, self-loops added, CSR sorted.
                   i0 puts the largest expected degree at the structural cutoff
                   sqrt(N * mean_degree), where Chung-Lu probabilities stay <= 1.
* ``band_random``  |i - j| <= half_width band plus uniform off-band entries.

Striped node order (``stripe_pos``): 64-node stripes dealt round-robin to
``owners`` (a power of two, 2 to 32) contiguous ranges of ``q = m / owners`` positions, m a
multiple of 64 * owners.  The SpMM's owner split relies on it: every XCD
gathers only the feature rows of one owner range.
"""
import numpy as np


class CSR:
    """CSR with int32 indices (row pointers int64 internally for big graphs)."""

    def __init__(self, indptr, indices, n, values=None):
        self.indptr = np.asarray(indptr, dtype=np.int64)
        self.indices = np.asarray(indices, dtype=np.int32)
        self.n = int(n)
        self.values = values

    @property
    def nnz(self):
        return int(self.indices.size)

    def rowptr32(self):
        if self.nnz >= 2**31:
            raise OverflowError("nnz does not fit int32 row pointers")
        return self.indptr.astype(np.int32)

    def degrees(self):
        return np.diff(self.indptr)

    def row_ids(self):
        return np.repeat(np.arange(self.n, dtype=np.int32), self.degrees())

    def padded(self, multiple):
        """Rows appended (empty) so the row count is a multiple of `multiple`."""
        n2 = -(-self.n // multiple) * multiple if self.n else multiple
        ip = np.concatenate([self.indptr, np.full(n2 - self.n, self.indptr[-1], np.int64)])
        return CSR(ip, self.indices, n2, self.values)

    def transpose_perm(self):
        """perm[e] = position of entry (col(e), row(e)) in the CSR, -1 if absent."""
        rows = self.row_ids().astype(np.int64)
        cols = self.indices.astype(np.int64)
        key = rows * self.n + cols
        tkey = cols * self.n + rows
        pos = np.searchsorted(key, tkey)
        pos = np.minimum(pos, max(key.size - 1, 0))
        found = key.size > 0
        ok = (key[pos] == tkey) if found else np.zeros(0, bool)
        return np.where(ok, pos, -1).astype(np.int64)


def _from_keys(keys, n):
    """CSR from sorted unique int64 keys row*n+col."""
    rows = (keys // n).astype(np.int64)
    cols = (keys - rows * n).astype(np.int32)
    indptr = np.zeros(n + 1, np.int64)
    indptr[1:] = np.cumsum(np.bincount(rows, minlength=n))
    return CSR(indptr, cols, n)


def chung_lu_weights(n, mean_degree, gamma):
    alpha = 1.0 / (gamma - 1.0)
    target_max = np.sqrt(n * mean_degree)
    # smallest i0 with max expected degree <= target_max (bisection on log i0)
    i = np.arange(n, dtype=np.float64)

    def max_deg(i0):
        w = (i + i0) ** -alpha
        return n * mean_degree * w[0] / w.sum()

    lo, hi = 1e-3, float(n)
    for _ in range(60):
        mid = np.sqrt(lo * hi)
        if max_deg(mid) > target_max:
            lo = mid
        else:
            hi = mid
    w = (i + hi) ** -alpha
    return w / w.sum()


def chung_lu(n, mean_degree, gamma=2.1, seed=0, target_nnz=None, self_loops=True):
    """Symmetric power-law CSR (see module docstring).

    mean_degree is the mean undirected degree before deduplication: n*mean/2
    endpoint pairs are drawn.  With target_nnz, extra pairs are drawn in up to
    four rounds until nnz (after symmetrisation, dedupe and self-loops) is
    within 0.5% of the target.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    p = chung_lu_weights(n, mean_degree, gamma)
    cdf = np.cumsum(p)
    cdf[-1] = 1.0

    def draw(m):
        u = np.searchsorted(cdf, rng.random(m)).astype(np.int64)
        v = np.searchsorted(cdf, rng.random(m)).astype(np.int64)
        keep = u != v
        u, v = u[keep], v[keep]
        return np.concatenate([u * n + v, v * n + u])

    loops = np.arange(n, dtype=np.int64) * (n + 1) if self_loops else np.zeros(0, np.int64)
    keys = np.unique(np.concatenate([draw(int(n * mean_degree / 2)), loops]))
    if target_nnz is not None:
        for _ in range(4):
            short = target_nnz - keys.size
            if abs(short) <= 0.005 * target_nnz or short <= 0:
                break
            # each drawn pair adds ~2 entries; duplicates grow with density
            dup = 1.0 - (keys.size - loops.size) / max(n * mean_degree, 1)
            m = int(short / 2 / max(1.0 - dup, 0.5) * 1.02) + 1
            keys = np.unique(np.concatenate([keys, draw(m)]))
    return _from_keys(keys, n)


def directed_chung_lu_symmetrized(n, target_nnz, gamma=2.1, seed=0):
    """Directed Chung-Lu draws symmetrised, with self-loops, nnz ~ target."""
    mean = max((target_nnz - n) / n, 1.0)
    return chung_lu(n, mean, gamma, seed, target_nnz=target_nnz)


def band_random(n, nnz_target, half_width=2, seed=0, value_seed=None):
    """|i-j| <= half_width band plus uniform random off-band entries to ~nnz_target
    (not symmetrised), values N(0,1) float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    band = []
    for d in range(-half_width, half_width + 1):
        i = np.arange(max(0, -d), min(n, n - d), dtype=np.int64)
        band.append(i * n + (i + d))
    keys = np.unique(np.concatenate(band))
    extra = nnz_target - keys.size
    while extra > 0:
        m = int(extra * 1.05) + 16
        r = rng.integers(0, n, size=m, dtype=np.int64)
        c = rng.integers(0, n, size=m, dtype=np.int64)
        off = np.abs(r - c) > half_width
        new = np.unique(r[off] * n + c[off])
        new = new[~np.isin(new, keys, assume_unique=True)]
        if new.size > extra:
            new = rng.choice(new, size=extra, replace=False)
        keys = np.unique(np.concatenate([keys, new]))
        extra = nnz_target - keys.size
    csr = _from_keys(keys, n)
    vrng = np.random.Generator(np.random.PCG64(seed + 1 if value_seed is None else value_seed))
    csr.values = vrng.standard_normal(csr.nnz).astype(np.float32)
    return csr


def gcn_norm_degree(csr):
    """d_i = 1/sqrt(row count) in fp32 (rows already include the self-loop)."""
    deg = csr.degrees().astype(np.float32)
    with np.errstate(divide="ignore"):
        return (1.0 / np.sqrt(deg)).astype(np.float32)


def gcn_norm_values(csr, d=None):
    """A_hat values d_i * d_j (fp32) in CSR order."""
    d = gcn_norm_degree(csr) if d is None else d
    return (d[csr.row_ids()] * d[csr.indices]).astype(np.float32)


def stripe_rows(n, owners):
    """Rows of a graph of n nodes in striped order: n rounded up to a multiple
    of 64 * owners (also a multiple of every window height)."""
    if owners not in (2, 4, 8, 16, 32):
        raise ValueError(f"owners must be a power of two from 2 to 32, got {owners}")
    unit = 64 * owners
    return max(-(-int(n) // unit) * unit, unit)


def stripe_pos(j, owners, m):
    """Striped position of node ids j (numpy integer array or torch tensor):
    ((j >> 6) mod owners) * q + (j >> (6 + log2 owners)) * 64 + (j mod 64),
    q = m / owners.  A bijection on [0, m)."""
    lg = {2: 1, 4: 2, 8: 3, 16: 4, 32: 5}[owners]
    q = m // owners
    return ((j >> 6) & (owners - 1)) * q + ((j >> (6 + lg)) << 6) + (j & 63)


def window_unique_columns(csr, window):
    """U = sum over row windows of distinct columns (for the byte model)."""
    rows = csr.row_ids().astype(np.int64) // window
    keys = rows * csr.n + csr.indices.astype(np.int64)
    return int(np.unique(keys).size)


def csr_segments(indptr, max_len=32):
    """Segments of CSR rows of at most `max_len` consecutive entries, in the
    layout of ``attgnn_segments`` (include/attgnn.h): int32 arrays seg_ptr
    [n_seg+1], seg_row [n_seg], seg_slot [n_seg] (-1 for a row's only segment,
    else its index among the segments of rows that have several) and row_seg
    [n+1], plus n_seg and n_slots."""
    indptr = np.asarray(indptr, np.int64)
    n = indptr.size - 1
    nseg = (np.diff(indptr) + max_len - 1) // max_len          # 0 for an empty row
    row_seg = np.zeros(n + 1, np.int64)
    row_seg[1:] = np.cumsum(nseg)
    n_seg = int(row_seg[-1])
    seg_row = np.repeat(np.arange(n, dtype=np.int64), nseg)
    within = np.arange(n_seg, dtype=np.int64) - row_seg[seg_row]
    seg_ptr = np.append(indptr[seg_row] + within * max_len, indptr[-1])
    multi = nseg[seg_row] > 1
    seg_slot = np.where(multi, np.cumsum(multi) - 1, -1)
    if n_seg + 1 >= 2**31 or indptr[-1] >= 2**31:
        raise OverflowError("segments do not fit int32")
    return {"seg_ptr": seg_ptr.astype(np.int32), "seg_row": seg_row.astype(np.int32),
            "seg_slot": seg_slot.astype(np.int32), "row_seg": row_seg.astype(np.int32),
            "n_seg": n_seg, "n_slots": int(multi.sum())}
