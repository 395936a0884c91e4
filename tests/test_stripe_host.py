"""Striped node order on the host: stripe_pos is a bijection whose owner
ranges and 16-row windows are what the owner-split SpMM assumes; a striped
structure built on the host keeps every part's columns ascending, so each
owner's columns form one segment; the striped entry points validate their
arguments before touching a device."""
import ctypes

import numpy as np
import pytest
import torch

from att_amd import _lib
from att_amd import graphs as G
from att_amd import ops


@pytest.mark.parametrize("owners", [2, 4, 8, 16, 32])
@pytest.mark.parametrize("units", [1, 3, 17])
def test_stripe_pos_bijection_and_ranges(owners, units):
    m = 64 * owners * units
    assert G.stripe_rows(m, owners) == m and G.stripe_rows(m - 1, owners) == m
    j = np.arange(m, dtype=np.int64)
    p = G.stripe_pos(j, owners, m)
    assert np.array_equal(np.sort(p), j)                       # a permutation of [0, m)
    q = m // owners
    assert np.array_equal(p // q, (j >> 6) % owners)            # owner = stripe index mod owners
    # stripes keep their order inside an owner range, nodes keep theirs inside a stripe
    for o in range(owners):
        mine = j[(j >> 6) % owners == o]
        assert np.array_equal(p[mine], o * q + np.arange(mine.size))
    # every 16-row window is contiguous in striped order
    w = p.reshape(-1, 16)
    assert np.array_equal(w - w[:, :1], np.broadcast_to(np.arange(16), w.shape))


def test_stripe_pos_planted_mistake_fails():
    """An owner count of 4 read as 2 in the shift (a plausible slip) is not
    a bijection with the stated owner ranges."""
    m = 64 * 4 * 5
    j = np.arange(m, dtype=np.int64)
    q = m // 4
    wrong = ((j >> 6) & 3) * q + ((j >> 7) << 6) + (j & 63)
    assert not np.array_equal(np.sort(wrong), j)


def _striped_structure(owners, part_size, seed):
    g = G.chung_lu(1500, 30, 2.1, seed=seed).padded(64 * owners)
    m = g.n
    scol = G.stripe_pos(g.indices.astype(np.int64), owners, m)
    order = np.lexsort((scol, g.row_ids()))
    vals = np.ones(g.nnz, np.float32)
    b = ops.blocks_build(g.rowptr32(), scol[order].astype(np.int32), "edge_f32", vals, torch.float16, 16, 32,
                         part_size, emit_empty=True, compact=True)
    return b, m


@pytest.mark.parametrize("owners", [2, 4, 8])
@pytest.mark.parametrize("part_size", [1, 3, 1000])
def test_striped_parts_have_ascending_owner_segments(owners, part_size):
    b, m = _striped_structure(owners, part_size, seed=owners)
    rp = b["rp"].numpy().astype(np.int64)
    col = b["col"].numpy().astype(np.int64)
    win = b["window"].numpy()
    q = m // owners
    for p in range(win.size):
        c = col[rp[2 * p]:rp[2 * p + 1]]
        assert np.all(c[1:] > c[:-1]), p                     # ascending within the part
        seg = np.searchsorted(c, np.arange(owners + 1) * q)  # the segments the device kernel finds
        counts = np.diff(seg)
        assert counts.sum() == c.size
        assert np.array_equal(counts, np.bincount(c // q, minlength=owners))


def test_striped_entry_points_reject_bad_arguments_without_device():
    L = _lib.lib()
    nul = ctypes.c_void_p(0)
    fake = ctypes.c_void_p(0x1000)   # never dereferenced: validation fails first

    def spmm(owners, m, n_x, dim=256, window=16, x=_lib.F16):
        return L.attgnn_spmm_compact_striped(fake, fake, fake, fake, fake, _lib.F16, fake, fake, 0, fake, owners,
                                             fake, x, n_x, fake, _lib.F16, m, dim, window, _lib.MODE_NATIVE, nul, 0,
                                             0, nul)
    assert spmm(3, 1024, 1024) == -1          # owners not a power of two
    assert spmm(64, 4096, 4096) == -1         # more than 32 owners
    assert spmm(4, 1000, 1000) == -1          # m not a multiple of 64 * owners
    assert spmm(4, 1024, 1000) == -1          # X shorter than m
    assert spmm(4, 1024, 1024, window=12) == -1
    assert spmm(4, 1024, 1024, x=_lib.BF16) == -4   # tile / X types differ
    assert L.attgnn_stripe_segments_device(fake, fake, fake, 10, 5, 1024, fake, nul) == -1
    assert L.attgnn_stripe_segments_device(fake, fake, fake, 10, 4, 1000, fake, nul) == -1
    # an owner tile per XCD per part (4 rows x 8 features x 64 lanes), or the usual split partials
    assert L.attgnn_spmm_striped_workspace_elems(10, 16, 256, 4) == 10 * 8 * 64 * 4 * 8
    assert L.attgnn_spmm_striped_workspace_elems(10, 16, 4096, 4) == L.attgnn_spmm_workspace_elems(10, 16, 4096)
