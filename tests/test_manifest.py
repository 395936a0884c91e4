"""CPU: the committed workload manifest (att_amd/benchmarks/manifest.json)
matches what the benchmark generators produce -- nodes, entries, maximum
degree, U16 (distinct columns over 16-row windows), condensed columns padded
to 32 per window and the CSR digest -- for configurations 1, 3 and 4 (the
114.6M-entry graph of configuration 2 is checked on the GPU box, where the
benchmark builds it, and reported in bench.py's line)."""
import numpy as np
import pytest

from att_amd import graphs as G
from att_amd.benchmarks import manifest as M


@pytest.mark.parametrize("name", ["cfg1_chung_lu_16k", "cfg3_attention_2.48M", "cfg4_band_random_2^20"])
def test_manifest_matches_generators(name):
    want = M.load()[name]
    call, gen = M.GENERATORS[name]
    assert want["call"] == call
    assert {"call": call, **M.describe(gen())} == want


def test_window_columns_brute_force():
    """U16 and the padded count against a loop over windows on a small graph."""
    g = G.chung_lu(700, 9, 2.1, seed=4)
    u, padded = M.window_columns(g)
    U = P = 0
    for w in range(0, g.n, 16):
        cols = set(g.indices[g.indptr[w]:g.indptr[min(w + 16, g.n)]].tolist())
        U += len(cols)
        P += -(-len(cols) // 32) * 32
    assert (u, padded) == (U, P)
    # a planted mistake -- 8-row windows -- gives another count
    assert M.window_columns(g, window=8)[0] != U
