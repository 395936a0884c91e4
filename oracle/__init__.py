"""TEST INFRASTRUCTURE ONLY -- the oracle.

CPU restatements of the reference's window-block preprocessing (C, in
``block_oracle.c``) and FP64 numpy closed forms of the sparse kernels and GNN
layers on its path (``dense.py``).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg may import this package, and only as the
checker; the product (``att_amd``) never imports it.
"""
