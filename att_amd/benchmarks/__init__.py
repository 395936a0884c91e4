"""Benchmarks beyond bench.py's headline line (each prints JSON lines):

  python -m att_amd.benchmarks.gat        GAT / GATv2 fwd+bwd vs plain FP32 PyTorch
  python -m att_amd.benchmarks.precision  SpMM precision sweep vs the FP64 oracle
  python -m att_amd.benchmarks.serving    continuous-batching inference loop
"""
