"""att_amd -- MI355X-native window-block sparse GNN path.

Window-block preprocessing (host C++), blocked SpMM / SDDMM on gfx950 MFMA
(HIP), GCN / GAT / GATv2 layers with hand-written backward kernels, and a
continuous-batching inference loop, behind the reference's extension API
(see ``att_amd.compat``), as ``torch.ops.attgnn.*`` once ``att_amd.torch_ops``
is imported, and over a C ABI (include/attgnn.h, libattgnn.so).
"""
from . import _lib  # noqa: F401

__all__ = ["ops", "compat", "graphs", "layers", "serve", "torch_ops"]
