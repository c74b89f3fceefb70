"""MI355X-native subgraph matcher with the call surfaces of
scorpiontornado/PSM_Benchmark (filters, BuildTable, matching orders,
enumeration) on hand-written HIP kernels for gfx950.

The compute path is libpsm.so (C ABI in include/psm_abi.h); this package is
its Python driver. Nothing here falls back to the CPU.
"""
from .graph_io import CSRGraph, csr_from_edges, load_text, save_text, load_binary, save_binary  # noqa: F401

__all__ = ["CSRGraph", "csr_from_edges", "load_text", "save_text", "load_binary", "save_binary"]
