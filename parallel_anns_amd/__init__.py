"""parallel_anns_amd: MI355X-native NSG top-M / PSS_v5 graph search.

The search path runs as hand-written HIP kernels for gfx950 behind the C-ABI
in include/panns_amd.h; `search.Index` is its Python face and
include/panns/Searching.h its C++ drop-in for PANNS::Searching.  `tools` holds
the inputs around the path (synthetic data, formats, exact k-NN, graphs).
"""
from ._native import NativeLibraryMissing  # noqa: F401

__all__ = ["NativeLibraryMissing", "search", "tools", "datasets"]
