"""MI355X (gfx950) backend for the JV fast-elimination edge test and the exact
3/4/5-opt leaf cost matrix of cudaEdgeElimination.

The product is the native library ``libee_hip.so`` (C ABI in ``include/``);
this package binds it for tests, benchmarks and Python callers.
"""
from ._lib import BackendMissing, EeError, InvalidArgument, OverflowFailure, Unavailable
from .graph import CEIL_2D, EUC_2D, GraphSnapshot

__all__ = [
    "BackendMissing", "EeError", "InvalidArgument", "OverflowFailure", "Unavailable",
    "GraphSnapshot", "EUC_2D", "CEIL_2D",
]
