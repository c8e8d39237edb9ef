"""MI355X-native rebuild of merijn/Belewitte's graph-kernel hot path.

Level-synchronous BFS variants and PageRank iteration variants as hand-written
HIP kernels for gfx950 behind a C ABI (include/belewitte_hip.h), a host graph
substrate (include/belewitte_graph.h), and the reference's plugin API
re-implemented in C++ (belewitte_amd/csrc/framework) for the drop-in
kernel-runner.
"""
from . import _native  # noqa: F401

__all__ = ["graph", "device", "bfs", "pagerank"]
