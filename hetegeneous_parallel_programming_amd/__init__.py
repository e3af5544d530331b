"""MI355X-native triangle counting (the bs_warp_vertex set-intersection path).

Host side: :mod:`.graph` (CSR, generators, orientation, partitioning) over
``libtc_host.so``.  Device side: :mod:`.solver` (count, partition count,
batched persistent kernel, serving loop) over ``libtc_hip.so`` -- hand-written
HIP kernels for gfx950 behind a C ABI (``include/tc.h``).
"""
from .graph import Graph, HaloPartition, MetisPart, PartitionedGraph, assemble_metis, batch_spec  # noqa: F401

__all__ = ["Graph", "HaloPartition", "MetisPart", "PartitionedGraph", "assemble_metis", "batch_spec"]
