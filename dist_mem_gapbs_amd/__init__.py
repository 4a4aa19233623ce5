"""dist_mem_gapbs_amd -- MI355X-native edge-traversal engine for DMM-GAPBS.

Direction-optimizing BFS (src/bfs.cc:57-188), batched 64-source BFS and pull
PageRank (src/pr.cc:41-70) as HIP kernels for gfx950 behind a C ABI
(include/dmm_gapbs.h); the host graph substrate (generator, builder, source
picker, partition) is native C++ in the same library.
"""
from ._lib import (DMM_PR_FP32_KAHAN, DMM_PR_FP32_SERIAL, DMM_PR_FP64, DmmError, LIB_PATH,
                   load)
from .engine import Context, device_count, stats_dict
from .graph import (HostGraph, build_graph, from_edgelist, generate_edgelist, partition,
                    pick_sources, read_serialized_graph)

__all__ = [
    "Context", "DmmError", "HostGraph", "LIB_PATH", "DMM_PR_FP32_SERIAL", "DMM_PR_FP32_KAHAN",
    "DMM_PR_FP64", "build_graph", "device_count", "from_edgelist", "generate_edgelist", "load",
    "partition", "pick_sources", "read_serialized_graph", "stats_dict",
]
