"""song-mi355x: SONG approximate nearest-neighbour graph search on MI355X.

The compute path is libsong.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/song.h).  This package is its Python front end:

* ``nsw_build``   -- CPU NSW graph build (kernalGraph.h:140-149)
* ``Index``       -- device-resident index + ``search`` (gpu_search,
                     kernalGraph.h:307-382 / hashKernelGraph.h:360-435)
* ``cpu_search``  -- the reference's CPU search baseline, Graph::search
                     (graph.h:33-115), on the host
* ``kernel_graph``-- the reference's KernelGraph / HashKernelGraph /
                     gpu_search interface on top of the above
* ``datasets``    -- seeded synthetic workloads and fvecs/ivecs/npy I/O
"""
from . import _native
from ._native import SongError, device_count
from .index import (TRACE_FIELDS, Index, SearchResult, bithash_matrix, bithash_project,
                    bithash_project_device, cpu_search, degrees_for_block, graph_load, graph_save, nsw_build)

__all__ = [
    "SongError", "device_count", "TRACE_FIELDS", "Index", "SearchResult", "bithash_matrix",
    "bithash_project", "bithash_project_device", "cpu_search", "degrees_for_block", "graph_load", "graph_save", "nsw_build", "_native",
]
