"""MI355X-native cooperative workgroup runtime (wgm).

Occupancy discovery, inter-workgroup barriers and a resizing scheduler for
persistent HIP kernels on gfx950, with persistent BFS / SSSP / colouring / MIS
/ PageRank, batched multi-graph launches and a serving loop. The compute path
is libwgm.so (hand-written HIP + a C ABI, include/wgm.h); this package is its
host-side Python mirror. Importing it loads libwgm.so and fails loudly if the
library has not been built.
"""
from . import _lib
from ._lib import WgmError, lib

lib()  # no silent fallback: the native library must be present

from . import graphs  # noqa: E402
from .runtime import DeviceGraph, Runtime  # noqa: E402

__all__ = ["Runtime", "DeviceGraph", "graphs", "WgmError", "lib"]
