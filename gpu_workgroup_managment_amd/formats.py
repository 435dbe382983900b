"""Graph and solution files of the reference's datasets, through libwgm's
native readers and writers (include/wgm.h, csrc/formats.cpp).

Formats (the reference's readers):
  "gr"      Galois .gr v1 binary (src/lonestar/graph_apps/bfs/bfs.h:95-169)
  "dimacs"  DIMACS shortest path, "p sp n m" / "a u v w"
            (src/pannotia/graph_parser/parse.cpp:177-281)
  "metis"   METIS adjacency lists (parse.cpp:58-175)
  "mm"      MatrixMarket coordinate (parse.cpp:467-580)
  "edges"   Lonestar text edge list (src/lonestar/graph_apps/sssp/sssp.h:49-94)

Both Galois binaries and DIMACS road graphs (USA-road-d.*.gr) use the ".gr"
extension; `detect_format` tells them apart by the first bytes.

Solution styles: "colon" ("%d:%d", Lonestar BFS/SSSP, bfs.h:336-365),
"colon_space" ("%d: %d", Pannotia color/SSSP, color.h:149-190), "plain"
("%d", Pannotia MIS, mis.h:170-195); floats for BC (bc.h:176-203).
"""
import ctypes
import os

import numpy as np

from . import _lib as L
from .graphs import _take

FORMATS = {"gr": L.FMT_GR, "dimacs": L.FMT_DIMACS, "metis": L.FMT_METIS, "mm": L.FMT_MM, "edges": L.FMT_EDGES}
STYLES = {"colon": L.SOL_COLON, "colon_space": L.SOL_COLON_SPACE, "plain": L.SOL_PLAIN}
EXTENSIONS = {".gr": None, ".dimacs": "dimacs", ".graph": "metis", ".metis": "metis", ".mtx": "mm", ".mm": "mm",
              ".edges": "edges", ".txt": "edges"}


def _check(rc):
    if rc != L.WGM_OK:
        raise L.WgmError(rc, (L.lib().wgm_format_error() or b"").decode(errors="replace"))


def detect_format(path):
    """Format name from the extension; a ".gr" file is Galois binary when it
    opens with the little-endian version word 1, DIMACS text otherwise."""
    ext = os.path.splitext(path)[1].lower()
    if ext not in EXTENSIONS:
        raise ValueError(f"cannot tell the format of {path}: give fmt=")
    if EXTENSIONS[ext]:
        return EXTENSIONS[ext]
    with open(path, "rb") as f:
        head = f.read(8)
    return "gr" if head == (1).to_bytes(8, "little") else "dimacs"


def read_graph(path, fmt=None, undirected=False, transpose=False):
    """Read a graph file into a CSR (wt is None for unweighted formats).

    undirected: add the reverse of every arc (the reference's
    directed = false); for METIS, each edge is listed at both ends and the
    header counts it once. transpose: the CSR of the transposed graph
    (parseCOO_transpose, parse.cpp:692-797, which the Pannotia SSSP reads)."""
    fmt = fmt or detect_format(path)
    c = L.Csr()
    flags = (L.READ_UNDIRECTED if undirected else 0) | (L.READ_TRANSPOSE if transpose else 0)
    _check(L.lib().wgm_read_graph(os.fsencode(path), FORMATS[fmt], flags, ctypes.byref(c)))
    return _take(c)


def write_graph(path, g, fmt, undirected=False):
    """Write a CSR as "gr", "dimacs" or "metis" (undirected: the METIS
    header counts each edge once)."""
    _check(L.lib().wgm_write_graph(os.fsencode(path), FORMATS[fmt], L.READ_UNDIRECTED if undirected else 0,
                                   ctypes.byref(g.as_c())))


def write_solution(path, values, style="colon"):
    v = np.ascontiguousarray(values, dtype=np.int32)
    _check(L.lib().wgm_write_solution(os.fsencode(path), STYLES[style], v.ctypes.data, v.size))


def check_solution(path, values, style="colon"):
    """diff_solution_file_int: (True, -1) if the file's first n values are
    `values`, else (False, first vertex that differs or is missing)."""
    v = np.ascontiguousarray(values, dtype=np.int32)
    bad = ctypes.c_int32(-1)
    rc = L.lib().wgm_check_solution(os.fsencode(path), STYLES[style], v.ctypes.data, v.size, ctypes.byref(bad))
    if rc in (L.WGM_ERR_MISMATCH, L.WGM_ERR_FORMAT):
        return False, bad.value
    _check(rc)
    return True, -1


def write_solution_f32(path, values):
    v = np.ascontiguousarray(values, dtype=np.float32)
    _check(L.lib().wgm_write_solution_f32(os.fsencode(path), v.ctypes.data, v.size))


def check_solution_f32(path, values, tol=8.0):
    """BC's check: every |file - value| <= tol (bc.h:194 uses 8)."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    bad = ctypes.c_int32(-1)
    rc = L.lib().wgm_check_solution_f32(os.fsencode(path), v.ctypes.data, v.size, tol, ctypes.byref(bad))
    if rc in (L.WGM_ERR_MISMATCH, L.WGM_ERR_FORMAT):
        return False, bad.value
    _check(rc)
    return True, -1
