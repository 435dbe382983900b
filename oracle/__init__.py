"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg. The product package never
imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

INF = 1_000_000_000

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        sig = {
            "or_bfs": (i32, [i32, _i32p, _i32p, i32, _u32p]),
            "or_sssp": (i32, [i32, _i32p, _i32p, _i32p, i32, _u32p]),
            "or_sssp_frontier": (i32, [i32, _i32p, _i32p, _i32p, i32, _u32p, ctypes.POINTER(i64)]),
            "or_color": (i32, [i32, _i32p, _i32p, _u32p, _i32p]),
            "or_mis": (i32, [i32, _i32p, _i32p, _u32p, _i32p]),
            "or_pagerank_f64": (i32, [i32, _i32p, _i32p, ctypes.c_double, i32, _f64p]),
            "or_pagerank_f32": (i32, [i32, _i32p, _i32p, ctypes.c_float, i32, _f32p]),
            "or_middle_square_fill": (i32, [_i32p, i32, i32, i32]),
            "or_matmul_i32": (None, [_i32p, _i32p, _i32p, i32, i32, i32]),
            "or_hash_diag": (i32, [_i32p, i32, i32]),
            "or_matmul_f64": (None, [_f32p, _f32p, _f64p, i32, i32, i32]),
            "or_reduce_min": (i32, [_i32p, i64]),
            "or_threads_pin": (i32, [i32]),
            "or_bfs_threads": (i32, [i32, _i32p, _i32p, i32, _u32p, i32]),
            "or_sssp_threads": (i32, [i32, _i32p, _i32p, _i32p, i32, _u32p, i32]),
            "or_fnv1a64": (ctypes.c_uint64, [ctypes.c_void_p, i64]),
            "or_reverse_csr": (i32, [i32, _i32p, _i32p, ctypes.c_void_p, _i32p, _i32p, ctypes.c_void_p]),
            "or_sssp_minplus": (i32, [i32, _i32p, _i32p, _i32p, i32, _i32p]),
            "or_bc": (i64, [i32, _i32p, _i32p, _i32p, _i32p, i32, _f64p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, name):
    if rc < 0:
        raise RuntimeError(f"{name} failed ({rc})")
    return rc


def bfs(row_ptr, col, src=0):
    n = len(row_ptr) - 1
    out = np.empty(n, np.uint32)
    levels = _check(lib().or_bfs(n, row_ptr, col, src, out), "or_bfs")
    return out, levels


def sssp(row_ptr, col, wt, src=0):
    n = len(row_ptr) - 1
    out = np.empty(n, np.uint32)
    _check(lib().or_sssp(n, row_ptr, col, wt, src, out), "or_sssp")
    return out


def sssp_frontier(row_ptr, col, wt, src=0):
    n = len(row_ptr) - 1
    out = np.empty(n, np.uint32)
    relax = ctypes.c_int64(0)
    levels = _check(lib().or_sssp_frontier(n, row_ptr, col, wt, src, out, ctypes.byref(relax)), "or_sssp_frontier")
    return out, levels, relax.value


def color(row_ptr, col, prio):
    n = len(row_ptr) - 1
    out = np.empty(n, np.int32)
    rounds = _check(lib().or_color(n, row_ptr, col, prio, out), "or_color")
    return out, rounds


def mis(row_ptr, col, prio):
    n = len(row_ptr) - 1
    out = np.empty(n, np.int32)
    rounds = _check(lib().or_mis(n, row_ptr, col, prio, out), "or_mis")
    return out, rounds


def pagerank_f64(row_ptr, col, damping=0.85, iters=50):
    n = len(row_ptr) - 1
    out = np.empty(n, np.float64)
    _check(lib().or_pagerank_f64(n, row_ptr, col, damping, iters, out), "or_pagerank_f64")
    return out


def pagerank_f32(row_ptr, col, damping=0.85, iters=50):
    n = len(row_ptr) - 1
    out = np.empty(n, np.float32)
    _check(lib().or_pagerank_f32(n, row_ptr, col, damping, iters, out), "or_pagerank_f32")
    return out


def middle_square_fill(rows, cols, seed):
    m = np.empty(rows * cols, np.int32)
    seed = lib().or_middle_square_fill(m, rows, cols, seed)
    return m.reshape(rows, cols), seed


def matmul_i32(a, b):
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    c = np.empty((a.shape[0], b.shape[1]), np.int32)
    lib().or_matmul_i32(a.ravel(), b.ravel(), c.ravel(), a.shape[0], a.shape[1], b.shape[1])
    return c


def hash_diag(m):
    m = np.ascontiguousarray(m, np.int32)
    return lib().or_hash_diag(m.ravel(), m.shape[0], m.shape[1])


def matmul_f64(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    c = np.empty((a.shape[0], b.shape[1]), np.float64)
    lib().or_matmul_f64(a.ravel(), b.ravel(), c.ravel(), a.shape[0], a.shape[1], b.shape[1])
    return c


def reduce_min(buf):
    buf = np.ascontiguousarray(buf, np.int32)
    return lib().or_reduce_min(buf, len(buf))


def threads_pin(on=True):
    """Pin the threaded restatement's threads, one per CPU of the caller's set."""
    lib().or_threads_pin(1 if on else 0)


def bfs_threads(row_ptr, col, src=0, threads=8):
    n = len(row_ptr) - 1
    out = np.empty(n, np.uint32)
    _check(lib().or_bfs_threads(n, row_ptr, col, src, out, threads), "or_bfs_threads")
    return out


def sssp_threads(row_ptr, col, wt, src=0, threads=8):
    n = len(row_ptr) - 1
    out = np.empty(n, np.uint32)
    _check(lib().or_sssp_threads(n, row_ptr, col, wt, src, out, threads), "or_sssp_threads")
    return out


def fnv1a64(data):
    """FNV-1a 64 of a bytes object or of a numpy array's raw (little-endian) bytes."""
    if isinstance(data, (bytes, bytearray)):
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        return lib().or_fnv1a64(ctypes.cast(buf, ctypes.c_void_p), len(data))
    a = np.ascontiguousarray(data)
    return lib().or_fnv1a64(a.ctypes.data_as(ctypes.c_void_p), a.nbytes)


BIG_NUM = 99999999


def _c(a, dtype=np.int32):
    return np.ascontiguousarray(a, dtype=dtype)


def reverse_csr(row_ptr, col, wt=None):
    """In-edge CSR (rrow, rcol, rwt): sources ascending within each row."""
    row_ptr, col = _c(row_ptr), _c(col)
    n, m = len(row_ptr) - 1, len(col)
    rrow = np.empty(n + 1, np.int32)
    rcol = np.empty(max(m, 1), np.int32)
    w = _c(wt) if wt is not None else None
    rwt = np.empty(max(m, 1), np.int32) if wt is not None else None
    _check(lib().or_reverse_csr(n, row_ptr, col, w.ctypes.data if w is not None else None, rrow, rcol,
                                rwt.ctypes.data if rwt is not None else None), "or_reverse_csr")
    return rrow, rcol[:m], (rwt[:m] if rwt is not None else None)


def sssp_minplus(row_ptr, col, wt, src=0):
    """Pannotia SSSP on the out-edge CSR (transposed here, as the reference
    reads its input): (dist int32 with BIG_NUM for unreached, sweeps)."""
    rrow, rcol, rwt = reverse_csr(row_ptr, col, wt)
    n = len(rrow) - 1
    out = np.empty(n, np.int32)
    sweeps = _check(lib().or_sssp_minplus(n, rrow, rcol, rwt if rwt is not None else np.zeros(1, np.int32), src, out),
                    "or_sssp_minplus")
    return out, sweeps


def bc(row_ptr, col, num_sources=None):
    """Pannotia BC in FP64 over sources [0, num_sources): (bc, total levels)."""
    row_ptr, col = _c(row_ptr), _c(col)
    n = len(row_ptr) - 1
    rrow, rcol, _ = reverse_csr(row_ptr, col)
    out = np.empty(n, np.float64)
    ns = n if num_sources is None else num_sources
    lv = lib().or_bc(n, row_ptr, col, rrow, rcol if len(rcol) else np.zeros(1, np.int32), ns, out)
    if lv < 0:
        raise RuntimeError(f"or_bc failed ({lv})")
    return out, int(lv)
