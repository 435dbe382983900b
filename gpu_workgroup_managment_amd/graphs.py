"""Host CSR graphs and the seeded generators of libwgm.

Conventions: 0-based int32 row_ptr[n+1] / col[m] with neighbours ascending,
optional int32 weights aligned with col, uint32 tie-free priorities.
Seeds by convention: ER 1, RMAT 2, weights 3, priorities 4, batch 5,
serving arrivals 6.
"""
import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L

SEED_ER, SEED_RMAT, SEED_WEIGHTS, SEED_PRIO, SEED_BATCH, SEED_ARRIVALS = 1, 2, 3, 4, 5, 6


@dataclass
class CSR:
    row_ptr: np.ndarray
    col: np.ndarray
    wt: Optional[np.ndarray] = None
    prio: Optional[np.ndarray] = None

    @property
    def n(self):
        return len(self.row_ptr) - 1

    @property
    def m(self):
        return int(self.row_ptr[-1])

    def degrees(self):
        return np.diff(self.row_ptr)

    def as_c(self):
        """wgm_csr view of the numpy buffers (keeps references alive via self)."""
        c = L.Csr()
        c.n = self.n
        c.m = self.m
        c.row_ptr = L.i32ptr(self.row_ptr)
        c.col = L.i32ptr(self.col)
        c.wt = L.i32ptr(self.wt) if self.wt is not None else ctypes.POINTER(ctypes.c_int32)()
        return c

    def nbytes(self):
        b = self.row_ptr.nbytes + self.col.nbytes
        if self.wt is not None:
            b += self.wt.nbytes
        if self.prio is not None:
            b += self.prio.nbytes
        return b


def from_edges(n, src, dst, wt=None, symmetric=False):
    """CSR from edge lists (duplicates and self-loops dropped)."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    if symmetric:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
        if wt is not None:
            wt = np.concatenate([wt, wt])
    keep = src != dst
    src, dst = src[keep], dst[keep]
    if wt is not None:
        wt = np.asarray(wt, np.int32)[keep]
    key = src * n + dst
    key, idx = np.unique(key, return_index=True)
    s = (key // n).astype(np.int32)
    d = (key % n).astype(np.int32)
    row_ptr = np.zeros(n + 1, np.int32)
    np.add.at(row_ptr, s + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32)
    w = wt[idx].astype(np.int32) if wt is not None else None
    return CSR(row_ptr, d, w)


def _take(c):
    """Copy a library-allocated wgm_csr into numpy and free it."""
    try:
        n, m = c.n, c.m
        rp = np.ctypeslib.as_array(c.row_ptr, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(c.col, shape=(max(m, 1),))[:m].copy()
        wt = None
        if bool(c.wt):
            wt = np.ctypeslib.as_array(c.wt, shape=(max(m, 1),))[:m].copy()
        return CSR(rp, col, wt)
    finally:
        L.lib().wgm_csr_free(ctypes.byref(c))


def splitmix64(seed, count):
    st = ctypes.c_uint64(seed)
    return [L.lib().wgm_splitmix64_next(ctypes.byref(st)) for _ in range(count)]


def erdos_renyi(n, pairs, seed=SEED_ER):
    c = L.Csr()
    L.check(L.lib().wgm_gen_er(n, pairs, seed, ctypes.byref(c)))
    return _take(c)


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=SEED_RMAT, truncate_n=0):
    out = L.Csr()
    L.check(L.lib().wgm_gen_rmat(scale, edge_factor, a, b, c, seed, truncate_n, ctypes.byref(out)))
    return _take(out)


def grid(rows, cols):
    c = L.Csr()
    L.check(L.lib().wgm_gen_grid(rows, cols, ctypes.byref(c)))
    return _take(c)


def add_weights(g, lo=1, hi=255, seed=SEED_WEIGHTS):
    c = g.as_c()
    c.wt = ctypes.POINTER(ctypes.c_int32)()
    L.check(L.lib().wgm_gen_weights(ctypes.byref(c), lo, hi, seed))
    try:
        g.wt = np.ctypeslib.as_array(c.wt, shape=(max(g.m, 1),))[: g.m].copy()
    finally:
        L.lib().wgm_free(ctypes.cast(c.wt, ctypes.c_void_p))
    return g


def add_priorities(g, seed=SEED_PRIO):
    prio = np.empty(g.n, np.uint32)
    c = g.as_c()
    L.check(L.lib().wgm_gen_priorities(ctypes.byref(c), seed, L.ptr(prio)))
    g.prio = prio
    return g


def batch_spec(count=512, seed=SEED_BATCH, lo_exp=10, hi_exp=18):
    items = (L.BatchItem * count)()
    L.check(L.lib().wgm_batch_spec(count, seed, lo_exp, hi_exp, items))
    return items


def batch_generate(items, threads=8):
    count = len(items)
    graphs = (L.Csr * count)()
    L.check(L.lib().wgm_batch_generate(count, items, graphs, threads))
    return [_take(graphs[i]) for i in range(count)]


# the configured workloads
def er_config(n=1 << 16, avg_degree=8, seed=SEED_ER):
    """ER: n vertices, n*avg_degree/2 distinct undirected pairs."""
    return erdos_renyi(n, n * avg_degree // 2, seed)


def rmat_config(scale=20, edge_factor=16, seed=SEED_RMAT):
    g = rmat(scale, edge_factor, 0.57, 0.19, 0.19, seed)
    return add_weights(g, 1, 255, SEED_WEIGHTS)


def grid_config(side=2048, seed=SEED_PRIO):
    return add_priorities(grid(side, side), seed)
