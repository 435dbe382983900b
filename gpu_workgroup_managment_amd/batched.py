"""Batched graph queries: many BFS / SSSP queries served by one persistent
launch pulling from a device work queue (libwgm wgm_batch_*), the per-graph
relaunch baseline, hipGraph replay and the autotuner."""
import ctypes

import numpy as np

from . import _lib as L
from . import graphs as G


def config(wg_size=256, wgs_per_cu=4, chunk=0, team_edges=1 << 18, max_team=16, use_graph=False, delta=0,
           sssp_team_edges=0, delta_high_degree=0, far_flags_lds=0, low_ell=0):
    c = L.BatchConfig()
    c.wg_size, c.wgs_per_cu, c.chunk = wg_size, wgs_per_cu, chunk
    c.team_edges, c.max_team, c.use_graph = team_edges, max_team, 1 if use_graph else 0
    c.delta, c.sssp_team_edges, c.delta_high_degree = delta, sssp_team_edges, delta_high_degree
    c.far_flags_lds = 1 if far_flags_lds else 0
    if low_ell not in (0, 1, 2, False, True):
        raise ValueError("low_ell is 0, 1 or 2")
    c.low_ell = int(low_ell)
    return c


class Batch:
    def __init__(self, rt, graphs, queries, sources=None):
        self.rt = rt
        self.graphs = list(graphs)
        self.queries = np.ascontiguousarray(queries, np.int32)
        self.sources = None if sources is None else np.ascontiguousarray(sources, np.int32)
        arr = (L.Csr * len(self.graphs))()
        for i, g in enumerate(self.graphs):
            arr[i] = g.as_c()
        self._h = ctypes.c_void_p()
        rt._check(L.lib().wgm_batch_create(rt.handle, len(self.graphs), arr, L.ptr(self.queries),
                                           L.ptr(self.sources), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            L.lib().wgm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nbytes(self):
        return L.lib().wgm_batch_bytes(self._h)

    def run(self, cfg=None, **kw):
        cfg = cfg if cfg is not None else config(**kw)
        st = L.BatchStats()
        self.rt._check(L.lib().wgm_batch_run(self.rt.handle, self._h, ctypes.byref(cfg), ctypes.byref(st)))
        return st.as_dict()

    def run_relaunch(self, cfg=None, **kw):
        cfg = cfg if cfg is not None else config(**kw)
        st = L.BatchStats()
        self.rt._check(L.lib().wgm_batch_run_relaunch(self.rt.handle, self._h, ctypes.byref(cfg), ctypes.byref(st)))
        return st.as_dict()

    def run_standalone(self, wg_size=1024, num_wgs=0):
        """Baseline: each query on wgm_bfs / wgm_sssp's standalone kernel, one launch per graph."""
        lc = L.Launch()
        lc.wg_size, lc.num_wgs = wg_size, num_wgs
        st = L.BatchStats()
        self.rt._check(L.lib().wgm_batch_run_standalone(self.rt.handle, self._h, ctypes.byref(lc), ctypes.byref(st)))
        return st.as_dict()

    def result(self, i):
        dist = np.empty(self.graphs[i].n, np.uint32)
        it, edges = ctypes.c_int32(), ctypes.c_int64()
        self.rt._check(L.lib().wgm_batch_result(self._h, i, L.ptr(dist), ctypes.byref(it), ctypes.byref(edges)))
        return dist, it.value, edges.value

    def job_times(self):
        """(start_us, end_us) per query of the last run, from the first query's start."""
        t0, t1 = np.empty(len(self.graphs)), np.empty(len(self.graphs))
        self.rt._check(L.lib().wgm_batch_job_times(self._h, L.ptr(t0), L.ptr(t1)))
        return t0, t1

    def autotune(self, reps=1):
        best, st = L.BatchConfig(), L.BatchStats()
        self.rt._check(L.lib().wgm_batch_autotune(self.rt.handle, self._h, reps, ctypes.byref(best), ctypes.byref(st)))
        return best, st.as_dict()


def workload(count=512, seed=G.SEED_BATCH, lo_exp=10, hi_exp=18, threads=8):
    """The batched workload: graphs, queries (BFS/SSSP) and spec items."""
    items = G.batch_spec(count, seed, lo_exp, hi_exp)
    graphs = G.batch_generate(items, threads)
    queries = np.array([it.query for it in items], np.int32)
    return items, graphs, queries
