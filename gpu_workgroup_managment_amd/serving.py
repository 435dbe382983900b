"""Python mirror of the query server (include/wgm.h, wgm_serve_*): resident
graphs, BFS / SSSP requests admitted under an HBM budget and run
concurrently by the resident kernel's persistent task, non-persistent
tasks through the resizing scheduler, and the open-loop Poisson driver."""
import ctypes

import numpy as np

from . import _lib as L

NP_MATMULT_I32, NP_MATMUL_F32, NP_REDUCE_MIN = 0, 1, 2
QUERY_BFS, QUERY_SSSP = 0, 1


def dist_digest(dist):
    """wgm_dist_digest of a uint32 distance array (host code)."""
    d = np.ascontiguousarray(dist, np.uint32)
    return int(L.lib().wgm_dist_digest(L.ptr(d), len(d)))


def config(wg_size=1024, num_wgs=0, max_inflight=256, reserve=16, budget_bytes=64 << 30, np_dim=256,
           np_length=1 << 20):
    c = L.ServeConfig()
    c.wg_size, c.num_wgs, c.max_inflight, c.reserve, c.budget_bytes = wg_size, num_wgs, max_inflight, reserve, budget_bytes
    c.np.kind, c.np.dim, c.np.length, c.np.seed = NP_MATMUL_F32, np_dim, np_length, 0
    return c


def plan(graphs, cfg=None):
    """(graph_bytes, arena_bytes) of a server on `graphs` (host only): the
    request-state arena never takes more than the budget the graphs leave."""
    arr = (L.Csr * len(graphs))()
    for i, g in enumerate(graphs):
        arr[i] = g.as_c()
    cfg = cfg if cfg is not None else config()
    gb, ab = ctypes.c_int64(), ctypes.c_int64()
    L.check(L.lib().wgm_serve_plan(len(graphs), arr, ctypes.byref(cfg), ctypes.byref(gb), ctypes.byref(ab)))
    return gb.value, ab.value


class Server:
    """wgm_serve: create uploads the graphs; start launches the resident kernel."""

    def __init__(self, rt, graphs, cfg=None):
        self.rt = rt
        self.graphs = list(graphs)
        self._arr = (L.Csr * len(self.graphs))()
        for i, g in enumerate(self.graphs):
            self._arr[i] = g.as_c()
        self.cfg = cfg if cfg is not None else config()
        self._h = ctypes.c_void_p()
        rt._check(L.lib().wgm_serve_create(rt.handle, len(self.graphs), self._arr, ctypes.byref(self.cfg),
                                           ctypes.byref(self._h)))
        self.participating = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def close(self):
        if self._h:
            L.lib().wgm_serve_destroy(self._h)  # stops (drains, quits) if running
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def graph_bytes(self):
        return int(L.lib().wgm_serve_graph_bytes(self._h))

    @property
    def arena_bytes(self):
        return int(L.lib().wgm_serve_arena_bytes(self._h))

    def start(self):
        p = ctypes.c_int32()
        self.rt._check(L.lib().wgm_serve_start(self._h, ctypes.byref(p)))
        self.participating = p.value
        return p.value

    def submit(self, graph, query=QUERY_BFS, src=0):
        i = ctypes.c_int64()
        self.rt._check(L.lib().wgm_serve_submit(self._h, graph, query, src, ctypes.byref(i)))
        return i.value

    def poll(self, max_records=256):
        buf = (L.ServeDone * max_records)()
        c = ctypes.c_int32()
        self.rt._check(L.lib().wgm_serve_poll(self._h, buf, max_records, ctypes.byref(c)))
        return [buf[i].as_dict() for i in range(c.value)]

    def outstanding(self):
        return int(L.lib().wgm_serve_outstanding(self._h))

    def inflight_bytes(self):
        return int(L.lib().wgm_serve_inflight_bytes(self._h))

    def drain(self, timeout_s=60.0):
        import time
        out = []
        t0 = time.perf_counter()
        while self.outstanding() > 0:
            out += self.poll()
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError(f"{self.outstanding()} requests outstanding")
        return out

    def task(self, kind, groups):
        t = L.TaskTimes()
        rc = L.lib().wgm_serve_task(self._h, kind, groups, ctypes.byref(t))
        if rc not in (L.WGM_OK, -6):
            self.rt._check(rc)
        return t.as_dict()

    def stop(self):
        self.rt._check(L.lib().wgm_serve_stop(self._h))

    def run_load(self, rate_per_s, requests, np_fraction=0.5, np_groups=16, seed=6, keep_records=True):
        ld = L.ServeLoad()
        ld.rate_per_s, ld.requests, ld.np_groups, ld.np_fraction, ld.seed = rate_per_s, requests, np_groups, np_fraction, seed
        rep = L.ServeReport()
        recs = (L.ServeDone * requests)() if keep_records else None
        n = ctypes.c_int32()
        self.rt._check(L.lib().wgm_serve_run_load(self._h, ctypes.byref(ld), ctypes.byref(rep), recs, ctypes.byref(n)))
        out = rep.as_dict()
        if keep_records:
            out["records"] = [recs[i].as_dict() for i in range(n.value)]
        return out
