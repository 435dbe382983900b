"""GPU: the BFS entry point that allocates nothing (wgm_bfs_noalloc), called
through the C ABI on caller-owned device buffers, against wgm_bfs on the
same graphs: bit-exact levels, level count and scanned edges, on two calls
in a row (the second on a smaller graph, so stale queue contents from the
first would show), and its size limits."""
import ctypes

import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import _lib as L
from gpu_workgroup_managment_amd import graphs as G

pytestmark = pytest.mark.gpu

MAX_N = 1 << 21
MAX_M = 1 << 26


def _noalloc(rt, dev, g, src=0, wg=256, num_wgs=0):
    rp = dev(g.row_ptr)
    col = dev(g.col if g.m else np.zeros(1, np.int32))
    dist = dev(np.full(g.n, 7, np.uint32))
    lc = L.Launch()
    lc.wg_size, lc.num_wgs = wg, num_wgs
    st = L.RunStats()
    try:
        rc = L.lib().wgm_bfs_noalloc(rt.handle, rp.ptr, col.ptr, g.n, g.m, src, ctypes.byref(lc), dist.ptr,
                                     ctypes.byref(st))
        rt._check(rc)
        return dist.get(), st.as_dict()
    finally:
        for b in (rp, col, dist):
            b.free()


@pytest.mark.parametrize("wg", [256, 1024])
def test_noalloc_bfs_matches_wgm_bfs(rt, dev, wg):
    big = G.rmat(16, 16, seed=7)  # 65,536 vertices, ~1M edges, hubs above the heavy threshold
    small = G.erdos_renyi(3000, 9000, seed=3)
    assert big.degrees().max() > 4 * wg
    for g, nw in ((big, 0), (small, 8)):
        got, st = _noalloc(rt, dev, g, 0, wg, nw)
        want, wst = rt.graph(g).bfs(0, wg_size=wg, num_wgs=nw)
        assert np.array_equal(got, want), f"{int((got != want).sum())} differ"
        assert st["iterations"] == wst["iterations"]
        assert st["edges_scanned"] == wst["edges_scanned"]
        ref, levels = oracle.bfs(g.row_ptr, g.col, 0)
        assert np.array_equal(got, ref) and st["iterations"] == levels


def test_noalloc_bfs_rejects_what_exceeds_its_scratch(rt, dev):
    tiny = dev(np.zeros(4, np.int32))
    try:
        lc = L.Launch()
        lc.wg_size, lc.num_wgs = 256, 8
        st = L.RunStats()
        f = L.lib().wgm_bfs_noalloc
        assert f(rt.handle, tiny.ptr, tiny.ptr, MAX_N + 1, 0, 0, ctypes.byref(lc), tiny.ptr, ctypes.byref(st)) == -1
        assert f(rt.handle, tiny.ptr, tiny.ptr, 3, MAX_M + 1, 0, ctypes.byref(lc), tiny.ptr, ctypes.byref(st)) == -1
        assert f(rt.handle, tiny.ptr, tiny.ptr, 3, 0, 3, ctypes.byref(lc), tiny.ptr, ctypes.byref(st)) == -1
    finally:
        tiny.free()
