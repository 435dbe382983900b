"""GPU: the host skeleton drives caller-written plugins through the vtables.

The persistent app is a Lonestar-style BFS written against the scheduler
stubs (tests/kernel_merge/bfs_level.hip), turned by kernel_merge into its
standalone form and, merged with the reference's matmult task written
against the stubs (tests/kernel_merge/matmult.hip), into a mega-kernel. The
plugins (tests/kernel_merge/plugins.hip) read the graph from a Galois .gr
file, check the levels against a "%d:%d" solution file written from the
oracle, and check every matmult task's djb2 diagonal hash against the host
product; wgm_run_persistent / wgm_run_merged call them in the reference's
order. Bit-exact, with and without the query barrier, under the task
stream's kills and forks."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import _lib as L
from gpu_workgroup_managment_amd import formats as F
from gpu_workgroup_managment_amd import graphs as G
from gpu_workgroup_managment_amd import skeleton as SK

pytestmark = pytest.mark.gpu
os.environ.setdefault("WGM_HOST_TIMEOUT_S", "30")

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def km():
    path = os.path.join(HERE, "kernel_merge", "libwgm_kmerge.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for f in ("km_mega_bfs", "km_standalone_bfs"):
        getattr(lib, f).restype = ctypes.c_void_p
    lib.km_bfs_app.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(L.PersistentApp)]
    lib.km_matmult_task.argtypes = [ctypes.c_int32, ctypes.POINTER(L.NpTask)]
    for f in ("km_bfs_checks", "km_bfs_free"):
        getattr(lib, f).argtypes = [ctypes.POINTER(L.PersistentApp)]
    for f in ("km_matmult_checks", "km_matmult_expected", "km_matmult_free"):
        getattr(lib, f).argtypes = [ctypes.POINTER(L.NpTask)]
    return lib


def _bfs_app(km, tmp_path, g, solution=None):
    gr = str(tmp_path / "graph.gr")
    sol = str(tmp_path / "graph.sol")
    F.write_graph(gr, g, "gr")
    ref, levels = oracle.bfs(g.row_ptr, g.col, 0)
    F.write_solution(sol, ref.astype(np.int32) if solution is None else solution, "colon")
    app = L.PersistentApp()
    assert km.km_bfs_app(gr.encode(), sol.encode(), ctypes.byref(app)) == 0
    return app, ref


@pytest.mark.parametrize("wg,wgs", [(256, 32), (1024, 0)])
def test_skeleton_run_persistent_bfs(rt, km, tmp_path, wg, wgs):
    g = G.rmat(14, 8, seed=4)
    app, _ = _bfs_app(km, tmp_path, g)
    rep = SK.run_persistent(rt, app, SK.config(threads_per_wg=wg, num_wgs=wgs, iterations=3))
    print(rep)
    assert rep["persistent_errors"] == 0 and rep["iterations"] == 3
    assert km.km_bfs_checks(ctypes.byref(app)) == 3
    assert 0 < rep["num_wgs"] <= rep["occupancy"] and (wgs == 0 or rep["num_wgs"] <= wgs)
    assert rep["min_ms"] > 0
    km.km_bfs_free(ctypes.byref(app))


def test_skeleton_check_catches_a_wrong_solution(rt, km, tmp_path):
    """The plugin's check really compares: one wrong level in the solution
    file fails every iteration."""
    g = G.rmat(12, 8, seed=4)
    ref, _ = oracle.bfs(g.row_ptr, g.col, 0)
    bad = ref.astype(np.int32)
    bad[np.flatnonzero(ref != oracle.INF)[-1]] += 1
    app, _ = _bfs_app(km, tmp_path, g, solution=bad)
    rep = SK.run_persistent(rt, app, SK.config(threads_per_wg=256, num_wgs=8, iterations=2))
    assert rep["persistent_errors"] == 2
    km.km_bfs_free(ctypes.byref(app))


@pytest.mark.parametrize("query", [False, True])
def test_skeleton_run_merged_bfs_with_matmult(rt, km, tmp_path, query):
    # a high-diameter graph keeps the persistent BFS in hundreds of resizing
    # barriers while the matmult tasks take workgroups away and back
    g = G.grid(200, 200)
    app, _ = _bfs_app(km, tmp_path, g)
    np_task = L.NpTask()
    assert km.km_matmult_task(256, ctypes.byref(np_task)) == 0
    cfg = SK.config(threads_per_wg=256, num_wgs=16, iterations=2, non_persistent_wgs=4, use_query_barrier=query,
                    max_tasks=400)
    rep = SK.run_merged(rt, app, np_task, km.km_mega_bfs(), cfg)
    print(rep)
    assert rep["persistent_errors"] == 0 and rep["np_errors"] == 0
    assert rep["tasks"] > 0 and km.km_matmult_checks(ctypes.byref(np_task)) == rep["tasks"]
    assert km.km_bfs_checks(ctypes.byref(app)) == 2
    assert rep["np_wgs"] == rep["num_wgs"] // 4
    with open(os.path.join(HERE, "golden", "fixtures.json")) as f:
        fx = json.load(f)["matmult"]
    assert fx["dim"] == 256 and fx["seed"] == 1234
    assert km.km_matmult_expected(ctypes.byref(np_task)) == fx["hash"]
    km.km_bfs_free(ctypes.byref(app))
    km.km_matmult_free(ctypes.byref(np_task))


def test_skeleton_standalone_bfs_is_the_kernel_merge_output(rt, km, tmp_path):
    """The standalone BFS the plugin runs is kernel_merge's emitted form, and
    the same kernel under wgm_launch_standalone gives the oracle's levels on
    an ER graph (a graph app through kernel_merge, bit-exact)."""
    g = G.erdos_renyi(20000, 80000, seed=6)
    app, ref = _bfs_app(km, tmp_path, g)
    assert ctypes.cast(km.km_standalone_bfs(), ctypes.c_void_p).value is not None
    rep = SK.run_persistent(rt, app, SK.config(threads_per_wg=512, num_wgs=0, iterations=1))
    assert rep["persistent_errors"] == 0
    km.km_bfs_free(ctypes.byref(app))
