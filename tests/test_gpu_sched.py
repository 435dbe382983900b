"""GPU: the resizing scheduler (mega-kernel). Persistent BFS/SSSP results
must be bit-exact with and without non-persistent tasks forcing workgroups
out (offer_kill) and back (offer_fork); every non-persistent task checks its
own result (djb2 diagonal hash / minimum / FP32 checksum)."""
import os

import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import graphs as G
from gpu_workgroup_managment_amd import scheduler as S

pytestmark = pytest.mark.gpu
os.environ.setdefault("WGM_HOST_TIMEOUT_S", "30")


@pytest.mark.parametrize("wg,wgs", [(256, 16), (1024, 4), (128, 8), (1024, 1)])
@pytest.mark.parametrize("kind", [S.NP_MATMULT_I32, S.NP_MATMUL_F32, S.NP_REDUCE_MIN])
def test_np_solo(rt, kind, wg, wgs):
    # FP32: MFMA path (wg % 256 == 0) and strided path (wg 128), position-weighted checksum
    with S.Mega(rt, np_kind=kind) as m:
        us, errs = m.run_solo(wg_size=wg, num_wgs=wgs, iterations=3)
    print(kind, wg, wgs, us, errs)
    assert errs == 0


def test_mega_launch_quit(rt):
    with S.Mega(rt, np_kind=S.NP_REDUCE_MIN) as m:
        p = m.launch(wg_size=256, num_wgs=8)
        m.quit()
    assert p == 7


@pytest.mark.parametrize("kind", [S.NP_MATMULT_I32, S.NP_MATMUL_F32, S.NP_REDUCE_MIN])
def test_mega_tasks_only(rt, kind):
    with S.Mega(rt, np_kind=kind) as m:
        p = m.launch(wg_size=256, num_wgs=16)
        oks = [m.send_task(g)["ok"] for g in (1, 3, p)]
        m.quit()
        st = m.stats()
    print(st)
    assert all(oks)
    assert st["tasks"] == 3 and st["errors"] == 0


@pytest.fixture(scope="module")
def er_graph(rt):
    g = G.add_weights(G.erdos_renyi(20000, 80000, seed=5), 1, 255, 3)
    return g, rt.graph(g)


@pytest.mark.parametrize("kind", [S.PERSIST_BFS, S.PERSIST_SSSP])
def test_mega_persistent_only(rt, er_graph, kind):
    g, dg = er_graph
    with S.Mega(rt, np_kind=S.NP_REDUCE_MIN) as m:
        p = m.launch(wg_size=256, num_wgs=12)
        m.send_persistent(dg, kind, 0, p)
        m.wait_persistent(30)
        m.quit()
        dist, levels = m.result(g.n)
    ref = oracle.bfs(g.row_ptr, g.col, 0)[0] if kind == S.PERSIST_BFS else oracle.sssp(g.row_ptr, g.col, g.wt, 0)
    assert np.array_equal(dist, ref)


@pytest.mark.parametrize("query", [False, True])
@pytest.mark.parametrize("kind", [S.PERSIST_BFS, S.PERSIST_SSSP])
def test_mega_persistent_with_resizing(rt, kind, query):
    # a high-diameter grid keeps the persistent task in many barriers
    side = 160
    g = G.grid(side, side)
    G.add_weights(g, 1, 255, 3)
    dg = rt.graph(g)
    with S.Mega(rt, np_kind=S.NP_MATMUL_F32) as m:
        p = m.launch(wg_size=256, num_wgs=16, query_barrier=query)
        m.send_persistent(dg, kind, 0, p)
        n_tasks = 0
        # task sizes cycle 4..11: some need kills, some fit in free workgroups
        while m.persistent_running() and n_tasks < 40:
            m.send_task(min(4 + n_tasks % 8, p - 1))
            n_tasks += 1
        m.wait_persistent(60)
        m.quit()
        st = m.stats()
        dist, levels = m.result(g.n)
    print(n_tasks, st)
    ref = oracle.bfs(g.row_ptr, g.col, 0)[0] if kind == S.PERSIST_BFS else oracle.sssp(g.row_ptr, g.col, g.wt, 0)
    assert np.array_equal(dist, ref), f"{int((dist != ref).sum())} differ"
    assert st["errors"] == 0


def test_run_merged_driver(rt):
    g = G.grid(128, 128)
    dg = rt.graph(g)
    with S.Mega(rt, np_kind=S.NP_MATMULT_I32) as m:
        st = S.run_merged(m, dg, S.PERSIST_BFS, 0, frequency_ms=0.2, np_ratio=4, wg_size=256, num_wgs=16)
        dist, _ = m.result(g.n)
    print(st)
    assert st["np_errors"] == 0
    r, c = np.divmod(np.arange(g.n), 128)
    assert np.array_equal(dist, (r + c).astype(np.uint32))


def _topo_ref(g, kind, iterations=50):
    if kind == S.PERSIST_COLOR:
        return oracle.color(g.row_ptr, g.col, g.prio)
    if kind == S.PERSIST_MIS:
        return oracle.mis(g.row_ptr, g.col, g.prio)
    return oracle.pagerank_f32(g.row_ptr, g.col, np.float32(0.85), iterations), iterations


@pytest.mark.parametrize("query", [False, True])
@pytest.mark.parametrize("kind", [S.PERSIST_COLOR, S.PERSIST_MIS, S.PERSIST_PAGERANK])
def test_mega_topology_tasks_with_resizing(rt, kind, query):
    """Colour, MIS and PageRank as the mega-kernel's persistent task while
    non-persistent tasks take workgroups away (offer_kill / the query
    barrier's shrink) and give them back (offer_fork) at every round's or
    iteration's resizing barriers: colours and sets bit-exact, ranks bitwise
    equal to the FP32 oracle, the round count equal to the oracle's."""
    # sized so that the task spans many milliseconds on 15 workgroups
    side, iters = (256, 400) if kind == S.PERSIST_PAGERANK else (768, 50)
    g = G.add_priorities(G.grid(side, side))
    ref, rounds = _topo_ref(g, kind, iters)
    dg = rt.graph(g)
    with S.Mega(rt, np_kind=S.NP_REDUCE_MIN) as m:
        p = m.launch(wg_size=256, num_wgs=16, query_barrier=query)
        m.send_persistent(dg, kind, 0, p, iterations=iters)
        n_tasks = 0
        # every workgroup holds the persistent task, so each task is served by
        # kills; sizes cycle 4..11
        while m.persistent_running() and n_tasks < 400:
            m.send_task(min(4 + n_tasks % 8, p - 1))
            n_tasks += 1
        m.wait_persistent(60)
        m.quit()
        st = m.stats()
        out, levels = m.result(g.n)
    print(kind, query, n_tasks, st)
    assert st["errors"] == 0 and n_tasks >= 2
    if kind == S.PERSIST_PAGERANK:
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), f"{int((out != ref).sum())} differ"
    else:
        assert np.array_equal(out, ref), f"{int((out != ref).sum())} differ"
    assert levels == rounds


@pytest.mark.parametrize("kind", [S.PERSIST_COLOR, S.PERSIST_MIS, S.PERSIST_PAGERANK])
def test_mega_topology_tasks_alone(rt, kind):
    """The same tasks with every workgroup and no interference, on a graph
    with degree-0 and degree-1 vertices (the colour round's degree-1 rule)."""
    g = G.add_priorities(G.erdos_renyi(6000, 5000, seed=9))
    ref, rounds = _topo_ref(g, kind)
    dg = rt.graph(g)
    with S.Mega(rt, np_kind=S.NP_REDUCE_MIN) as m:
        p = m.launch(wg_size=1024, num_wgs=0)
        m.send_persistent(dg, kind, 0, p)
        m.wait_persistent(60)
        m.quit()
        out, levels = m.result(g.n)
    if kind == S.PERSIST_PAGERANK:
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    else:
        assert np.array_equal(out, ref)
    assert levels == rounds
