"""GPU: the query server. Every query's distance array must equal the
oracle's (checked through the order-independent digest, which covers all n
entries, plus reached / scanned counts) whether it ran alone or in a mix
with non-persistent tasks that force workgroups out of the server mid-query
(preemption at level boundaries) and back; every task checks its own
result; admission keeps in-flight state within the HBM budget."""
import os

import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import graphs as G
from gpu_workgroup_managment_amd import serving as S

pytestmark = pytest.mark.gpu
os.environ.setdefault("WGM_HOST_TIMEOUT_S", "30")


def _pool():
    """Small graphs drawn as in the batched workload (RMAT / grid), weighted,
    plus two long-diameter grids that keep queries in many levels."""
    items = G.batch_spec(12, G.SEED_BATCH, 10, 13)
    gs = G.batch_generate(items, 8)
    for g in gs:
        G.add_weights(g, 1, 255, 3)
    for side in (96, 128):
        gs.append(G.add_weights(G.grid(side, side), 1, 255, 3))
    return gs


@pytest.fixture(scope="module")
def pool():
    gs = _pool()
    want = {}
    for i, g in enumerate(gs):
        b = oracle.bfs(g.row_ptr, g.col, 0)[0]
        d = oracle.sssp(g.row_ptr, g.col, g.wt, 0)
        for q, dist in ((S.QUERY_BFS, b), (S.QUERY_SSSP, d)):
            reached = dist != oracle.INF
            want[(i, q)] = (S.dist_digest(dist), int(reached.sum()), int(g.degrees()[reached].sum()))
    return gs, want


def _check(recs, want):
    assert recs
    for r in recs:
        dig, reached, scanned = want[(r["graph"], r["query"])]
        assert (r["digest"], r["reached"], r["edges_scanned"]) == (dig, reached, scanned), r


def test_serve_queries_alone(rt, pool):
    gs, want = pool
    with S.Server(rt, gs, S.config(wg_size=1024, max_inflight=32)) as srv:
        p = srv.start()
        print("participating", p, "graph bytes", srv.graph_bytes)
        recs = []
        for i in range(len(gs)):
            for q in (S.QUERY_BFS, S.QUERY_SSSP):
                srv.submit(i, q)
                recs += srv.drain()  # one at a time: each query alone on the server
        srv.stop()
    _check(recs, want)
    assert len(recs) == 2 * len(gs)


def test_serve_mix_with_tasks_preempts_and_matches(rt, pool):
    gs, want = pool
    big = [len(gs) - 1, len(gs) - 2]  # the grids
    # a small server (23 workgroups) so that queries occupy every serving workgroup
    with S.Server(rt, gs, S.config(wg_size=1024, num_wgs=24, max_inflight=64, reserve=4)) as srv:
        p = srv.start()
        assert p == 23
        recs, tasks = [], []
        for rep in range(3):
            for i in range(len(gs)):
                srv.submit(i, S.QUERY_SSSP if (i + rep) % 2 else S.QUERY_BFS)
            for i in big:
                srv.submit(i, S.QUERY_BFS)
                srv.submit(i, S.QUERY_SSSP)
            # tasks wider than the reserve take workgroups from running queries
            for k, kind in enumerate((S.NP_MATMUL_F32, S.NP_REDUCE_MIN, S.NP_MATMULT_I32)):
                tasks.append(srv.task(kind, 8 + 6 * k))
                recs += srv.poll()
            recs += srv.drain()
        srv.stop()
    print("tasks", tasks[:3], "preempted", sum(r["preempted"] for r in recs))
    assert all(t["ok"] for t in tasks)
    # the device-clock breakdown covers the execution seen by the host
    for t in tasks:
        parts = t["dev_assign_us"] + t["dev_start_us"] + t["dev_body_us"] + t["dev_retire_us"]
        assert t["dev_body_us"] > 0 and parts <= t["execution_us"] + 1.0, t
    _check(recs, want)
    assert len(recs) == 3 * (len(gs) + 4)
    assert sum(r["preempted"] for r in recs) > 0  # queries really yielded and resumed


def test_serve_admission_budget(rt, pool):
    gs, want = pool
    probe = S.Server(rt, gs, S.config(max_inflight=8))
    gbytes = probe.graph_bytes
    probe.close()
    nmax = max(g.n for g in gs)
    state = ((4 * nmax + 2 + 63) // 64) * 64 * 4
    budget = gbytes + 2 * state  # room for two of the largest requests at a time
    cfg = S.config(max_inflight=32, budget_bytes=budget)
    with S.Server(rt, gs, cfg) as srv:
        # the real allocation: graphs + request-state arena within the budget
        assert (srv.graph_bytes, srv.arena_bytes) == S.plan(gs, cfg)
        assert srv.graph_bytes + srv.arena_bytes <= budget
        srv.start()
        peak = 0
        for i in range(len(gs)):
            srv.submit(i, S.QUERY_BFS)
            peak = max(peak, srv.inflight_bytes())
        recs = srv.drain()
        srv.stop()
    assert peak <= budget - gbytes
    _check(recs, want)
    assert len(recs) == len(gs)


def test_serve_admission_budget_sssp(rt, pool):
    """An SSSP request also holds its near-far queues and flags: 7n + 4 words
    (rounded up to 64), so a budget sized for two of the largest SSSP
    requests admits at most two at a time and every result is still exact."""
    gs, want = pool
    probe = S.Server(rt, gs, S.config(max_inflight=8))
    gbytes = probe.graph_bytes
    probe.close()
    nmax = max(g.n for g in gs)
    state = ((7 * nmax + 4 + 63) // 64) * 64 * 4
    budget = gbytes + 2 * state
    cfg = S.config(max_inflight=32, budget_bytes=budget)
    with S.Server(rt, gs, cfg) as srv:
        assert srv.graph_bytes + srv.arena_bytes <= budget
        srv.start()
        peak = 0
        for i in range(len(gs)):
            srv.submit(i, S.QUERY_SSSP)
            peak = max(peak, srv.inflight_bytes())
        recs = srv.drain()
        srv.stop()
    assert peak <= budget - gbytes
    _check(recs, want)
    assert len(recs) == len(gs) and all(r["query"] == S.QUERY_SSSP for r in recs)


def test_serve_unweighted_graphs_bfs_only(rt):
    """A server whose graphs carry no weights sizes its request slots for
    BFS, serves BFS exactly and refuses SSSP."""
    gs = [G.grid(48, 40), G.grid(64, 64)]
    with S.Server(rt, gs, S.config(wg_size=256, max_inflight=8)) as srv:
        srv.start()
        for i in range(len(gs)):
            srv.submit(i, S.QUERY_BFS)
        with pytest.raises(Exception) as e:
            srv.submit(0, S.QUERY_SSSP)
        assert "weights" in str(e.value)
        recs = srv.drain()
        srv.stop()
    assert len(recs) == len(gs)
    for r in recs:
        g = gs[r["graph"]]
        b = oracle.bfs(g.row_ptr, g.col, 0)[0]
        reached = b != oracle.INF
        assert (r["digest"], r["reached"], r["edges_scanned"]) == (
            S.dist_digest(b), int(reached.sum()), int(g.degrees()[reached].sum()))


def test_serve_rejects_what_never_fits(rt, pool):
    gs, _ = pool
    probe = S.Server(rt, gs, S.config(max_inflight=8))
    gbytes = probe.graph_bytes
    probe.close()
    with S.Server(rt, gs, S.config(max_inflight=8, budget_bytes=gbytes + 4096)) as srv:
        with pytest.raises(Exception) as e:
            srv.submit(len(gs) - 1, S.QUERY_BFS)
        assert "BUDGET" in str(e.value)


def test_serve_open_loop_load(rt, pool):
    gs, want = pool
    with S.Server(rt, gs, S.config(wg_size=1024, max_inflight=128, reserve=16)) as srv:
        srv.start()
        rep = srv.run_load(rate_per_s=4000.0, requests=400, np_fraction=0.5, np_groups=16, seed=6)
        srv.stop()
    recs = rep.pop("records")
    print(rep)
    assert rep["task_errors"] == 0
    assert rep["queries"] + rep["tasks"] == 400
    _check(recs, want)
