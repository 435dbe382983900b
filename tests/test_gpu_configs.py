"""GPU parity at the configured sizes: every workload the benchmark quotes
runs here on its full input and is checked against the oracle.

- BFS on ER 2^16 (average degree 8), levels bit-exact;
- Lonestar SSSP and BFS on RMAT scale 20 (edge factor 16, w in [1, 255]) vs
  Dijkstra / serial BFS, and the Pannotia min-plus SSSP vs its Jacobi oracle;
- colour, MIS and PageRank on the 2048 x 2048 grid (colours and sets
  bit-exact, PageRank bitwise equal to the FP32 oracle with its L1 error
  against the FP64 oracle asserted);
- the 512-graph batch, every query bit-exact, in one persistent launch
  (hipGraph) and one launch per graph;
- serving on the 512-graph pool at offered load 0.5, every finished query's
  distance digest, reach and scanned edges equal to the oracle's;
- the non-persistent tasks: the library's inputs, outputs, published
  summaries and expected values against the oracle and the matmult fixture.

Every input is first pinned to the generator fixtures (FNV-1a digests).
"""
import json
import os

import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import batched as B
from gpu_workgroup_managment_amd import graphs as G
from gpu_workgroup_managment_amd import scheduler as S
from gpu_workgroup_managment_amd import serving as SV

pytestmark = pytest.mark.gpu
os.environ.setdefault("WGM_HOST_TIMEOUT_S", "30")

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "fixtures.json")) as _f:
    FIX = json.load(_f)
THREADS = min(16, os.cpu_count() or 1)
# the benchmark's launch shape for the 512-graph batch
BATCH_SHAPE = dict(wg_size=1024, wgs_per_cu=2, chunk=256, team_edges=1 << 20, max_team=16, delta=512,
                   delta_high_degree=12, sssp_team_edges=786432)


def _first_diff(a, b):
    idx = np.flatnonzero(a != b)
    return f"{len(idx)} differ, first at {idx[0]}: got {a[idx[0]]} want {b[idx[0]]}" if len(idx) else ""


def _pin(g, name):
    want = FIX["graphs"][name]
    got = {"n": g.n, "m": g.m, "row_ptr": oracle.fnv1a64(g.row_ptr), "col": oracle.fnv1a64(g.col)}
    if "wt" in want:
        got["wt"] = oracle.fnv1a64(g.wt)
    if "prio" in want:
        got["prio"] = oracle.fnv1a64(g.prio)
    assert got == want, f"{name}: generated input differs from its fixture"


def test_config_er_bfs(rt):
    g = G.er_config()
    _pin(g, "er_config")
    dg = rt.graph(g)
    ref, levels = oracle.bfs(g.row_ptr, g.col, 0)
    reached = ref != oracle.INF
    for wg, nw in ((1024, 0), (256, 0), (1024, 64)):
        dist, st = dg.bfs(0, wg_size=wg, num_wgs=nw)
        assert np.array_equal(dist, ref), _first_diff(dist, ref)
        assert st["iterations"] == levels
        assert st["edges_scanned"] == int(g.degrees()[reached].sum())
    dg.close()


@pytest.fixture(scope="module")
def rmat20(rt):
    g = G.rmat_config()
    _pin(g, "rmat_config")
    dg = rt.graph(g)
    yield g, dg
    dg.close()


def test_config_rmat20_sssp(rt, rmat20):
    g, dg = rmat20
    ref = oracle.sssp(g.row_ptr, g.col, g.wt, 0)  # Dijkstra, 64-bit sums
    reached = ref != oracle.INF
    for wg, nw, mode in ((1024, 64, 0), (1024, 0, 0), (1024, 64, 1), (256, 0, 1)):
        dist, st = dg.sssp(0, wg_size=wg, num_wgs=nw, mode=mode)  # mode 0: pull here; 1: push
        assert np.array_equal(dist, ref), _first_diff(dist, ref)
        assert st["edges_scanned"] == int(g.degrees()[reached].sum())


def test_config_rmat20_bfs(rt, rmat20):
    g, dg = rmat20
    ref, levels = oracle.bfs(g.row_ptr, g.col, 0)
    dist, st = dg.bfs(0, wg_size=1024, num_wgs=0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    assert st["iterations"] == levels
    assert st["edges_scanned"] == int(g.degrees()[ref != oracle.INF].sum())


def test_config_rmat20_minplus(rt, rmat20):
    g, dg = rmat20
    ref, _ = oracle.sssp_minplus(g.row_ptr, g.col, g.wt, 0)
    dist, st = dg.sssp_minplus(0, wg_size=1024, num_wgs=0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    # the same distances as Dijkstra, with BIG_NUM for unreached vertices
    dj = oracle.sssp(g.row_ptr, g.col, g.wt, 0).astype(np.int64)
    assert np.array_equal(np.where(dj == oracle.INF, oracle.BIG_NUM, dj), dist.astype(np.int64))


@pytest.fixture(scope="module")
def grid2048(rt):
    g = G.grid_config()
    _pin(g, "grid_config")
    dg = rt.graph(g)
    yield g, dg
    dg.close()


def test_config_grid_color(rt, grid2048):
    g, dg = grid2048
    ref, rounds = oracle.color(g.row_ptr, g.col, g.prio)
    col, st = dg.color(wg_size=1024, num_wgs=0)
    assert np.array_equal(col, ref), _first_diff(col, ref)
    assert st["iterations"] == rounds


def test_config_grid_mis(rt, grid2048):
    g, dg = grid2048
    ref, rounds = oracle.mis(g.row_ptr, g.col, g.prio)
    s, st = dg.mis(wg_size=1024, num_wgs=0)
    assert np.array_equal(s, ref), _first_diff(s, ref)
    assert st["iterations"] == rounds


def test_config_grid_pagerank(rt, grid2048):
    g, dg = grid2048
    ref32 = oracle.pagerank_f32(g.row_ptr, g.col, np.float32(0.85), 50)
    ref64 = oracle.pagerank_f64(g.row_ptr, g.col, 0.85, 50)
    r, st = dg.pagerank(0.85, 50, wg_size=1024, num_wgs=0)
    assert np.array_equal(r.view(np.uint32), ref32.view(np.uint32)), _first_diff(r, ref32)
    l1 = float(np.abs(r.astype(np.float64) - ref64).sum())
    print(f"PageRank 2048^2, 50 iterations: L1 vs FP64 = {l1:.3e}")
    # FP32 rounding of 4M terms of ~1/n each, over 50 iterations
    assert l1 < 1e-5


@pytest.fixture(scope="module")
def batch512():
    items, graphs, queries = B.workload(512, threads=THREADS)
    raw = bytes(memoryview(items).cast("B"))
    assert oracle.fnv1a64(raw) == FIX["batch"]["items"]
    assert sum(g.n for g in graphs) == FIX["batch"]["sum_n"]
    refs = []
    for g, q in zip(graphs, queries):
        if q == SV.QUERY_BFS:
            refs.append(oracle.bfs_threads(g.row_ptr, g.col, 0, THREADS))
        else:
            refs.append(oracle.sssp_threads(g.row_ptr, g.col, g.wt, 0, THREADS))
    return graphs, queries, refs


def _check_batch(b, graphs, refs):
    bad = []
    for i, g in enumerate(graphs):
        dist, _, edges = b.result(i)
        if not np.array_equal(dist, refs[i]) or edges != int(g.degrees()[refs[i] != oracle.INF].sum()):
            bad.append(i)
    assert not bad, f"{len(bad)} of {len(graphs)} queries differ, first {bad[:8]}"


def test_config_batched512(rt, batch512):
    graphs, queries, refs = batch512
    b = B.Batch(rt, graphs, queries)
    cfg = B.config(use_graph=True, **BATCH_SHAPE)
    for rep in range(2):  # the second run replays the captured graph with the job table reordered by duration
        st = b.run(cfg)
        assert st["launches"] == 1
        _check_batch(b, graphs, refs)
    st = b.run(B.config(use_graph=True, far_flags_lds=1, **BATCH_SHAPE))
    _check_batch(b, graphs, refs)
    st = b.run(B.config(use_graph=True, far_flags_lds=1, low_ell=1, **BATCH_SHAPE))
    _check_batch(b, graphs, refs)
    st = b.run(B.config(use_graph=True, far_flags_lds=1, low_ell=2, **BATCH_SHAPE))
    _check_batch(b, graphs, refs)
    st = b.run_relaunch(B.config(**dict(BATCH_SHAPE, team_edges=1 << 16)))
    assert st["launches"] == len(graphs)
    _check_batch(b, graphs, refs)
    st = b.run_standalone(wg_size=1024)
    assert st["launches"] == len(graphs)
    _check_batch(b, graphs, refs)
    b.close()


def test_config_serving_on_batch_pool(rt, batch512):
    """Open-loop serving at offered load 0.5 on the 512-graph pool (the
    benchmark's serving leg): every query that finishes in the mix with the
    non-persistent tasks carries the oracle's distances."""
    graphs, queries, refs = batch512
    want = {}
    for i, g in enumerate(graphs):
        reached = refs[i] != oracle.INF
        want[i] = (SV.dist_digest(refs[i]), int(reached.sum()), int(g.degrees()[reached].sum()))
    np_groups = 16
    with SV.Server(rt, graphs, SV.config(wg_size=1024, max_inflight=256, reserve=np_groups)) as srv:
        srv.start()
        solo = []
        for i in range(20):
            t = srv.task(SV.NP_MATMUL_F32 if i % 2 == 0 else SV.NP_REDUCE_MIN, np_groups)
            assert t["ok"]
            solo.append(t["response_us"] + t["execution_us"])
        s_np = sorted(solo)[len(solo) // 2] * 1e-6
        rate = 0.5 / (0.5 * s_np)
        n = max(400, int(rate * 0.4))
        rep = srv.run_load(rate_per_s=rate, requests=n, np_fraction=0.5, np_groups=np_groups, seed=6)
        srv.stop()
    recs = rep.pop("records")
    print({k: v for k, v in rep.items()})
    assert rep["task_errors"] == 0 and rep["tasks"] > 0
    assert len(recs) == rep["queries"] > 0
    for r in recs:
        assert r["query"] == queries[r["graph"]]
        assert (r["digest"], r["reached"], r["edges_scanned"]) == want[r["graph"]], r


@pytest.mark.parametrize("kind", [S.NP_MATMULT_I32, S.NP_MATMUL_F32, S.NP_REDUCE_MIN])
def test_np_tasks_vs_oracle(rt, kind):
    """The non-persistent tasks' inputs, outputs and summaries against the
    oracle (a39/a40 of the reference: matmult's djb2 diagonal hash and the
    reduction minimum), run alone and inside the resident mega-kernel."""
    fx = FIX["matmult"]
    with S.Mega(rt, np_kind=kind) as m:
        us, errs = m.run_solo(wg_size=256, num_wgs=16, iterations=1)
        assert errs == 0
        solo_val, solo_chk = m.np_result()
        p = m.launch(wg_size=1024, num_wgs=24)
        t = m.send_task(p)
        m.quit()
        assert t["ok"]
        val, chk = m.np_result()
        exp_val, exp_chk, tol = m.np_expected(kind)
        if kind == S.NP_MATMULT_I32:
            a = m.np_copy(kind, S.NP_IN_A)
            b = m.np_copy(kind, S.NP_IN_B)
            c = m.np_copy(kind, S.NP_OUT_C)
            ra, seed = oracle.middle_square_fill(fx["dim"], fx["dim"], fx["seed"])
            rb, _ = oracle.middle_square_fill(fx["dim"], fx["dim"], seed)
            assert np.array_equal(a, ra) and np.array_equal(b, rb)
            rc = oracle.matmul_i32(a, b)
            assert np.array_equal(c, rc), _first_diff(c, rc)
            assert solo_val == val == exp_val == oracle.hash_diag(rc) == fx["hash"]
        elif kind == S.NP_REDUCE_MIN:
            buf = m.np_copy(kind, S.NP_IN_A)
            assert solo_val == val == exp_val == oracle.reduce_min(buf)
        else:
            a = m.np_copy(kind, S.NP_IN_A)
            b = m.np_copy(kind, S.NP_IN_B)
            c = m.np_copy(kind, S.NP_OUT_C)
            ref = oracle.matmul_f64(a, b)
            # the FP32 dot-product bound, element by element: K u sum_k |a_ik b_kj|
            bound = a.shape[1] * 2.0 ** -24 * (np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64))
            assert np.all(np.abs(c.astype(np.float64) - ref) <= bound)
            u = (np.arange(a.shape[0]) % 7 + 1).astype(np.float64)
            v = (np.arange(a.shape[0]) % 5 + 1).astype(np.float64)
            want = float(u @ ref @ v)
            assert solo_val == val == exp_val == 1
            assert abs(exp_chk - want) <= 1e-9 * abs(want)
            assert abs(chk - want) <= tol and abs(solo_chk - want) <= tol
