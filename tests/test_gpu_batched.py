"""GPU parity of batched persistent BFS/SSSP: every query bit-exact vs the oracle."""
import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import batched as B
from gpu_workgroup_managment_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    items, graphs, queries = B.workload(count=48, seed=11, lo_exp=6, hi_exp=13)
    refs = []
    for g, q in zip(graphs, queries):
        refs.append(oracle.bfs(g.row_ptr, g.col, 0)[0] if q == L.QUERY_BFS else oracle.sssp(g.row_ptr, g.col, g.wt, 0))
    return graphs, queries, refs


def _check_all(batch, graphs, refs):
    for i, g in enumerate(graphs):
        dist, iters, edges = batch.result(i)
        assert np.array_equal(dist, refs[i]), f"graph {i}: {int((dist != refs[i]).sum())} differ"
        assert edges == int(g.degrees()[refs[i] != oracle.INF].sum())


@pytest.mark.parametrize("delta", [0, 64, 2048])
@pytest.mark.parametrize("team_edges,chunk", [(1 << 30, 0), (512, 0), (512, 64)])
def test_batched_persistent(rt, small_batch, team_edges, chunk, delta):
    # delta > 0: SSSP near-far bands (single workgroups and teams); 64 forces many band splits
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    st = b.run(wg_size=256, wgs_per_cu=2, team_edges=team_edges, chunk=chunk, max_team=8, delta=delta)
    print(st)
    assert st["launches"] == 1
    _check_all(b, graphs, refs)
    b.close()


@pytest.mark.parametrize("delta,delta_high_degree", [(0, 1), (1024, 1), (512, 8)])
def test_batched_band_per_degree_class(rt, small_batch, delta, delta_high_degree):
    """Graphs with an out-degree above 4 take their own band width (with
    delta 0 only they run near-far): exact either way. At width 1 every band
    is one distance and weights are >= 1, so each reached vertex's edges are
    relaxed exactly once, at its final distance."""
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    b.run(wg_size=256, wgs_per_cu=2, team_edges=512, max_team=8, delta=delta, delta_high_degree=delta_high_degree)
    _check_all(b, graphs, refs)
    b.close()
    hi = [i for i, g in enumerate(graphs) if queries[i] == L.QUERY_SSSP and g.degrees().max() > 4]
    assert hi
    sub = B.Batch(rt, [graphs[i] for i in hi], queries[hi])
    st = sub.run(wg_size=256, wgs_per_cu=2, team_edges=512, max_team=8, delta=delta,
                 delta_high_degree=delta_high_degree)
    for k, i in enumerate(hi):
        assert np.array_equal(sub.result(k)[0], refs[i])
    if delta_high_degree == 1:
        assert st["edges_relaxed"] == st["edges_scanned"]
    else:
        assert st["edges_relaxed"] >= st["edges_scanned"]
    sub.close()


def test_batched_autotune_returns_an_exact_config(rt, small_batch):
    """The autotuner's choice is one of the shapes and bands it sweeps
    (workgroup size, band of the low-degree graphs, band of the others) and
    serves every query exactly."""
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    best, st = b.autotune(reps=1)
    cfg = best.as_dict()
    print(cfg, st["wall_ms"])
    assert cfg["wg_size"] in (64, 128, 256, 512, 1024)
    assert cfg["wg_size"] * cfg["wgs_per_cu"] <= 2048
    assert cfg["chunk"] in (32, 64, 256, 1024, 4096)
    assert cfg["team_edges"] in (1 << 14, 1 << 16, 1 << 18, 1 << 20, 1 << 22)
    assert cfg["delta"] in (0, 256, 512, 1024, 2048, 4096)
    assert cfg["delta_high_degree"] in (0, 4, 8, 12, 16, 32, 128)
    assert cfg["sssp_team_edges"] in (0, 1 << 16, 1 << 18, 1 << 19, 786432, 1 << 20)
    assert cfg["far_flags_lds"] in (0, 1) and cfg["low_ell"] in (0, 1, 2)
    assert st["wall_ms"] > 0
    b.run(best)
    _check_all(b, graphs, refs)
    b.close()


def test_batched_relaunch(rt, small_batch):
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    st = b.run_relaunch(wg_size=256, wgs_per_cu=2, team_edges=2048, max_team=4)
    assert st["launches"] == len(graphs)
    _check_all(b, graphs, refs)
    b.close()


def test_batched_hipgraph_replays(rt, small_batch):
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    cfg = B.config(wg_size=128, wgs_per_cu=4, team_edges=4096, max_team=4, use_graph=True)
    for rep in range(4):
        b.run(cfg)
        _check_all(b, graphs, refs)
        # eager work between replays
        b.run_relaunch(wg_size=64, wgs_per_cu=1, team_edges=1 << 30)
        _check_all(b, graphs, refs)
    b.close()


@pytest.mark.parametrize("wg", [64, 1024])
def test_batched_wg_sizes(rt, small_batch, wg):
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    b.run(wg_size=wg, wgs_per_cu=1, team_edges=1024, max_team=16)
    _check_all(b, graphs, refs)
    b.close()


def test_batched_job_times_work_and_order(rt, small_batch):
    """Per-query device stamps, relaxation counts and the duration-ordered
    job table: BFS relaxes each reached vertex's edges exactly once, SSSP at
    least once; after a timed run the table is rebuilt longest-first and the
    results stay exact, also for a separate SSSP team granularity."""
    graphs, queries, refs = small_batch
    bfs = [i for i, q in enumerate(queries) if q == L.QUERY_BFS]
    b = B.Batch(rt, [graphs[i] for i in bfs], queries[bfs])
    st = b.run(wg_size=256, wgs_per_cu=2, team_edges=512, max_team=8)
    assert st["edges_relaxed"] == st["edges_scanned"]
    b.close()
    b = B.Batch(rt, graphs, queries)
    for rep in range(3):  # first run by edge count, then by measured duration
        st = b.run(wg_size=256, wgs_per_cu=2, team_edges=512, max_team=8, delta=64, sssp_team_edges=256)
        _check_all(b, graphs, refs)
        assert st["edges_relaxed"] >= st["edges_scanned"]
        t0, t1 = b.job_times()
        assert t0.shape == (len(graphs),) and np.all(t0 >= 0) and np.all(t1 >= t0)
        assert t0.min() == 0.0
    b.close()


@pytest.mark.parametrize("wg", [256, 1024])
def test_batched_standalone_baseline(rt, small_batch, wg):
    """The per-graph baseline on wgm_bfs / wgm_sssp's standalone kernel:
    one launch per query on the batch's arrays, every query exact."""
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    st = b.run_standalone(wg_size=wg)
    assert st["launches"] == len(graphs)
    _check_all(b, graphs, refs)
    # the batched launch afterwards is unaffected by the baseline's state
    b.run(wg_size=256, wgs_per_cu=2, team_edges=512, max_team=8, delta=64)
    _check_all(b, graphs, refs)
    b.close()


@pytest.mark.parametrize("delta", [64, 512])
def test_batched_far_flags_in_lds(rt, small_batch, delta):
    """Low-degree SSSP queries on one 1024-thread workgroup with their near-far
    flags in the LDS bitmap: exact, with narrow and wide bands."""
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    for flags in (1, 0, 1):
        b.run(wg_size=1024, wgs_per_cu=1, team_edges=512, max_team=8, delta=delta, delta_high_degree=8,
              far_flags_lds=flags)
        _check_all(b, graphs, refs)
    b.close()


@pytest.mark.parametrize("wg,delta", [(1024, 512), (1024, 0), (256, 64)])
def test_batched_packed_rows(rt, small_batch, wg, delta):
    """Low-degree queries on one workgroup reading packed rows ((target << 8) |
    weight, four words per vertex) instead of the CSR: exact for BFS and SSSP,
    alternating with the CSR step on the same batch."""
    graphs, queries, refs = small_batch
    b = B.Batch(rt, graphs, queries)
    for ell in (1, 0, 2, 1):
        b.run(wg_size=wg, wgs_per_cu=1, team_edges=512, max_team=8, delta=delta, delta_high_degree=8,
              far_flags_lds=1, low_ell=ell)
        _check_all(b, graphs, refs)
    b.close()


@pytest.mark.parametrize("delta", [512, 4096, 1])
def test_batched_lds_combine(rt, delta):
    """Low-degree SSSP on one 1024-thread workgroup with its distances in the
    L2 (low_ell 2): plain loads and stores, candidates folded per target in
    the LDS combine tables. Wide bands give levels of several passes (the
    barrier between passes), unit weights give ties (one winner per slot).
    Exact against the oracle, alternating with the atomic step."""
    from gpu_workgroup_managment_amd import graphs as G
    graphs = [G.add_weights(G.grid(300, 300), seed=11), G.grid(200, 200), G.add_weights(G.grid(64, 1000), seed=12),
              G.grid(120, 90)]
    graphs[1].wt = np.ones(graphs[1].m, np.int32)
    queries = np.array([L.QUERY_SSSP, L.QUERY_SSSP, L.QUERY_SSSP, L.QUERY_BFS], np.int32)
    refs = [oracle.sssp(g.row_ptr, g.col, g.wt, 0) for g in graphs[:3]] + \
        [oracle.bfs(graphs[3].row_ptr, graphs[3].col, 0)[0]]
    b = B.Batch(rt, graphs, queries)
    for ell in (2, 1, 2):
        b.run(wg_size=1024, wgs_per_cu=1, delta=delta, far_flags_lds=1, low_ell=ell)
        _check_all(b, graphs, refs)
    b.close()


def test_batched_lds_combine_largest_bitmap(rt):
    """A 512 x 512 grid SSSP (2^18 vertices, the most the LDS bitmap holds) on
    the L2-distance step beside a 513 x 512 grid, one vertex row too many,
    which takes the atomic step: both exact."""
    from gpu_workgroup_managment_amd import graphs as G
    graphs = [G.add_weights(G.grid(512, 512), seed=21), G.add_weights(G.grid(513, 512), seed=22)]
    queries = np.array([L.QUERY_SSSP, L.QUERY_SSSP], np.int32)
    refs = [oracle.sssp(g.row_ptr, g.col, g.wt, 0) for g in graphs]
    b = B.Batch(rt, graphs, queries)
    b.run(wg_size=1024, wgs_per_cu=1, delta=512, far_flags_lds=1, low_ell=2)
    _check_all(b, graphs, refs)
    b.close()


def test_batched_packed_rows_skip_wide_weights(rt):
    """A low-degree graph with a weight above 255 keeps the CSR step (its
    weights do not fit the packed word); its neighbours in the batch take the
    packed rows; every query exact."""
    from gpu_workgroup_managment_amd import graphs as G
    wide = G.add_weights(G.grid(40, 40), seed=3)
    wide.wt[7] = 300
    graphs = [G.add_weights(G.grid(30, 20), seed=1), wide, G.grid(25, 25)]
    queries = np.array([L.QUERY_SSSP, L.QUERY_SSSP, L.QUERY_BFS], np.int32)
    refs = [oracle.sssp(graphs[0].row_ptr, graphs[0].col, graphs[0].wt, 0),
            oracle.sssp(wide.row_ptr, wide.col, wide.wt, 0), oracle.bfs(graphs[2].row_ptr, graphs[2].col, 0)[0]]
    b = B.Batch(rt, graphs, queries)
    # packed rows (four words per vertex, 64-word aligned) for the first and last graph only
    rows = sum((4 * g.n + 63) // 64 * 64 for g in (graphs[0], graphs[2]))
    assert b.nbytes() == 4 * (sum(g.n + 1 + 2 * g.m for g in graphs) + rows)
    b.run(wg_size=1024, wgs_per_cu=1, delta=64, far_flags_lds=1, low_ell=1)
    _check_all(b, graphs, refs)
    b.close()


@pytest.mark.parametrize("team_edges", [1 << 30, 2048])
def test_batched_packed_edges_skip_wide_weights(rt, team_edges):
    """SSSP on higher-degree graphs reads one packed word per edge ((target <<
    8) | weight) on one workgroup and in teams; a graph with a weight above 255
    keeps col and wt; BFS graphs get no packed edges. Every query exact."""
    from gpu_workgroup_managment_amd import graphs as G
    wide = G.add_weights(G.rmat(10, 8, seed=5), seed=7)
    wide.wt[3] = 4096
    graphs = [G.add_weights(G.rmat(11, 8, seed=2), seed=1), wide, G.rmat(10, 8, seed=6)]
    queries = np.array([L.QUERY_SSSP, L.QUERY_SSSP, L.QUERY_BFS], np.int32)
    refs = [oracle.sssp(graphs[0].row_ptr, graphs[0].col, graphs[0].wt, 0),
            oracle.sssp(wide.row_ptr, wide.col, wide.wt, 0), oracle.bfs(graphs[2].row_ptr, graphs[2].col, 0)[0]]
    assert all(g.degrees().max() > 4 for g in graphs)
    b = B.Batch(rt, graphs, queries)
    packed = (graphs[0].m + 63) // 64 * 64
    assert b.nbytes() == 4 * (sum(g.n + 1 + 2 * g.m for g in graphs) + packed)
    for flags in (1, 0):
        b.run(wg_size=1024, wgs_per_cu=1, team_edges=team_edges, max_team=8, delta=64, delta_high_degree=8,
              far_flags_lds=flags, low_ell=1)
        _check_all(b, graphs, refs)
    b.close()
