"""GPU parity: persistent BFS / SSSP / colour / MIS / PageRank vs the oracle."""
import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import _lib as L
from gpu_workgroup_managment_amd import graphs as G

pytestmark = pytest.mark.gpu


def _first_diff(a, b):
    idx = np.flatnonzero(a != b)
    return f"{len(idx)} differ, first at {idx[0]}: got {a[idx[0]]} want {b[idx[0]]}" if len(idx) else ""


@pytest.mark.parametrize("n,pairs,wgs", [(1024, 2048, 4), (4096, 16384, 64), (1 << 16, 1 << 18, 0)])
def test_bfs_er(rt, n, pairs, wgs):
    g = G.erdos_renyi(n, pairs, seed=1)
    dist, st = rt.graph(g).bfs(0, wg_size=256, num_wgs=wgs)
    ref, levels = oracle.bfs(g.row_ptr, g.col, 0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    assert st["iterations"] == levels
    reached = ref != oracle.INF
    assert st["edges_scanned"] == int(g.degrees()[reached].sum())
    print(st)


@pytest.mark.parametrize("wg,wgs", [(256, 16), (1024, 0)])
def test_bfs_directed_rmat_bottom_up(rt, wg, wgs):
    """A directed RMAT graph: its wide levels run bottom-up over the in-edges
    (which differ from the out-edges here), the narrow ones top-down; levels,
    level count and scanned edges are the oracle's."""
    g = G.rmat(16, 16, seed=3)
    src = int(np.argmax(g.degrees()))  # a hub: the second level is most of the graph
    ref, levels = oracle.bfs(g.row_ptr, g.col, src)
    assert int(np.bincount(ref[ref != oracle.INF]).max()) > g.n // 16  # some level is wide
    dist, st = rt.graph(g).bfs(src, wg_size=wg, num_wgs=wgs)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    assert st["iterations"] == levels
    assert st["edges_scanned"] == int(g.degrees()[ref != oracle.INF].sum())


def test_bfs_grid_closed_form(rt):
    side = 300
    g = G.grid(side, side)
    dist, st = rt.graph(g).bfs(0, wg_size=256, num_wgs=32)
    r, c = np.divmod(np.arange(side * side), side)
    assert np.array_equal(dist, (r + c).astype(np.uint32))
    assert st["iterations"] == 2 * side - 1


@pytest.mark.parametrize("wg", [64, 512, 1024])
def test_bfs_wg_sizes(rt, wg):
    g = G.rmat(12, 8, seed=7)
    dist, _ = rt.graph(g).bfs(0, wg_size=wg, num_wgs=16)
    ref, _ = oracle.bfs(g.row_ptr, g.col, 0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)


def test_bfs_handshake(rt):
    g = G.erdos_renyi(4096, 16384, seed=3)
    dist, st = rt.graph(g).bfs(0, wg_size=256, num_wgs=32, handshake=True)
    ref, _ = oracle.bfs(g.row_ptr, g.col, 0)
    assert np.array_equal(dist, ref)
    assert st["host_ms"] > 0


@pytest.mark.parametrize("mode", [L.SSSP_AUTO, L.SSSP_PUSH, L.SSSP_PULL])
@pytest.mark.parametrize("scale,wgs", [(10, 8), (14, 0)])
def test_sssp_rmat(rt, scale, wgs, mode):
    """Both relaxation directions give Dijkstra's distances (AUTO pulls here:
    the out-degree is above 4), and the scanned-edge count of the reached
    vertices either way."""
    g = G.add_weights(G.rmat(scale, 16, seed=2), 1, 255, 3)
    dist, st = rt.graph(g).sssp(0, wg_size=256, num_wgs=wgs, mode=mode)
    ref = oracle.sssp(g.row_ptr, g.col, g.wt, 0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    assert st["edges_scanned"] == int(g.degrees()[ref != oracle.INF].sum())
    print(st)


@pytest.mark.parametrize("wg", [64, 192, 1024])
def test_sssp_push_workgroup_sizes(rt, wg):
    """Push SSSP spreads a batch's edges several to a thread per pass, each
    slot finding its source entry by a bisection of ceil(log2 wg) steps: exact
    at a workgroup size that is not a power of two, at the smallest and at the
    largest, on a skewed and on a low-degree graph."""
    for g in (G.add_weights(G.rmat(12, 16, seed=5), 1, 255, 6), G.add_weights(G.grid(90, 70), 1, 255, 7)):
        dist, st = rt.graph(g).sssp(0, wg_size=wg, num_wgs=0, mode=L.SSSP_PUSH)
        ref = oracle.sssp(g.row_ptr, g.col, g.wt, 0)
        assert np.array_equal(dist, ref), (wg, g.n, _first_diff(dist, ref))
        assert st["edges_scanned"] == int(g.degrees()[ref != oracle.INF].sum())


@pytest.mark.parametrize("mode", [L.SSSP_AUTO, L.SSSP_PULL])
def test_sssp_grid_unit_weights(rt, mode):
    side = 128
    g = G.grid(side, side)
    g.wt = np.ones(g.m, np.int32)
    dist, st = rt.graph(g).sssp(0, wg_size=256, num_wgs=16, mode=mode)
    r, c = np.divmod(np.arange(side * side), side)
    assert np.array_equal(dist, (r + c).astype(np.uint32))


@pytest.mark.parametrize("side,wgs", [(64, 4), (512, 0)])
def test_color_grid(rt, side, wgs):
    g = G.add_priorities(G.grid(side, side))
    col, st = rt.graph(g).color(wg_size=256, num_wgs=wgs)
    ref, rounds = oracle.color(g.row_ptr, g.col, g.prio)
    assert np.array_equal(col, ref), _first_diff(col, ref)
    assert st["iterations"] == rounds


def test_color_er_with_degree_one(rt):
    g = G.add_priorities(G.erdos_renyi(5000, 4000, seed=9))  # many degree-0/1 vertices
    col, _ = rt.graph(g).color(wg_size=128, num_wgs=8)
    ref, _ = oracle.color(g.row_ptr, g.col, g.prio)
    assert np.array_equal(col, ref), _first_diff(col, ref)


def test_color_rmat(rt):
    """Vertices of high degree: the edges past the first four of a vertex
    are walked after the grouped loads."""
    g = G.rmat(12, 8, seed=5)
    src = np.repeat(np.arange(g.n), g.degrees())
    g = G.add_priorities(G.from_edges(g.n, src, g.col, symmetric=True))
    assert g.degrees().max() > 64
    for wg, wgs in ((256, 32), (1024, 0)):
        col, _ = rt.graph(g).color(wg_size=wg, num_wgs=wgs)
        ref, _ = oracle.color(g.row_ptr, g.col, g.prio)
        assert np.array_equal(col, ref), _first_diff(col, ref)


@pytest.mark.parametrize("side,wgs", [(64, 4), (512, 0)])
def test_mis_grid(rt, side, wgs):
    g = G.add_priorities(G.grid(side, side))
    s, st = rt.graph(g).mis(wg_size=256, num_wgs=wgs)
    ref, rounds = oracle.mis(g.row_ptr, g.col, g.prio)
    assert np.array_equal(s, ref), _first_diff(s, ref)
    assert st["iterations"] == rounds


def test_mis_rmat(rt):
    g = G.rmat(12, 8, seed=5)
    # make it undirected
    src = np.repeat(np.arange(g.n), g.degrees())
    g = G.add_priorities(G.from_edges(g.n, src, g.col, symmetric=True))
    s, _ = rt.graph(g).mis(wg_size=256, num_wgs=32)
    ref, _ = oracle.mis(g.row_ptr, g.col, g.prio)
    assert np.array_equal(s, ref), _first_diff(s, ref)


@pytest.mark.parametrize("side,wgs", [(64, 4), (256, 0)])
def test_pagerank_grid(rt, side, wgs):
    g = G.grid(side, side)
    r, st = rt.graph(g).pagerank(0.85, 50, wg_size=256, num_wgs=wgs)
    ref32 = oracle.pagerank_f32(g.row_ptr, g.col, 0.85, 50)
    ref64 = oracle.pagerank_f64(g.row_ptr, g.col, 0.85, 50)
    assert np.array_equal(r.view(np.uint32), ref32.view(np.uint32)), _first_diff(r, ref32)
    l1 = float(np.abs(r.astype(np.float64) - ref64).sum())
    print(f"PageRank {side}^2: L1 vs FP64 = {l1:.3e}")
    assert l1 < 1e-4


def _hub_chain(levels=6, leaves=(3000, 6000), seed=5, wmax=256):
    """Two hubs per level, each pointing at its own block of leaves; every leaf
    points at both hubs of the next level, plus random edges. Each frontier
    then holds several entries far above the heavy threshold (4 * bd, at
    least 1024 edges) next to many small ones."""
    rng = np.random.default_rng(seed)
    hubs = np.arange(2 * levels).reshape(levels, 2)
    nxt = 2 * levels
    src, dst = [], []
    for lv in range(levels):
        for k in range(2):
            cnt = leaves[k]
            blk = np.arange(nxt, nxt + cnt)
            nxt += cnt
            src.append(np.full(cnt, hubs[lv, k]))
            dst.append(blk)
            if lv + 1 < levels:
                for h in hubs[lv + 1]:
                    src.append(blk)
                    dst.append(np.full(cnt, h))
    n = nxt
    extra = 4 * n
    src.append(rng.integers(2 * levels, n, extra))
    dst.append(rng.integers(0, n, extra))
    src, dst = np.concatenate(src), np.concatenate(dst)
    wt = rng.integers(1, wmax, src.size).astype(np.int32)
    return G.from_edges(n, src, dst, wt)


@pytest.mark.parametrize("wg", [64, 256])
def test_frontier_heavy_list_of_several_chunks(rt, wg):
    """A level of 300 heavy entries (1100-1500 edges each, above the 1024-edge
    threshold of 64- and 256-thread workgroups) is more than one workgroup's
    width of the heavy list: the pass scans it in several chunks, each with its
    own claim counter. BFS levels and SSSP distances stay exact."""
    rng = np.random.default_rng(11)
    nhub, leaves = 300, 60000
    n = 1 + nhub + leaves
    degs = rng.integers(1100, 1501, nhub)
    src = [np.zeros(nhub, np.int64)] + [np.full(d, 1 + h) for h, d in enumerate(degs)]
    dst = [np.arange(1, 1 + nhub)] + [rng.integers(1 + nhub, n, d) for d in degs]
    src.append(rng.integers(1 + nhub, n, 3 * leaves))
    dst.append(rng.integers(1, n, 3 * leaves))
    src, dst = np.concatenate(src), np.concatenate(dst)
    g = G.from_edges(n, src, dst, rng.integers(1, 5000, src.size).astype(np.int32))
    assert (g.degrees() > 1024).sum() >= nhub
    dg = rt.graph(g)
    dist, st = dg.bfs(0, wg_size=wg, num_wgs=8)
    ref, levels = oracle.bfs(g.row_ptr, g.col, 0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    assert st["iterations"] == levels
    dist, st = dg.sssp(0, wg_size=wg, num_wgs=8, mode=L.SSSP_PUSH)
    ref = oracle.sssp(g.row_ptr, g.col, g.wt, 0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    assert st["edges_scanned"] == int(g.degrees()[ref != oracle.INF].sum())


@pytest.mark.parametrize("wg", [64, 256, 1024])
def test_frontier_heavy_entries(rt, wg):
    """Entries of thousands of edges are deferred and their edges claimed in
    pieces of one list by every workgroup; BFS levels and SSSP distances stay
    exact, the latter with weights large enough for many near-far bands."""
    g = _hub_chain()
    assert g.degrees().max() > 4 * 1024
    dg = rt.graph(g)
    dist, st = dg.bfs(0, wg_size=wg, num_wgs=16)
    ref, levels = oracle.bfs(g.row_ptr, g.col, 0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    assert st["iterations"] == levels
    g = _hub_chain(wmax=20000)
    dg = rt.graph(g)
    dist, st = dg.sssp(0, wg_size=wg, num_wgs=16, mode=L.SSSP_PUSH)
    ref = oracle.sssp(g.row_ptr, g.col, g.wt, 0)
    assert np.array_equal(dist, ref), _first_diff(dist, ref)
    assert st["edges_scanned"] == int(g.degrees()[ref != oracle.INF].sum())
