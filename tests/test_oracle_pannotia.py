"""CPU: the oracle's Pannotia SSSP (min-plus sweeps) and BC (Brandes, FP64),
pinned independently of their own formulas:

* the in-edge CSR against scipy's transpose;
* min-plus distances against Dijkstra (itself pinned against scipy), the
  sweep count against its closed form on a path;
* BC against a brute-force sum over ordered pairs (s, t) of
  sigma_sv * sigma_vt / sigma_st from per-source BFS path counts (no
  dependency accumulation), and against closed forms on paths, stars,
  complete graphs and odd cycles."""
from collections import deque

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from gpu_workgroup_managment_amd import graphs as G


def directed(n, m, seed, wlo=1, whi=20):
    rng = np.random.default_rng(seed)
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = src != dst
    return G.from_edges(n, src[keep], dst[keep], rng.integers(wlo, whi + 1, keep.sum()))


def test_reverse_csr_is_the_transpose():
    g = directed(300, 2500, 1)
    rrow, rcol, rwt = oracle.reverse_csr(g.row_ptr, g.col, g.wt)
    src = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    a = sp.coo_matrix((g.wt.astype(np.int64) + 1, (src, g.col)), shape=(g.n, g.n)).T.tocsr()
    b = sp.csr_matrix((rwt.astype(np.int64) + 1, rcol, rrow), shape=(g.n, g.n))
    assert (a != b).nnz == 0
    assert all(np.all(np.diff(rcol[rrow[v]:rrow[v + 1]]) >= 0) for v in range(g.n))  # sources ascending


@pytest.mark.parametrize("make", [lambda: directed(2000, 9000, 2), lambda: G.add_weights(G.rmat(11, 8)),
                                  lambda: G.add_weights(G.grid(30, 20), lo=1, hi=9)])
def test_minplus_equals_dijkstra(make):
    g = make()
    d, sweeps = oracle.sssp_minplus(g.row_ptr, g.col, g.wt, 0)
    ref = oracle.sssp(g.row_ptr, g.col, g.wt, 0).astype(np.int64)
    ref[ref == oracle.INF] = oracle.BIG_NUM
    assert np.array_equal(d.astype(np.int64), ref)
    assert sweeps >= 2


def test_minplus_sweeps_closed_form():
    k = 12  # directed path 0 -> 1 -> ... -> k: one hop per Jacobi sweep, then one sweep without change
    g = G.from_edges(k + 1, np.arange(k), np.arange(1, k + 1), np.full(k, 3))
    d, sweeps = oracle.sssp_minplus(g.row_ptr, g.col, g.wt, 0)
    assert d.tolist() == [3 * i for i in range(k + 1)] and sweeps == k + 1
    d, _ = oracle.sssp_minplus(g.row_ptr, g.col, g.wt, 5)
    assert d.tolist() == [oracle.BIG_NUM] * 5 + [3 * i for i in range(k + 1 - 5)]


def brute_bc(g):
    """sum over ordered (s, t), s != v != t, of sigma_sv sigma_vt / sigma_st."""
    n = g.n
    dist = np.full((n, n), -1, np.int64)
    cnt = np.zeros((n, n))
    for s in range(n):
        dist[s, s], cnt[s, s] = 0, 1.0
        q = deque([s])
        while q:
            u = q.popleft()
            for w in g.col[g.row_ptr[u]:g.row_ptr[u + 1]]:
                if dist[s, w] < 0:
                    dist[s, w] = dist[s, u] + 1
                    q.append(w)
                if dist[s, w] == dist[s, u] + 1:
                    cnt[s, w] += cnt[s, u]
    out = np.zeros(n)
    for v in range(n):
        for s in range(n):
            for t in range(n):
                if len({s, t, v}) == 3 and dist[s, t] > 0 and dist[s, v] >= 0 and dist[v, t] >= 0 \
                        and dist[s, v] + dist[v, t] == dist[s, t]:
                    out[v] += cnt[s, v] * cnt[v, t] / cnt[s, t]
    return out


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_bc_against_pair_sum(seed):
    rng = np.random.default_rng(seed)
    n = 28
    src, dst = rng.integers(0, n, 70), rng.integers(0, n, 70)
    keep = src != dst
    g = G.from_edges(n, src[keep], dst[keep])  # directed
    got, _ = oracle.bc(g.row_ptr, g.col)
    np.testing.assert_allclose(got, brute_bc(g), rtol=1e-12, atol=1e-12)


def test_bc_closed_forms():
    n = 9
    path = G.from_edges(n, np.arange(n - 1), np.arange(1, n), symmetric=True)
    assert oracle.bc(path.row_ptr, path.col)[0].tolist() == [2.0 * i * (n - 1 - i) for i in range(n)]
    k = 7
    star = G.from_edges(k + 1, np.zeros(k, np.int64), np.arange(1, k + 1), symmetric=True)
    assert oracle.bc(star.row_ptr, star.col)[0].tolist() == [float(k * (k - 1))] + [0.0] * k
    s, d = np.nonzero(~np.eye(6, dtype=bool))
    kn = G.from_edges(6, s, d)
    assert not oracle.bc(kn.row_ptr, kn.col)[0].any()
    r = 4  # odd cycle of 2r+1 vertices: every vertex r(r-1)
    cyc = G.from_edges(2 * r + 1, np.arange(2 * r + 1), (np.arange(2 * r + 1) + 1) % (2 * r + 1), symmetric=True)
    assert oracle.bc(cyc.row_ptr, cyc.col)[0].tolist() == [float(r * (r - 1))] * (2 * r + 1)


def test_bc_partial_sources_accumulate():
    g = G.rmat(9, 4)
    full, _ = oracle.bc(g.row_ptr, g.col)
    part, _ = oracle.bc(g.row_ptr, g.col, 100)
    rest = full - part
    assert np.all(rest >= -1e-9) and part.sum() > 0
    assert np.array_equal(oracle.bc(g.row_ptr, g.col, g.n)[0], full)
