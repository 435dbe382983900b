"""CPU: the seeded generators of libwgm (host code, no GPU) against the
published splitmix64 vectors, the structural invariants of each graph
family and the committed fixtures (tests/golden/fixtures.json); and the
oracle's answers on the small workloads against the same fixtures."""
import json
import os

import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import graphs as G

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "fixtures.json")) as _f:
    FX = json.load(_f)

import sys  # noqa: E402

sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402


def test_splitmix64_published_vector():
    # splitmix64 (Steele, Lea, Flood 2014; Vigna's C reference), seed 1234567
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert G.splitmix64(1234567, 5) == want
    assert FX["splitmix64_seed_1234567"] == want


def _check_csr(g, symmetric):
    rp, col = g.row_ptr, g.col
    assert rp[0] == 0 and np.all(np.diff(rp) >= 0) and rp[-1] == len(col)
    assert col.dtype == np.int32 and rp.dtype == np.int32
    assert np.all((col >= 0) & (col < g.n))
    src = np.repeat(np.arange(g.n), np.diff(rp))
    assert not np.any(src == col), "self-loop"
    key = src.astype(np.int64) * g.n + col
    assert np.all(np.diff(key) > 0), "neighbours ascending, no duplicates"
    if symmetric:
        rev = np.sort(col.astype(np.int64) * g.n + src)
        assert np.array_equal(rev, key)


def test_er_invariants():
    g = G.erdos_renyi(5000, 20000, 1)
    _check_csr(g, symmetric=True)
    assert g.m == 40000  # distinct undirected pairs, both directions


def test_rmat_invariants():
    g = G.rmat(12, 16, seed=2)
    _check_csr(g, symmetric=False)
    assert g.n == 4096 and 0 < g.m <= 16 * 4096
    deg = g.degrees()
    assert deg[0] > 10 * deg.mean()  # a=0.57 skews towards vertex 0
    t = G.rmat(12, 16, seed=2, truncate_n=3000)
    _check_csr(t, symmetric=False)
    assert t.n == 3000


def test_grid_closed_form():
    rows, cols = 7, 11
    g = G.grid(rows, cols)
    _check_csr(g, symmetric=True)
    assert g.m == 2 * (rows * (cols - 1) + cols * (rows - 1))
    v = 3 * cols + 4
    assert g.col[g.row_ptr[v]:g.row_ptr[v + 1]].tolist() == [v - cols, v - 1, v + 1, v + cols]


def test_weights_and_priorities():
    g = G.add_priorities(G.add_weights(G.erdos_renyi(3000, 9000, 1), 1, 255, 3), 4)
    assert g.wt.min() >= 1 and g.wt.max() <= 255
    assert len(np.unique(g.wt)) == 255
    src = np.repeat(np.arange(g.n), g.degrees())
    assert not np.any(g.prio[src] == g.prio[g.col]), "adjacent priorities must differ"
    g2 = G.add_weights(G.erdos_renyi(3000, 9000, 1), 1, 255, 3)
    assert np.array_equal(g.wt, g2.wt)  # deterministic


def test_batch_spec():
    items = G.batch_spec(512, 5)
    ns = np.array([it.n for it in items])
    assert ns.min() >= 1 << 10 and ns.max() <= 1 << 18
    kinds = [it.kind for it in items]
    assert kinds == [i % 2 for i in range(512)]  # even RMAT, odd grid
    for it in items:
        if it.kind == 1:
            assert it.rows * it.cols == it.n or it.rows == it.cols
    # log-uniform: the median of log2 n is near the middle of [10, 18]
    assert 13 < np.median(np.log2(ns)) < 15
    assert sum(ns) == FX["batch"]["sum_n"]


@pytest.mark.parametrize("name", ["er_small", "rmat_small", "rmat_trunc", "grid_small"])
def test_small_fixtures(name):
    g = make_golden.small_graphs()[name]
    assert make_golden.digest(g) == FX["graphs"][name]
    assert make_golden.answers(name, g) == FX["answers"][name]


def test_config_fixtures():
    for name, g in make_golden.config_graphs().items():
        assert make_golden.digest(g) == FX["graphs"][name], name
    assert FX["graphs"]["grid_config"]["m"] == 2 * (2048 * 2047 * 2)
    assert FX["graphs"]["er_config"]["m"] == 524288


def test_batch_and_matmult_fixtures():
    assert make_golden.batch_digest() == FX["batch"]
    assert make_golden.matmult_answer() == FX["matmult"]
