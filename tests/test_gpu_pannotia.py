"""GPU parity: Pannotia SSSP (min-plus sweeps) and Pannotia BC (Brandes,
FP64) through the C ABI against the oracle.

SSSP distances are integers: bit-exact, BIG_NUM where unreached, and the
in-place sweeps never need more sweeps than the reference's Jacobi form. BC
is FP64 with a fixed summation order on both sides: bit-exact, whatever the
grid. At full size (RMAT scale 20) the min-plus distances are checked
against the Lonestar SSSP on the same graph (a different algorithm and
kernel) plus the edge-relaxation property."""
import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import _lib as L
from gpu_workgroup_managment_amd import graphs as G

pytestmark = pytest.mark.gpu


def directed(n, m, seed, whi=50):
    rng = np.random.default_rng(seed)
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = src != dst
    return G.from_edges(n, src[keep], dst[keep], rng.integers(1, whi + 1, keep.sum()))


SSSP_CASES = {
    "rmat14": lambda: G.add_weights(G.rmat(14, 8), lo=1, hi=255),
    "grid96": lambda: G.add_weights(G.grid(96, 80), lo=1, hi=9),
    "directed": lambda: directed(20000, 50000, 9),  # sparse: many unreached vertices
}


@pytest.mark.parametrize("name", list(SSSP_CASES))
@pytest.mark.parametrize("wg,nw", [(256, 0), (1024, 0), (64, 7)])
def test_minplus_parity(rt, name, wg, nw):
    g = SSSP_CASES[name]()
    ref, jacobi_sweeps = oracle.sssp_minplus(g.row_ptr, g.col, g.wt, 0)
    dg = rt.graph(g)
    try:
        dist, st = dg.sssp_minplus(0, wg_size=wg, num_wgs=nw)
    finally:
        dg.close()
    bad = np.flatnonzero(dist != ref)
    assert bad.size == 0, f"{name}: {bad.size} wrong, first v={bad[0]}: {dist[bad[0]]} != {ref[bad[0]]}"
    assert 1 <= st["iterations"] <= jacobi_sweeps
    print(name, wg, nw, "sweeps", st["iterations"], "jacobi", jacobi_sweeps, "unreached",
          int((dist == L.BIG_NUM).sum()), f"{st['kernel_ms']:.2f} ms")


def test_minplus_full_size_against_lonestar_sssp(rt):
    g = G.add_weights(G.rmat(20, 16), lo=1, hi=255)
    dg = rt.graph(g)
    try:
        mp, st = dg.sssp_minplus(0)
        ls, _ = dg.sssp(0)
    finally:
        dg.close()
    ls = ls.astype(np.int64)
    ls[ls == oracle.INF] = L.BIG_NUM
    assert np.array_equal(mp.astype(np.int64), ls)
    src = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    reach = mp[src] != L.BIG_NUM
    assert np.all(mp[g.col[reach]] <= mp[src[reach]] + g.wt[reach])  # no edge can still relax
    print("rmat20 sweeps", st["iterations"], f"{st['kernel_ms']:.2f} ms",
          f"{st['edges_scanned'] / st['kernel_ms'] / 1e6:.2f} GTEPS(sweep edges)")


BC_CASES = {
    "rmat11": lambda: G.rmat(11, 8),
    "grid24": lambda: G.grid(24, 20),
    "directed": lambda: directed(1500, 6000, 4),
}


@pytest.mark.parametrize("name", list(BC_CASES))
@pytest.mark.parametrize("wg,nw", [(256, 0), (1024, 0), (64, 5)])
def test_bc_parity_bit_exact(rt, name, wg, nw):
    g = BC_CASES[name]()
    ref, levels = oracle.bc(g.row_ptr, g.col)
    dg = rt.graph(g)
    try:
        bc, st = dg.bc(wg_size=wg, num_wgs=nw)
    finally:
        dg.close()
    bad = np.flatnonzero(bc != ref)
    assert bad.size == 0, f"{name}: {bad.size} differ, first v={bad[0]}: {bc[bad[0]]!r} != {ref[bad[0]]!r}"
    assert st["iterations"] == levels
    print(name, wg, nw, "levels", levels, f"{st['kernel_ms']:.2f} ms")


def test_bc_partial_sources_and_closed_form(rt):
    g = G.rmat(12, 8)
    ref, _ = oracle.bc(g.row_ptr, g.col, 300)
    dg = rt.graph(g)
    try:
        bc, _ = dg.bc(300)
        again, _ = dg.bc(300, wg_size=512)
    finally:
        dg.close()
    assert np.array_equal(bc, ref) and np.array_equal(again, ref)
    n = 65
    path = G.from_edges(n, np.arange(n - 1), np.arange(1, n), symmetric=True)
    dp = rt.graph(path)
    try:
        pb, _ = dp.bc()
    finally:
        dp.close()
    assert pb.tolist() == [2.0 * i * (n - 1 - i) for i in range(n)]


def test_minplus_negative_cycle_is_an_error(rt):
    g = G.from_edges(4, np.array([0, 1, 2, 3]), np.array([1, 2, 3, 1]), np.array([2, 1, 1, -5]))
    dg = rt.graph(g)
    try:
        with pytest.raises(L.WgmError, match="negative cycle"):
            dg.sssp_minplus(0)
    finally:
        dg.close()
