"""CPU: graph and solution files (include/wgm.h file formats, csrc/formats.cpp).

Known answers come from the formats' specifications, written out by hand in
each test (a .gr v1 binary packed field by field, the METIS manual's
7-vertex example, small DIMACS / MatrixMarket / edge-list files); round
trips and transposes are checked against the generators and scipy; the
solution styles follow the reference's fscanf patterns."""
import struct

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from gpu_workgroup_managment_amd import _lib as L
from gpu_workgroup_managment_amd import formats as F
from gpu_workgroup_managment_amd import graphs as G


def rows(g):
    """Adjacency as a list of [(col, wt)] per vertex, in CSR order."""
    w = g.wt if g.wt is not None else np.zeros(g.m, np.int32)
    return [list(zip(g.col[g.row_ptr[v]:g.row_ptr[v + 1]].tolist(), w[g.row_ptr[v]:g.row_ptr[v + 1]].tolist()))
            for v in range(g.n)]


def same(a, b):
    return (np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col)
            and ((a.wt is None and b.wt is None) or np.array_equal(a.wt, b.wt)))


def edge_set(g):
    src = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    w = g.wt if g.wt is not None else np.zeros(g.m, np.int32)
    order = np.lexsort((w, g.col, src))
    return np.stack([src[order], g.col[order], w[order]])


# ------------------------------------------------------------------ .gr v1
def pack_gr(n, adj, weights=None):
    """A Galois .gr v1 file from its layout (bfs.h:115-129): version, edge
    data size, n, m; cumulative out-edge ends (u64); destinations (u32);
    a u32 pad when m is odd; then the u32 edge data."""
    m = sum(len(a) for a in adj)
    out = struct.pack("<4Q", 1, 4 if weights else 0, n, m)
    end = 0
    for a in adj:
        end += len(a)
        out += struct.pack("<Q", end)
    out += b"".join(struct.pack("<I", d) for a in adj for d in a)
    if weights:
        if m % 2:
            out += b"\0\0\0\0"
        out += b"".join(struct.pack("<i", w) for ws in weights for w in ws)
    return out


def test_gr_known_answer(tmp_path):
    adj = [[1, 2], [2], [], [0, 1]]  # m = 5 (odd: padded before the edge data)
    wts = [[7, 3], [9], [], [1, 4]]
    p = tmp_path / "g.gr"
    p.write_bytes(pack_gr(4, adj, wts))
    g = F.read_graph(str(p))
    assert g.row_ptr.tolist() == [0, 2, 3, 3, 5]
    assert g.col.tolist() == [1, 2, 2, 0, 1]
    assert g.wt.tolist() == [7, 3, 9, 1, 4]
    # the last vertex keeps its last edge (the reference's psrc[n] = m drops it)
    assert rows(g)[3] == [(0, 1), (1, 4)]
    t = F.read_graph(str(p), transpose=True)
    assert rows(t) == [[(3, 1)], [(0, 7), (3, 4)], [(0, 3), (1, 9)], []]


def test_gr_without_edge_data(tmp_path):
    p = tmp_path / "u.gr"
    p.write_bytes(pack_gr(3, [[1], [0, 2], [1]]))
    g = F.read_graph(str(p), fmt="gr")
    assert g.wt is None and g.col.tolist() == [1, 0, 2, 1] and g.row_ptr.tolist() == [0, 1, 3, 4]


@pytest.mark.parametrize("make", [lambda: G.add_weights(G.rmat(10, 8)), lambda: G.grid(17, 9),
                                  lambda: G.add_weights(G.erdos_renyi(1001, 3001))])
def test_gr_round_trip(tmp_path, make):
    g = make()
    p = tmp_path / "r.gr"
    F.write_graph(str(p), g, "gr")
    assert F.detect_format(str(p)) == "gr"
    assert same(F.read_graph(str(p)), g)


@pytest.mark.parametrize("mutate,msg", [
    (lambda b: b[:40], "header needs"),
    (lambda b: struct.pack("<Q", 2) + b[8:], "version 2"),
    (lambda b: b[:32] + struct.pack("<Q", 4) + b[40:], "not monotone"),
    (lambda b: b[:-28] + struct.pack("<I", 9) + b[-24:], "out of range"),  # last destination
])
def test_gr_rejects_malformed(tmp_path, mutate, msg):
    good = pack_gr(4, [[1, 2], [2], [], [0, 1]], [[7, 3], [9], [], [1, 4]])
    p = tmp_path / "bad.gr"
    p.write_bytes(mutate(good))
    with pytest.raises(L.WgmError, match=msg) as e:
        F.read_graph(str(p), fmt="gr")
    assert e.value.code == L.WGM_ERR_FORMAT


def test_missing_file_is_io_error(tmp_path):
    with pytest.raises(L.WgmError) as e:
        F.read_graph(str(tmp_path / "none.graph"), fmt="metis")
    assert e.value.code == L.WGM_ERR_IO


# ------------------------------------------------------------------ DIMACS
DIMACS = """c 9th DIMACS shortest-path format
c a comment
p sp 4 5
a 1 2 10
a 1 3 5
a 3 2 2
a 4 1 7
a 2 4 1
"""


def test_dimacs_known_answer(tmp_path):
    p = tmp_path / "d.gr"
    p.write_text(DIMACS)
    assert F.detect_format(str(p)) == "dimacs"  # a DIMACS road graph also ends in .gr
    g = F.read_graph(str(p))
    assert rows(g) == [[(1, 10), (2, 5)], [(3, 1)], [(1, 2)], [(0, 7)]]
    # undirected: each arc is followed by its reverse, then a stable sort by row (parse.cpp:217-250)
    u = F.read_graph(str(p), undirected=True)
    assert rows(u) == [[(1, 10), (2, 5), (3, 7)], [(0, 10), (2, 2), (3, 1)], [(0, 5), (1, 2)], [(0, 7), (1, 1)]]
    t = F.read_graph(str(p), transpose=True)
    assert rows(t) == [[(3, 7)], [(0, 10), (2, 2)], [(0, 5)], [(1, 1)]]


def test_dimacs_round_trip_and_transpose_vs_scipy(tmp_path):
    g = G.add_weights(G.rmat(11, 8, truncate_n=1500), lo=1, hi=1000)
    p = tmp_path / "r.dimacs"
    F.write_graph(str(p), g, "dimacs")
    assert same(F.read_graph(str(p)), g)
    t = F.read_graph(str(p), transpose=True)
    src = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    a = sp.coo_matrix((g.wt.astype(np.int64) + 1, (src, g.col)), shape=(g.n, g.n)).T.tocsr()
    b = sp.csr_matrix((t.wt.astype(np.int64) + 1, t.col, t.row_ptr), shape=(t.n, t.n))
    assert (a != b).nnz == 0
    # within a transposed row, arcs keep file order (ascending source here)
    assert all(np.all(np.diff(t.col[t.row_ptr[v]:t.row_ptr[v + 1]]) >= 0) for v in range(t.n))


@pytest.mark.parametrize("text,msg", [
    ("a 1 2 3\np sp 2 1\n", "before the problem line"),
    ("p sp 2 2\na 1 2 3\n", "1 arcs read"),
    ("p sp 2 1\na 1 3 3\n", "out of 1..2"),
    ("p sp 2 1\na 1 2\n", "needs <u> <v> <w>"),
    ("p sp 2 1\nx 1 2 3\n", "unknown line type"),
])
def test_dimacs_rejects_malformed(tmp_path, text, msg):
    p = tmp_path / "bad.dimacs"
    p.write_text(text)
    with pytest.raises(L.WgmError, match=msg):
        F.read_graph(str(p))


# ------------------------------------------------------------------- METIS
# The METIS manual's example graph (7 vertices, 11 undirected edges).
METIS = """% METIS manual, figure 2(a)
7 11
5 3 2
1 3 4
5 4 2 1
2 3 6 7
1 3 6
5 4 7
6 4
"""


def test_metis_known_answer(tmp_path):
    p = tmp_path / "m.graph"
    p.write_text(METIS)
    g = F.read_graph(str(p), undirected=True)
    assert g.wt is None and g.n == 7 and g.m == 22
    assert [[c for c, _ in r] for r in rows(g)] == [[4, 2, 1], [0, 2, 3], [4, 3, 1, 0], [1, 2, 5, 6], [0, 2, 5],
                                                    [4, 3, 6], [5, 3]]
    a = sp.csr_matrix((np.ones(g.m), g.col, g.row_ptr), shape=(7, 7))
    assert (a != a.T).nnz == 0  # both ends listed
    with pytest.raises(L.WgmError, match="header implies 11"):
        F.read_graph(str(p))  # directed: the header's 11 must be the entry count


def test_metis_isolated_vertex_and_weights(tmp_path):
    p = tmp_path / "iso.graph"
    p.write_text("4 1 1\n2 5\n1 5\n\n\n")  # one weighted edge; vertices 3 and 4 have none
    g = F.read_graph(str(p), undirected=True)
    # a complete row_ptr (the reference's emits no start for empty rows, parse.cpp:148-161)
    assert g.row_ptr.tolist() == [0, 1, 2, 2, 2] and g.col.tolist() == [1, 0] and g.wt.tolist() == [5, 5]
    q = tmp_path / "vw.graph"
    q.write_text("2 1 011\n1 2 3\n1 1 3\n")
    with pytest.raises(L.WgmError, match="not supported"):
        F.read_graph(str(q), undirected=True)


@pytest.mark.parametrize("weighted", [False, True])
def test_metis_round_trip(tmp_path, weighted):
    g = G.grid(13, 11)
    if weighted:  # symmetric weights, as an undirected METIS file holds them
        src = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
        g.wt = ((np.minimum(src, g.col) * 31 + np.maximum(src, g.col)) % 97 + 1).astype(np.int32)
    p = tmp_path / "g.graph"
    F.write_graph(str(p), g, "metis", undirected=True)
    assert p.read_text().splitlines()[0] == f"{g.n} {g.m // 2}" + (" 1" if weighted else "")
    assert same(F.read_graph(str(p), undirected=True), g)


# ------------------------------------------------------ MatrixMarket, edges
def test_matrix_market(tmp_path):
    p = tmp_path / "a.mtx"
    p.write_text("%%MatrixMarket matrix coordinate integer general\n% comment\n3 3 4\n1 2 4\n2 2 9\n3 1 6\n2 3 1\n")
    g = F.read_graph(str(p))
    assert rows(g) == [[(1, 4)], [(2, 1)], [(0, 6)]]  # the self-loop 2-2 is dropped (parse.cpp:517-520)
    u = F.read_graph(str(p), undirected=True)
    assert rows(u) == [[(1, 4), (2, 6)], [(0, 4), (2, 1)], [(0, 6), (1, 1)]]
    q = tmp_path / "p.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 2\n")
    assert F.read_graph(str(q)).wt is None


def test_edge_list(tmp_path):
    p = tmp_path / "e.edges"
    p.write_text("3 3\n0 1 5\n2 0 1\n0 2 8\n")
    g = F.read_graph(str(p))
    assert rows(g) == [[(1, 5), (2, 8)], [], [(0, 1)]]
    p.write_text("3 2\n0 1 5\n2 0\n")
    with pytest.raises(L.WgmError, match="some edges only"):
        F.read_graph(str(p))


# --------------------------------------------------------------- solutions
def test_solution_styles(tmp_path):
    v = np.array([0, 3, 1000000000, 7], np.int32)
    for style, first in (("colon", "0:0"), ("colon_space", "1: 0"), ("plain", "0")):
        p = tmp_path / f"{style}.txt"
        F.write_solution(str(p), v, style)
        assert p.read_text().splitlines()[0] == first
        assert F.check_solution(str(p), v, style) == (True, -1)
        w = v.copy()
        w[2] = 5
        assert F.check_solution(str(p), w, style) == (False, 2)
        assert F.check_solution(str(p), np.append(v, 1), style) == (False, 4)  # the file is short
    # fscanf("%d:%d") and ("%d: %d") read either spacing; the index is not compared
    p = tmp_path / "mixed.txt"
    p.write_text("5:0\n9: 3\n  2 :1000000000\n7:7\n")
    assert F.check_solution(str(p), v, "colon") == (True, -1)
    assert F.check_solution(str(p), v, "colon_space") == (True, -1)
    p.write_text("0:0\n1 3\n")
    assert F.check_solution(str(p), v, "colon") == (False, 1)


def test_solution_f32(tmp_path):
    v = np.array([0.5, 120.25, 3e6], np.float32)
    p = tmp_path / "bc.txt"
    F.write_solution_f32(str(p), v)
    assert F.check_solution_f32(str(p), v) == (True, -1)
    assert F.check_solution_f32(str(p), v + np.float32(7.5)) == (True, -1)  # within BC's tolerance 8
    assert F.check_solution_f32(str(p), v + np.array([0, 8.5, 0], np.float32)) == (False, 1)
    assert F.check_solution_f32(str(p), v, tol=0.0) == (True, -1)


def test_file_to_oracle_solution(tmp_path):
    """A DIMACS file read here gives the oracle the same graph as the
    generator's CSR, and its SSSP solution file checks out."""
    g = G.add_weights(G.erdos_renyi(3000, 9000), lo=1, hi=100)
    p = tmp_path / "g.dimacs"
    F.write_graph(str(p), g, "dimacs")
    h = F.read_graph(str(p))
    d1 = oracle.sssp(g.row_ptr, g.col, g.wt, 0)
    d2 = oracle.sssp(h.row_ptr, h.col, h.wt, 0)
    assert np.array_equal(d1, d2)
    s = tmp_path / "g.sol"
    F.write_solution(str(s), d1.astype(np.int32), "colon")
    assert F.check_solution(str(s), d2.astype(np.int32), "colon") == (True, -1)
