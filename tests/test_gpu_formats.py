"""GPU: graphs read from the reference's file formats run through the HIP
path, and their results go through the reference's solution-file check.
The solution files are written from the oracle's answers on the CSR the
generator produced (not from the file), so reader and device are checked
together."""
import numpy as np
import pytest

import oracle
from gpu_workgroup_managment_amd import formats as F
from gpu_workgroup_managment_amd import graphs as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", ["gr", "dimacs"])
def test_file_graph_bfs_sssp_solutions(rt, tmp_path, fmt):
    g = G.add_weights(G.rmat(14, 8), lo=1, hi=255)
    path = tmp_path / f"g.{'gr' if fmt == 'gr' else 'dimacs'}"
    F.write_graph(str(path), g, fmt)
    h = F.read_graph(str(path))
    ref_bfs, _ = oracle.bfs(g.row_ptr, g.col, 0)
    ref_sssp = oracle.sssp(g.row_ptr, g.col, g.wt, 0)
    sol_bfs, sol_sssp = tmp_path / "bfs.sol", tmp_path / "sssp.sol"
    F.write_solution(str(sol_bfs), ref_bfs.astype(np.int32), "colon")
    F.write_solution(str(sol_sssp), ref_sssp.astype(np.int32), "colon")
    dg = rt.graph(h)
    try:
        dist, _ = dg.bfs(0)
        assert F.check_solution(str(sol_bfs), dist.astype(np.int32), "colon") == (True, -1)
        dist, _ = dg.sssp(0)
        assert F.check_solution(str(sol_sssp), dist.astype(np.int32), "colon") == (True, -1)
    finally:
        dg.close()


def test_metis_graph_color_mis_solutions(rt, tmp_path):
    g = G.add_priorities(G.grid(96, 80))
    path = tmp_path / "g.graph"
    F.write_graph(str(path), g, "metis", undirected=True)
    h = F.read_graph(str(path), undirected=True)
    h.prio = g.prio
    ref_color, _ = oracle.color(g.row_ptr, g.col, g.prio)
    ref_mis, _ = oracle.mis(g.row_ptr, g.col, g.prio)
    dg = rt.graph(h)
    try:
        colors, _ = dg.color()
        mis, _ = dg.mis()
    finally:
        dg.close()
    sol = tmp_path / "color.sol"
    F.write_solution(str(sol), ref_color, "colon_space")
    assert F.check_solution(str(sol), colors.astype(np.int32), "colon_space") == (True, -1)
    sol = tmp_path / "mis.sol"
    F.write_solution(str(sol), ref_mis, "plain")
    assert F.check_solution(str(sol), mis.astype(np.int32), "plain") == (True, -1)
