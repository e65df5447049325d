"""run_all over the device containers (byo_amd.run, restating
run_unweighted.h:930-1034): every runner of the sequence prints its lines
and a time, on the reference's own input graph, for the CSR and the CPMA;
with -i the Batch_insert_runner sweep runs and leaves the edge set as it
found it."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

INPUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs", "rMatGraph_J_5_100")

RUN_ALL = ["Map"] + ["Map_withRemote_%d" % w for w in (0, 1, 2, 4, 8, 16, 32, 64)] + [
    "BFS", "WorkEfficientDensestSubgraph", "SimpleUnionCC", "Coloring", "KCore", "LDD", "MIS", "PageRank",
    "SSBetweennessCentrality", "Spanner", "TriangleCount"]


@pytest.mark.parametrize("extra", [[], ["--cpma"], ["--fp32"]])
def test_run_all_on_reference_input(extra, capsys):
    from byo_amd import run as R
    rc = R.main(["-s", "-rounds", "1", "-src", "10"] + extra + [INPUT])
    out = capsys.readouterr().out
    assert rc == 1
    for alg in RUN_ALL:
        assert "# # # %s, " % alg in out, alg
    assert "### Num triangles = 107" in out
    assert "num components = 4" in out
    assert out.count("# time per iter:") == len(RUN_ALL)


def test_batch_insert_runner_restores_the_graph(capsys):
    import byo_amd as B
    from byo_amd import run as R
    g = O.read_adjacency_graph(INPUT)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    tm = {}
    R.batch_insert_runner(tm, G, rounds=2, max_batch=1000)
    for b in (1, 10, 100, 1000):
        assert "batch_insert_%d" % b in tm and "batch_remove_%d" % b in tm
    off, ids = G.export_csr()
    # every batch is removed again: only edges the graph already had and the
    # batch held are gone (remove is a set difference)
    E = O.graph_keys(g)
    s, d = O.update_rmat(g.n, 1000)
    ups = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
    assert np.array_equal(O.graph_keys(O.Graph(off, ids)), np.setdiff1d(E, ups))
    G.close()
    assert R.main(["-s", "-i", "-rounds", "1", "-max_batch", "100", "--cpma", INPUT]) == 1
    out = capsys.readouterr().out
    assert "# # # batch_insert_100, " in out and "# # # batch_remove_1, " in out
