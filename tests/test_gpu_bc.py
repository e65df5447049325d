"""GPU parity of single-source betweenness centrality (byo_bc) with the
oracle's restatement of SSBetweennessCentrality_BFS (itself pinned by brute
force in test_oracle.py): FP64 values to 1e-9 relative (sums run in another
order), NaN exactly where the source does not reach, on the reference's
input files, RMAT, grid and many-component graphs, under every direction
policy, and on a CPMA graph after updates."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")


@pytest.fixture(scope="module")
def B():
    import byo_amd
    assert byo_amd.device_count() > 0, "no GPU visible"
    return byo_amd


def check(g, G, src):
    print("bc", g.n, g.m, src, flush=True)
    got, ms = G.bc(src)
    ref = O.bc(g, src)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan), src
    assert np.allclose(got[~nan], ref[~nan], rtol=1e-9, atol=1e-9), src
    assert ms > 0
    return got


@pytest.mark.parametrize("name", ["rMatGraph_J_5_100", "star.txt", "triangles.txt"])
def test_reference_inputs(B, name):
    g = O.read_adjacency_graph(os.path.join(INPUTS, name))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    for src in sorted({0, 10 % g.n, g.n - 1}):
        check(g, G, src)
    G.close()


def test_rmat_sources_and_policies(B):
    g = O.rmat_graph(15)
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    srcs = O.bfs_sources(g, 4, seed=21)
    for s in srcs:
        check(g, G, s)
    try:
        for policy in (2, 3):  # all push, all pull: the levels and so the values do not change
            G.set_direction_policy(policy)
            check(g, G, srcs[0])
    finally:
        G.set_direction_policy(0)
    G.close()


def test_grid_and_components(B):
    for g in (O.grid3d_graph(14), O.er_graph(13, 6000, seed=8)):
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        deg = np.diff(g.offsets.astype(np.int64))
        iso = int(np.nonzero(deg == 0)[0][0]) if (deg == 0).any() else 0
        for s in (0, g.n // 3, iso):
            check(g, G, s)
        G.close()


def test_after_updates(B):
    g = O.rmat_graph(11)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    s, d = O.update_rmat(g.n, 3000, 4)
    one = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
    rev = (d.astype(np.uint64) << np.uint64(32)) | s.astype(np.uint64)
    G.insert_batch(np.concatenate([one, rev]))
    off, ids = G.export_csr()
    h = O.Graph(off, ids)
    check(h, G, int(O.bfs_sources(h, 1, seed=2)[0]))
    G.close()
