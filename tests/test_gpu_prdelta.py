"""GPU parity of PageRankDelta (byo_pagerank_delta) with the oracle's
restatement (pinned in test_oracle.py): the same rounds and an L1 distance
of at most 1e-9 on the reference inputs, RMAT (dense rounds then sparse
pushes), a grid and many-component graphs; one round equals the first
Jacobi iterate, and with every vertex active r rounds equal r Jacobi
iterations of the FP64 PageRank; the same on the degree-relabelled layout."""
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


def check(g, G, eps=1e-4, leps=0.01, iters=20):
    print("prdelta", g.n, g.m, eps, leps, iters, flush=True)
    p, r, ms = G.pagerank_delta(eps, leps, iters)
    ref, rr = O.pagerank_delta(g, eps, leps, iters)
    assert r == rr
    assert np.abs(p - ref).sum() <= 1e-9
    return p, r


@pytest.mark.parametrize("name", ["rMatGraph_J_5_100", "star.txt", "triangles.txt"])
def test_reference_inputs(B, name):
    g = O.read_adjacency_graph(os.path.join(INPUTS, name))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    check(g, G)
    check(g, G, 1e-9, 0.001, 60)
    G.close()


def test_rmat_grid_er(B):
    for g in (O.rmat_graph(15), O.grid3d_graph(16), O.er_graph(13, 9000, seed=3)):
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        _, r = check(g, G)
        assert r >= 2
        check(g, G, 1e-7, 0.001, 50)  # deeper: more sparse rounds
        G.close()


def test_jacobi_pins(B):
    g = O.rmat_graph(13)
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    p, r, _ = G.pagerank_delta(0.0, 0.01, 1)
    assert r == 1 and np.abs(p - O.pagerank(g, 0.85, 1)[0]).sum() <= 1e-12
    p, r, _ = G.pagerank_delta(0.0, -1e300, 15)
    assert r == 15 and np.abs(p - O.pagerank(g, 0.85, 15)[0]).sum() <= 1e-10
    G.close()


def test_relabelled_layout(B):
    # the rounds on the degree-relabelled graph, ranks mapped back
    for g in (O.rmat_graph(15), O.er_graph(13, 9000, seed=3)):
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        G.set_tuning("pr_relabel", 1)
        check(g, G)
        check(g, G, 1e-7, 0.001, 50)
        G.close()
