"""GPU parity of the approximate densest subgraph (byo_densest_subgraph)
with the oracle's restatement (pinned in test_oracle.py by closed forms and
brute force): the same density, bit for bit, and the same round count on
closed-form graphs, the reference's input files, RMAT, a grid,
many-component graphs and a CPMA graph after updates, at the default and a
coarse eps."""
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


def check(g, G, eps=0.001):
    print("densest", g.n, g.m, eps, flush=True)
    d, rounds, ms = G.densest_subgraph(eps)
    rd, rr = O.densest(g, eps)
    assert d == rd and rounds == rr
    return d


def test_closed_forms(B):
    k = 9
    src = np.repeat(np.arange(k, dtype=np.uint32), k)
    dst = np.tile(np.arange(k, dtype=np.uint32), k)
    g = O.sym_csr(k, src, dst)
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    assert check(g, G) == k - 1
    G.close()
    s = 7
    g = O.sym_csr(s + 1, np.zeros(s, np.uint32), np.arange(1, s + 1, dtype=np.uint32))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    assert check(g, G) == 2 * s / (s + 1)
    G.close()


@pytest.mark.parametrize("name", ["rMatGraph_J_5_100", "star.txt", "triangles.txt"])
def test_reference_inputs(B, name):
    g = O.read_adjacency_graph(os.path.join(INPUTS, name))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    check(g, G)
    check(g, G, 0.5)
    G.close()


def test_rmat_grid_er(B):
    for g in (O.rmat_graph(15), O.rmat_graph(12, 32), O.grid3d_graph(16), O.er_graph(14, 20000, seed=5)):
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        check(g, G)
        check(g, G, 0.25)
        G.close()


def test_after_updates(B):
    g = O.rmat_graph(11)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    s, d = O.update_rmat(g.n, 3000, 4)
    one = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
    rev = (d.astype(np.uint64) << np.uint64(32)) | s.astype(np.uint64)
    G.insert_batch(np.concatenate([one, rev]))
    off, ids = G.export_csr()
    check(O.Graph(off, ids), G)
    G.close()
