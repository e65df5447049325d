"""GPU parity of the k-core peel (byo_kcore) with the oracle's coreness
(pinned by test_kcore.cc's vectors and the brute-force definition):
bit-exact on the reference unit cases, its input files, RMAT, grid and
many-component graphs, and on a CPMA graph after updates."""
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


def check(g, G):
    print("kcore", g.n, g.m, flush=True)
    core, kmax, rounds, ms = G.kcore()
    ref = O.kcore(g)
    assert np.array_equal(core, ref)
    assert kmax == (int(ref.max()) if g.n else 0)
    assert rounds >= 1 or g.n == 0
    return core


def test_reference_unit_cases(B):
    g = O.sym_csr(3, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    assert list(check(g, G)) == [0, 0, 0]
    G.close()
    e = np.array([[0, 1], [2, 3], [3, 4], [3, 5], [4, 5], [5, 6]], np.uint32)
    g = O.sym_csr(8, e[:, 0], e[:, 1])
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    assert list(check(g, G)) == [1, 1, 1, 2, 2, 2, 1, 0]
    G.close()


@pytest.mark.parametrize("name", ["rMatGraph_J_5_100", "star.txt", "triangles.txt"])
def test_reference_inputs(B, name):
    g = O.read_adjacency_graph(os.path.join(INPUTS, name))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    check(g, G)
    G.close()


@pytest.mark.parametrize("make", ["rmat16", "rmat12_dense", "grid", "er"])
def test_graphs(B, make):
    g = {"rmat16": lambda: O.rmat_graph(16),
         "rmat12_dense": lambda: O.rmat_graph(12, 64),
         "grid": lambda: O.grid3d_graph(16),
         "er": lambda: O.er_graph(14, 20000, seed=2)}[make]()
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    check(g, G)
    check(g, G)  # state is rebuilt per call
    G.close()


def test_after_updates(B):
    g = O.rmat_graph(12)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    s, d = O.update_rmat(g.n, 5000, 6)
    one = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
    rev = (d.astype(np.uint64) << np.uint64(32)) | s.astype(np.uint64)
    G.insert_batch(np.concatenate([one, rev]))
    off, ids = G.export_csr()
    check(O.Graph(off, ids), G)
    G.close()
