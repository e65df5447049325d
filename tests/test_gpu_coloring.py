"""GPU parity of the greedy LLF colouring (byo_coloring) with the oracle's
restatement (pinned in test_oracle.py by hand-worked cases and the
definition): colours bit-exact on hand cases, the reference's input files,
RMAT, a grid, many-component graphs and a CPMA graph after updates, and the
definition checked directly on the smaller ones."""
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


def check(g, G, full_check=True):
    print("coloring", g.n, g.m, flush=True)
    got, rounds, ms = G.coloring()
    ref = O.coloring(g)
    assert np.array_equal(got, ref)
    if full_check:
        assert O.coloring_check(g, got) == 0
    assert g.n == 0 or rounds >= 1
    return got


def test_hand_cases(B):
    t = O.sym_csr(4, np.array([0, 1, 0], np.uint32), np.array([1, 2, 2], np.uint32))
    G = B.CsrGraph.from_csr(t.offsets, t.ids)
    got = check(t, G)
    assert sorted(int(x) for x in got[:3]) == [0, 1, 3] and got[3] == 0
    G.close()
    star = O.sym_csr(6, np.zeros(5, np.uint32), np.arange(1, 6, dtype=np.uint32))
    G = B.CsrGraph.from_csr(star.offsets, star.ids)
    assert list(check(star, G)) == [0, 2, 2, 2, 2, 2]
    G.close()


def test_wide_windows(B):
    # K_4200 minus nothing: the last vertex in order sees 4199 coloured
    # neighbours, so its colour search crosses the 4096-colour window
    k = 4200
    src = np.repeat(np.arange(k, dtype=np.uint32), k)
    dst = np.tile(np.arange(k, dtype=np.uint32), k)
    g = O.sym_csr(k, src, dst)
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    got = check(g, G, full_check=False)
    assert sorted(int(x) for x in got) == list(range(k - 1)) + [k]
    G.close()


@pytest.mark.parametrize("name", ["rMatGraph_J_5_100", "star.txt", "triangles.txt"])
def test_reference_inputs(B, name):
    g = O.read_adjacency_graph(os.path.join(INPUTS, name))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    check(g, G)
    G.close()


def test_rmat_grid_er(B):
    for g in (O.rmat_graph(15), O.rmat_graph(11, 32), O.grid3d_graph(16), O.er_graph(14, 20000, seed=5)):
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        check(g, G, full_check=g.n <= 1 << 11)
        G.close()


def test_after_updates(B):
    g = O.rmat_graph(11)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    s, d = O.update_rmat(g.n, 3000, 4)
    one = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
    rev = (d.astype(np.uint64) << np.uint64(32)) | s.astype(np.uint64)
    G.insert_batch(np.concatenate([one, rev]))
    off, ids = G.export_csr()
    check(O.Graph(off, ids), G, full_check=False)
    G.close()
