"""GPU parity of the spanner (byo_spanner) with the oracle's restatement
(pinned in test_oracle.py against its definition over the LDD and G's
components): the same sorted edge set, bit for bit, on the reference's
input files, RMAT, a grid, many-component graphs and a CPMA graph after
updates, for k = 4 (run_all's) and k = 2."""
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


def check(g, G, k=4, full_check=True):
    print("spanner", g.n, g.m, k, flush=True)
    got, ms = G.spanner(k)
    ref = O.spanner(g, k)
    assert np.array_equal(got, ref)
    if full_check:
        assert O.spanner_check(g, got, k) == 0
    return got


@pytest.mark.parametrize("name", ["rMatGraph_J_5_100", "star.txt", "triangles.txt"])
def test_reference_inputs(B, name):
    g = O.read_adjacency_graph(os.path.join(INPUTS, name))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    check(g, G)
    check(g, G, 2)
    G.close()


def test_tiny_graphs(B):
    for n in (1, 2):
        g = O.sym_csr(n, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        got, _ = G.spanner()
        assert len(got) == 0
        G.close()


def test_rmat_grid_er(B):
    for g in (O.rmat_graph(15), O.rmat_graph(11, 32), O.grid3d_graph(16), O.er_graph(14, 20000, seed=5)):
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        check(g, G, full_check=g.n <= 1 << 11)
        check(g, G, 2, full_check=False)
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


def test_error_paths(B):
    import ctypes as C
    g = O.rmat_graph(9)
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    lib = B.lib()
    cnt, ms = C.c_uint64(), C.c_double()
    # k = 0 is rejected; a buffer smaller than the spanner is rejected with
    # the size needed reported
    assert lib.byo_spanner(G._h, 0, None, 0, C.byref(cnt), C.byref(ms)) == B.E_INVAL
    assert lib.byo_spanner(G._h, 4, None, 0, C.byref(cnt), C.byref(ms)) == B.OK
    need = cnt.value
    small = np.empty(max(need - 1, 1), np.uint64)
    rc = lib.byo_spanner(G._h, 4, small.ctypes.data_as(C.c_void_p), need - 1, C.byref(cnt), C.byref(ms))
    assert rc == B.E_INVAL and cnt.value == need
    d = C.c_double()
    assert lib.byo_densest_subgraph(G._h, -1.0, C.byref(d), None, None) == B.E_INVAL
    assert lib.byo_coloring(G._h, None, None, None) == B.E_INVAL
    assert lib.byo_mis(G._h, None, None, None) == B.E_INVAL
    G.close()
