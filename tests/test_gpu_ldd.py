"""GPU parity of the low-diameter decomposition (byo_ldd) with the oracle
(pinned by the MPX layering properties in test_oracle.py): bit-exact
cluster ids on the reference inputs, RMAT (pull and push steps), a grid,
many-component graphs and other beta values."""
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


def check(g, G, beta=0.2):
    print("ldd", g.n, g.m, beta, flush=True)
    cl, rounds, ms = G.ldd(beta)
    ref, rd = O.ldd(g, beta)
    assert np.array_equal(cl, ref)
    assert rounds == int(rd.max()) or g.n <= 1


@pytest.mark.parametrize("name", ["rMatGraph_J_5_100", "star.txt", "triangles.txt"])
def test_reference_inputs(B, name):
    g = O.read_adjacency_graph(os.path.join(INPUTS, name))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    check(g, G)
    G.close()


@pytest.mark.parametrize("beta", [0.05, 0.2, 1.0])
def test_graphs(B, beta):
    for g in (O.rmat_graph(15), O.grid3d_graph(20), O.er_graph(14, 12000, seed=6)):
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        check(g, G, beta)
        G.close()
