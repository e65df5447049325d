"""GPU checks of the direction-selector MLP kernels against the CPU oracle's
double-precision forward/backward (orc_mlp_forward_backward), plus a short
training run whose loss must fall."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import byo_amd
    assert byo_amd.device_count() > 0, "no GPU visible"
    return byo_amd


def test_gradients_match_oracle(B):
    rng = np.random.default_rng(0)
    P = (rng.standard_normal(O.MLP_NPARAM) * 0.3).astype(np.float32)
    x = rng.standard_normal((300, 6)).astype(np.float32)
    y = rng.integers(0, 3, 300).astype(np.int32)
    T = B.Trainer(P)
    loss, g, lg = T.grad(x, y)
    lref, gref, lgref = O.mlp_forward_backward(P, x, y)
    assert abs(loss - lref) < 1e-5 * max(1, abs(lref))
    assert np.allclose(lg, lgref, rtol=1e-4, atol=1e-5)
    assert np.allclose(g, gref, rtol=1e-3, atol=2e-6), np.abs(g - gref).max()
    T.close()


def test_training_reduces_loss_and_predict(B):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((512, 6)).astype(np.float32)
    y = (x[:, 0] > 0).astype(np.int32) + (x[:, 1] > 1).astype(np.int32)
    P = (rng.standard_normal(O.MLP_NPARAM) * 0.2).astype(np.float32)
    T = B.Trainer(P, lr=1e-2)
    first = T.grad(x, y)[0]
    for _ in range(300):
        T.step(x, y)
    last = T.grad(x, y)[0]
    assert last < 0.5 * first
    pred = T.predict(x)
    assert (pred == y).mean() > 0.9
    # device params reproduce the host oracle's logits
    lref, _, lg = O.mlp_forward_backward(T.params(), x, y)
    assert abs(lref - last) < 1e-4
    T.close()


@pytest.mark.parametrize("engine", ["single", "batch"])
def test_learned_policy_is_bit_exact(B, engine):
    """The committed selector of each traversal engine keeps BFS levels and
    parent trees valid on graphs it never saw (a 48^3 grid, RMAT-16)."""
    import os
    from byo_amd import selector as S
    if not os.path.exists(S.ENGINES[engine][1]):
        pytest.skip("no trained %s-engine selector weights committed" % engine)
    for G, g in ((B.CsrGraph.grid3d(48), O.grid3d_graph(48)), (B.CsrGraph.rmat(16), O.rmat_graph(16))):
        S.load(G)
        G.set_direction_policy(B.DIR_MLP)
        srcs = O.bfs_sources(g, 8, seed=21)
        for s in srcs[:3]:
            par, st = G.bfs(s)
            lv = O.bfs_levels(g, s)
            assert np.array_equal(G.bfs_levels(), lv)
            assert O.check_bfs_tree(g, s, par, lv) == 0
        for k in (8, 64):
            srcs = O.bfs_sources(g, k, seed=21)
            par, _ = G.bfs_multi(srcs)
            for j, s in enumerate(srcs):
                assert O.check_bfs_tree(g, s, par[j], O.bfs_levels(g, s)) == 0
        G.close()


HELD_OUT = {"rmat24_ef32": lambda B: B.CsrGraph.rmat(24, 32), "er23_deg8": lambda B: B.CsrGraph.erdos_renyi(23, 4 << 23),
            "grid256": lambda B: B.CsrGraph.grid3d(256)}


@pytest.mark.parametrize("graph", sorted(HELD_OUT))
@pytest.mark.parametrize("engine", ["single", "batch"])
def test_learned_policy_on_held_out_graphs(B, engine, graph):
    """Config C5: the selector, trained on RMAT scale 18-22 timings, against
    the M/20 rule on the held-out graphs (RMAT scale 24 ef 32, ER n = 2^23
    degree 8, the 256^3 grid), with the shared traversal at the benchmark's
    8 sources and the batched job's 64.  Levels are identical under both
    policies; where the network changes a direction the traversal is
    faster (speedup >= 1); where it keeps the rule's directions the work is
    the same and the time differs only by the network's evaluation (bounded
    at 3 %); on RMAT-24 ef 32 with 64 sources, where the rule pulls a level
    too early, the gain reaches the 1.2x target."""
    import os
    from byo_amd import selector as S
    if not os.path.exists(S.ENGINES[engine][1]):
        pytest.skip("no trained %s-engine selector weights committed" % engine)
    G = HELD_OUT[graph](B)
    widths = (8, 64) if engine == "batch" else (4,)
    for k in widths:
        srcs = S.sources(G.degrees(), k, seed=5)
        res = S.compare(G, srcs, engine=engine, repeats=5)
        print("selector", engine, graph, "sources", k, "speedup %.3f" % res["speedup"],
              "rule %.3f ms learned %.3f ms" % (res["threshold"], res["learned"]),
              "dirs", res["learned_dirs"][:12], "rule dirs", res["threshold_dirs"][:12], flush=True)
        assert res["levels_identical"]
        if res["same_directions"]:
            assert res["speedup"] >= 0.97
        else:
            assert res["speedup"] >= 1.0
        if engine == "batch" and graph == "rmat24_ef32" and k == 64:
            assert res["speedup"] >= 1.2
    G.close()
