"""GPU parity of the batched job (byo_batch_job: up to 64 BFS sources in one
shared traversal, then up to 64 personalized-PageRank seeds, in one
persistent launch) and of the persistent personalized PageRank
(byo_ppr_multi): every BFS tree valid with the oracle's levels, every PPR
vector within 1e-6 L1 of the oracle's FP64 Jacobi PPR, and the job's PPR
bit-identical to byo_ppr_multi (same per-lane summation order)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import byo_amd
    assert byo_amd.device_count() > 0, "no GPU visible"
    return byo_amd


@pytest.fixture(scope="module")
def s16(B):
    g = O.rmat_graph(16)
    return g, B.CsrGraph.from_csr(g.offsets, g.ids)


def check_trees(g, srcs, par, stats):
    deg = np.diff(g.offsets.astype(np.int64))
    for j, s in enumerate(srcs):
        lv = O.bfs_levels(g, s)
        assert O.check_bfs_tree(g, s, par[j], lv) == 0, "source %d" % s
        assert stats[j]["reached"] == int((lv != O.NONE).sum())
        assert stats[j]["edges_reached"] == int(deg[lv != O.NONE].sum())


def check_ppr(g, seeds, p, iters):
    for j, s in enumerate(seeds):
        ref = O.ppr(g, s, 0.85, iters)
        assert np.abs(p[j].astype(np.float64) - ref).sum() <= 1e-6, "seed %d" % s


@pytest.mark.parametrize("kb,kp", [(64, 32), (8, 5), (33, 64), (1, 1)])
def test_job_matches_oracle(B, s16, kb, kp):
    g, G = s16
    srcs = O.bfs_sources(g, kb, seed=3 + kb)
    seeds = O.bfs_sources(g, kp, seed=7 + kp)
    print("job", kb, kp, flush=True)
    par, stats, p, ms = G.batch_job(srcs, seeds, 0.85, 30)
    assert ms > 0
    check_trees(g, srcs, par, stats)
    check_ppr(g, seeds, p, 30)
    p2, _ = G.ppr(seeds, 0.85, 30)
    assert np.array_equal(p, p2)


def test_job_one_side_only(B, s16):
    g, G = s16
    seeds = O.bfs_sources(g, 3, seed=1)
    par, stats, p, _ = G.batch_job([], seeds, 0.85, 12)
    check_ppr(g, seeds, p, 12)
    srcs = O.bfs_sources(g, 10, seed=2)
    par, stats, p, _ = G.batch_job(srcs, [], 0.85, 30)
    check_trees(g, srcs, par, stats)
    assert np.array_equal(G.bfs_levels(), O.bfs_levels(g, srcs[-1]))


def test_ppr_persistent_fixtures(B):
    import os
    inputs = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")
    for f in ("rMatGraph_J_5_100", "star.txt"):
        g = O.read_adjacency_graph(os.path.join(inputs, f))
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        seeds = list(range(min(g.n, 7)))
        p, _ = G.ppr(seeds, 0.85, 20)
        check_ppr(g, seeds, p, 20)
        G.close()
