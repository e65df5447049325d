"""GPU parity of the batched BFS (up to 64 sources sharing one traversal per
launch, byo_bfs_multi): for every source, a valid parent tree whose levels
equal the oracle's single-source BFS levels, and exact per-source reached
vertex and edge counts -- across mask widths (<= 8, <= 32, <= 64 sources),
chunking above 64, duplicate sources, forced and learned directions, and
graphs with isolated vertices and many components."""
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


@pytest.fixture(scope="module")
def s16(B):
    g = O.rmat_graph(16)
    return g, B.CsrGraph.from_csr(g.offsets, g.ids)


def check_batch(g, G, srcs):
    print("batch", g.n, g.m, len(srcs), flush=True)
    par, stats = G.bfs_multi(srcs)
    deg = np.diff(g.offsets.astype(np.int64))
    for j, s in enumerate(srcs):
        lv = O.bfs_levels(g, s)
        assert O.check_bfs_tree(g, s, par[j], lv) == 0, "source %d (#%d)" % (s, j)
        reached = lv != O.NONE
        assert stats[j]["reached"] == int(reached.sum())
        assert stats[j]["edges_reached"] == int(deg[reached].sum())
        assert stats[j]["levels"] == int(lv[reached].max()) + 1
    return par, stats


@pytest.mark.parametrize("k", [2, 8, 9, 32, 33, 64])
def test_batch_widths(B, s16, k):
    g, G = s16
    check_batch(g, G, O.bfs_sources(g, k, seed=40 + k))


def test_batch_more_than_64_sources(B, s16):
    g, G = s16
    srcs = O.bfs_sources(g, 70, seed=5)
    check_batch(g, G, srcs)
    # the last source's tree is the one byo_bfs_levels reports
    assert np.array_equal(G.bfs_levels(), O.bfs_levels(g, srcs[-1]))


def test_batch_duplicates_and_isolated(B, s16):
    g, G = s16
    deg = np.diff(g.offsets.astype(np.int64))
    iso = [int(v) for v in np.nonzero(deg == 0)[0][:3]]
    srcs = [10, 10] + iso + [10] + O.bfs_sources(g, 3, seed=2)
    check_batch(g, G, srcs)


@pytest.mark.parametrize("policy", [2, 3])
def test_batch_forced_directions(B, s16, policy):
    g, G = s16
    G.set_direction_policy(policy)
    try:
        check_batch(g, G, O.bfs_sources(g, 8, seed=9))
        check_batch(g, G, O.bfs_sources(g, 40, seed=9))
    finally:
        G.set_direction_policy(0)


def test_batch_schedules(B, s16):
    g, G = s16
    try:
        for srcs in (O.bfs_sources(g, 16, seed=11), O.bfs_sources(g, 5, seed=12), O.bfs_sources(g, 50, seed=13)):
            # 4: bitmap pull (union-frontier bitmap before the mask gather)
            for sched in ([0, 0, 1, 1, 3, 0], [3, 3, 3, 3, 3, 3], [1, 0, 1, 0, 1, 0, 1], [255, 255, 3, 0, 1],
                          [4] * 12, [0, 0, 4, 1, 4, 0, 4], [4, 3, 4, 0, 1, 4, 3, 4], [0, 0, 0, 4, 4, 4, 0, 0]):
                G.set_direction_schedule(sched)
                _, stats = check_batch(g, G, srcs)
                dl = max((s["dir"] for s in stats), key=len)  # the shared per-level record
                for i, d in enumerate(sched):
                    if d != 255 and i < len(dl) and dl[i] != 2:
                        assert dl[i] == d
        # the one-source engine runs a scheduled bitmap pull as its (bitmap) pull
        G.set_direction_schedule([0, 4, 4, 0, 4])
        src = O.bfs_sources(g, 1, seed=14)[0]
        par, st = G.bfs(src)
        assert O.check_bfs_tree(g, src, par, O.bfs_levels(g, src)) == 0
        assert all(d in (0, 1, 2) for d in st["dir"][:5])
    finally:
        G.set_direction_policy(0)


def test_batch_fixtures_and_grid(B):
    graphs = [O.read_adjacency_graph(os.path.join(INPUTS, f)) for f in ("rMatGraph_J_5_100", "star.txt")]
    graphs.append(O.grid3d_graph(12))
    graphs.append(O.er_graph(12, 3000, seed=4))  # many small components
    for g in graphs:
        G = B.CsrGraph.from_csr(g.offsets, g.ids)
        k = min(g.n, 12)
        srcs = [int(v) for v in np.linspace(0, g.n - 1, k).astype(np.int64)]
        check_batch(g, G, srcs)
        G.close()


def test_batch_no_edges(B):
    g = O.sym_csr(5, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    par, stats = G.bfs_multi([1, 3, 1])
    assert list(par[0]) == [O.NONE, 1, O.NONE, O.NONE, O.NONE]
    assert list(par[1]) == [O.NONE, O.NONE, O.NONE, 3, O.NONE]
    assert [s["reached"] for s in stats] == [1, 1, 1]
    G.close()


def test_batch_matches_sequential_engine(B, s16):
    """The shared traversal and one-source-at-a-time traversals give the same
    levels and counts."""
    g, G = s16
    srcs = O.bfs_sources(g, 8, seed=17)
    _, st_b = G.bfs_multi(srcs, want_parents=False)
    G.set_tuning("bfs_engine", 1)
    try:
        par_s, st_s = G.bfs_multi(srcs)
    finally:
        G.set_tuning("bfs_engine", 0)
    for j, s in enumerate(srcs):
        assert st_b[j]["reached"] == st_s[j]["reached"]
        assert st_b[j]["edges_reached"] == st_s[j]["edges_reached"]
        assert st_b[j]["levels"] == st_s[j]["levels"]
        assert O.check_bfs_tree(g, s, par_s[j], O.bfs_levels(g, s)) == 0


def test_batch_follows_updates(B):
    """Batched BFS on a CPMA graph after symmetric insert and remove batches."""
    g = O.rmat_graph(12)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    s, d = O.update_rmat(g.n, 4000, 9)
    one = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
    rev = (d.astype(np.uint64) << np.uint64(32)) | s.astype(np.uint64)
    batch = np.concatenate([one, rev])  # symmetric: the pull reads out-lists as in-lists
    for step in ("base", "insert", "remove"):
        if step == "insert":
            G.insert_batch(batch)
        elif step == "remove":
            G.remove_batch(np.concatenate([one[: len(one) // 2], rev[: len(rev) // 2]]))
        off, ids = G.export_csr()
        h = O.Graph(off, ids)
        print(step, h.m, flush=True)
        check_batch(h, G, O.bfs_sources(h, 9, seed=3))
    G.close()


def test_batch_learned_policy(B, s16):
    from byo_amd import selector as S
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained selector weights")
    g, G = s16
    S.load(G)
    G.set_direction_policy(B.DIR_MLP)
    try:
        check_batch(g, G, O.bfs_sources(g, 8, seed=23))
        check_batch(g, G, O.bfs_sources(g, 64, seed=23))
    finally:
        G.set_direction_policy(0)
