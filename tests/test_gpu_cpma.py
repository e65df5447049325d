"""GPU parity of the CPMA container: after every insert_batch / remove_batch
the edge set equals the closed form E u unique(B) / E minus unique(B) as
sorted uint64 keys (bit-exact), degrees and M follow, and BFS / PageRank /
CC rerun on the updated graph agree with the oracle on the same edge set."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import byo_amd
    assert byo_amd.device_count() > 0, "no GPU visible"
    return byo_amd


def keys_of(G):
    off, ids = G.export_csr()
    g = O.Graph(off, ids)
    return O.graph_keys(g), g


def upd_keys(n, count, first=0):
    s, d = O.update_rmat(n, count, first)
    return (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)


def test_build_roundtrip(B):
    g = O.rmat_graph(14)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    k, h = keys_of(G)
    assert np.array_equal(k, O.graph_keys(g))
    assert G.M() == g.m and G.N() == g.n
    assert np.array_equal(G.degrees(), g.degrees())
    C = B.CsrGraph.from_csr(g.offsets, g.ids)
    G2 = B.CpmaGraph.from_graph(C)
    assert np.array_equal(keys_of(G2)[0], O.graph_keys(g))
    assert G.get_memory_size() > 8 * g.m


def test_batches_follow_set_semantics(B):
    """Batch_insert_runner (run_unweighted.h:781-893): sizes 1..1e5 in
    powers of 10, insert then remove the same (shuffled) batch."""
    g = O.rmat_graph(14)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    E = O.graph_keys(g)
    rng = np.random.default_rng(0)
    b = 1
    while b <= 100000:
        batch = upd_keys(g.n, b)
        G.insert_batch(batch)
        E = np.union1d(E, batch)
        k, _ = keys_of(G)
        assert np.array_equal(k, E), ("insert", b)
        assert G.M() == len(E)
        shuffled = batch.copy()
        rng.shuffle(shuffled)
        G.remove_batch(shuffled)
        E = np.setdiff1d(E, batch)
        k, h = keys_of(G)
        assert np.array_equal(k, E), ("remove", b)
        assert np.array_equal(G.degrees(), h.degrees())
        b *= 10


def test_overflow_rebalance_and_growth(B):
    """A batch concentrated on few vertices overflows leaves and forces the
    global rebalance; a large random batch grows the leaf array."""
    g = O.rmat_graph(12)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    E = O.graph_keys(g)
    hub = np.arange(0, g.n, dtype=np.uint64)
    batch = (np.uint64(7) << np.uint64(32)) | hub  # vertex 7 gains n edges
    G.insert_batch(batch)
    E = np.union1d(E, batch)
    assert np.array_equal(keys_of(G)[0], E)
    rng = np.random.default_rng(1)
    big = rng.integers(0, g.n, (400000, 2), dtype=np.uint64)
    bk = (big[:, 0] << np.uint64(32)) | big[:, 1]
    G.insert_batch(bk)
    E = np.union1d(E, bk)
    assert np.array_equal(keys_of(G)[0], E)
    # remove almost everything: shrink path
    G.remove_batch(E[: len(E) - 1000])
    E = E[len(E) - 1000:]
    assert np.array_equal(keys_of(G)[0], E)
    G.remove_batch(E)
    assert G.M() == 0
    assert len(keys_of(G)[0]) == 0
    G.insert_batch(bk[:100])
    assert np.array_equal(keys_of(G)[0], np.unique(bk[:100]))


def test_duplicates_self_loops_and_bad_ids(B):
    g = O.rmat_graph(10)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    E = O.graph_keys(g)
    batch = np.array([(3 << 32) | 3, (3 << 32) | 3, (5 << 32) | 9, (5 << 32) | 9], np.uint64)
    G.insert_batch(batch)
    E = np.union1d(E, batch)
    assert np.array_equal(keys_of(G)[0], E)
    with pytest.raises(B.ByoError) as e:
        G.insert_batch(np.array([(np.uint64(g.n) << np.uint64(32)) | np.uint64(1)], np.uint64))
    assert e.value.code == B.E_INVAL
    assert np.array_equal(keys_of(G)[0], E)  # unchanged after the rejected batch
    G.insert_batch(np.zeros(0, np.uint64))
    G.remove_batch(np.zeros(0, np.uint64))
    assert np.array_equal(keys_of(G)[0], E)


def test_algorithms_after_updates(B):
    g = O.rmat_graph(14)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    # symmetric batches keep the graph undirected, so BFS pull (which reads
    # out-neighbours as in-neighbours on a symmetric Graph<>) stays defined
    one = upd_keys(g.n, 25000)
    rev = ((one & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (one >> np.uint64(32))
    batch = np.concatenate([one, rev])
    G.insert_batch(batch)
    G.remove_batch(np.concatenate([one[:10000], rev[:10000]]))
    _, h = keys_of(G)
    assert np.array_equal(G.degrees(), h.degrees())
    src = 10
    lv = O.bfs_levels(h, src)
    par, st = G.bfs(src)
    assert O.check_bfs_tree(h, src, par, lv) == 0
    p_ref, _ = O.pagerank(h, 0.85, 20, 0.0)
    p, _, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP64)
    assert np.abs(p - p_ref).sum() < 1e-12
    s, _ = G.map_sum()
    assert s == O.map_sum(h)


# ---- several batches as one captured HIP graph (byo_insert_batches_dev) ----

def _graph_insert(B, G, batches):
    keys = np.concatenate(batches)
    offs = np.concatenate([[0], np.cumsum([len(b) for b in batches])]).astype(np.uint64)
    d = B.DeviceBuffer(keys)
    try:
        return G.insert_batches_dev(d.ptr, offs)
    finally:
        d.free()


def test_windowed_rebalance_many_hot_leaves(B):
    """Batches that overflow many leaves at once (RMAT-skewed hot spots and
    one vertex gaining every edge) keep the exact edge set and degrees, one
    call at a time and in a captured sequence."""
    g = O.rmat_graph(13)
    E = O.graph_keys(g)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    for i, b in enumerate([upd_keys(g.n, 30000, 0), upd_keys(g.n, 60000, 30000),
                           (np.uint64(3) << np.uint64(32)) | np.arange(g.n, dtype=np.uint64)]):
        G.insert_batch(b)
        E = np.union1d(E, b)
        k, _ = keys_of(G)
        assert np.array_equal(k, E), i
        assert np.array_equal(G.degrees(), np.bincount((E >> np.uint64(32)).astype(np.int64), minlength=g.n))
    # then removals over rebuilt windows
    rem = upd_keys(g.n, 60000, 30000)
    G.remove_batch(rem)
    E = np.setdiff1d(E, rem)
    assert np.array_equal(keys_of(G)[0], E)


def test_captured_batches_equal_sequential(B):
    """Uneven batches in one graph give the edge set, M and degrees of the
    same batches inserted one call at a time."""
    g = O.rmat_graph(14)
    sizes = [10, 1000, 3, 20000, 500, 7, 4096]
    batches, first = [], 0
    for b in sizes:
        batches.append(upd_keys(g.n, b, first))
        first += b
    G1 = B.CpmaGraph.from_csr(g.offsets, g.ids)
    graphs, ms = _graph_insert(B, G1, batches)
    print("graphs", graphs, "ms", ms, flush=True)
    assert graphs >= 1 and ms > 0
    G2 = B.CpmaGraph.from_csr(g.offsets, g.ids)
    for b in batches:
        G2.insert_batch(b)
    E = O.graph_keys(g)
    for b in batches:
        E = np.union1d(E, b)
    k1, _ = keys_of(G1)
    k2, _ = keys_of(G2)
    assert np.array_equal(k1, E) and np.array_equal(k2, E)
    assert G1.M() == len(E) == G2.M()
    assert np.array_equal(G1.degrees(), G2.degrees())


def test_captured_batches_overflow_and_bad_ids(B):
    """Leaves that would overflow are rebuilt in windows inside the graph; a
    batch that makes the whole array too dense stops the graph, goes through
    the host's global rebalance, and the rest follows in a new graph; a batch
    with an id >= n raises after the batches before it were applied."""
    g = O.rmat_graph(12)
    hub = (np.uint64(5) << np.uint64(32)) | np.arange(g.n, dtype=np.uint64)  # one vertex gains n edges
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    batches = [upd_keys(g.n, 100, 0), hub, upd_keys(g.n, 300, 100)]
    graphs, _ = _graph_insert(B, G, batches)
    assert graphs == 1  # windowed rebalance, no host fallback
    E = O.graph_keys(g)
    for b in batches:
        E = np.union1d(E, b)
    assert np.array_equal(keys_of(G)[0], E)
    assert np.array_equal(G.degrees(), np.bincount((E >> np.uint64(32)).astype(np.int64), minlength=g.n))
    # grow the array by far more than its slack in one batch: global rebalance
    dense = ((np.arange(64, dtype=np.uint64)[:, None] << np.uint64(32)) |
             np.arange(g.n, dtype=np.uint64)[None, :]).ravel()
    batches = [upd_keys(g.n, 50, 900), dense, upd_keys(g.n, 70, 1200)]
    graphs, _ = _graph_insert(B, G, batches)
    assert graphs >= 2
    for b in batches:
        E = np.union1d(E, b)
    assert np.array_equal(keys_of(G)[0], E)
    bad = np.array([(np.uint64(g.n) << np.uint64(32)) | np.uint64(1)], np.uint64)
    later = upd_keys(g.n, 50, 5000)
    ok = upd_keys(g.n, 40, 7000)
    with pytest.raises(B.ByoError):
        _graph_insert(B, G, [ok, bad, later])
    E2 = np.union1d(E, ok)
    assert np.array_equal(keys_of(G)[0], E2)


# ---- algorithms over the PMA itself (slot index + dst words, no CSR copy) ----

def _sym(keys):
    rev = ((keys & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (keys >> np.uint64(32))
    return np.concatenate([keys, rev])


@pytest.mark.parametrize("scale", [12, 14])
def test_algorithms_read_the_pma(B, scale):
    """BFS (one source, 8 sources, forced directions), PageRank FP32/FP64,
    both connectivity algorithms, triangle counting and the edge scan run on
    the CPMA's own leaves -- empty slots at leaf ends, empty leaves after
    removals, rebuilt windows -- and agree with the oracle on the same edge
    set; none of them materialises the CSR export."""
    g = O.rmat_graph(scale)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    E = O.graph_keys(g)
    ins = _sym(upd_keys(g.n, 6000))  # RMAT-skewed: overflows hot leaves
    G.insert_batch(ins)
    E = np.union1d(E, ins)
    # remove every edge of a band of vertices: whole leaves empty out
    band = E[((E >> np.uint64(32)) >= 200) & ((E >> np.uint64(32)) < 900)]
    band = np.union1d(band, _sym(band))
    G.remove_batch(band)
    E = np.setdiff1d(E, band)
    rem = _sym(upd_keys(g.n, 3000, 100))
    G.remove_batch(rem)
    E = np.setdiff1d(E, rem)
    h = O.keys_to_csr(g.n, E)
    G.profile(True)
    G.profile_reset()
    srcs = O.bfs_sources(h, 8, seed=5)
    for src in srcs[:3]:
        lv = O.bfs_levels(h, src)
        par, st = G.bfs(src)
        assert O.check_bfs_tree(h, src, par, lv) == 0, src
        _, dirlog, fs, fd = O.bfs_diropt(h, src)
        assert st["dir"] == list(dirlog) and st["fsize"] == list(fs) and st["fdeg"] == list(fd)
        assert np.array_equal(G.bfs_levels(), lv)
    deg = np.diff(h.offsets)
    for pol in (B.DIR_THRESHOLD, B.DIR_SPARSE, B.DIR_DENSE):
        G.set_direction_policy(pol)
        pars, sts = G.bfs_multi(srcs)
        for j, src in enumerate(srcs):
            lv = O.bfs_levels(h, src)
            assert O.check_bfs_tree(h, src, pars[j], lv) == 0, (pol, src)
            assert sts[j]["reached"] == int((lv != O.NONE).sum())
            assert sts[j]["edges_reached"] == int(deg[lv != O.NONE].sum())
        par, _ = G.bfs(srcs[0])
        assert O.check_bfs_tree(h, srcs[0], par, O.bfs_levels(h, srcs[0])) == 0, pol
    G.set_direction_schedule([0, 3, 3, 1, 0])  # push, forward, forward, pull, push
    par, _ = G.bfs(srcs[1])
    assert O.check_bfs_tree(h, srcs[1], par, O.bfs_levels(h, srcs[1])) == 0
    G.set_direction_schedule([0, 0, 4, 4, 1, 0])  # bitmap pulls on the PMA view
    pars, _ = G.bfs_multi(srcs)
    for j, src in enumerate(srcs):
        assert O.check_bfs_tree(h, src, pars[j], O.bfs_levels(h, src)) == 0, ("bitmap pull", src)
    G.set_direction_policy(B.DIR_THRESHOLD)
    p_ref, _ = O.pagerank(h, 0.85, 20, 0.0)
    p64, _, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP64)
    assert np.abs(p64 - p_ref).sum() < 1e-12
    p32, _, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP32)
    assert np.abs(p32 - p_ref).sum() <= 1e-6  # the stated FP32 tolerance
    # the per-XCD source slices, built from the slot view (holes dropped)
    G.set_tuning("pr_xcd", 1)
    p64, _, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP64)
    assert np.abs(p64 - p_ref).sum() < 1e-12
    p32, _, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP32)
    assert np.abs(p32 - p_ref).sum() <= 1e-6
    G.set_tuning("pr_xcd", 0)
    lab = O.cc(h)
    assert np.array_equal(G.cc()[0], lab)
    assert np.array_equal(G.cc_labelprop()[0], lab)
    assert G.tc()[0] == O.tc(h)
    assert G.map_sum()[0] == O.map_sum(h)
    assert np.array_equal(G.degrees(), deg.astype(np.uint32))
    assert G.profile_get("cpma_csr_export")[0] == 0  # no CSR copy was built
    assert G.profile_get("cpma_index")[0] >= 1
    # the export (used only now) holds the same edge set
    assert np.array_equal(keys_of(G)[0], E)
    assert G.profile_get("cpma_csr_export")[0] == 1
    G.close()
