"""GPU parity at the sizes the benchmark's headline numbers are measured on.

* The device update stream (byo_update_rmat_keys) equals the oracle's
  restatement of Batch_insert_runner's rMat stream (run_unweighted.h:730-806)
  bit for bit, at n = 2^16 and 2^24 and from offsets first > 0.
* CPMA insert_batch / remove_batch at b = 1e6 and 1e7 (device-resident keys,
  the benchmark's path: full-grid radix sort, windowed rebalances over
  thousands of leaves, leaf-array growth) keep the edge set equal to E u B /
  E minus B after every batch (Batch_insert_runner, run_unweighted.h:781-893),
  on RMAT-18, RMAT-20 and the benchmark's RMAT-24, where BFS and PageRank
  then rerun on the PMA after a symmetric batch.
* RMAT scale 22 ef 16 (the benchmark graph): the device generator equals the
  oracle's graph, the 8-source shared traversal gives valid parent trees
  with the oracle's levels for every source, CC labels and the triangle count
  equal the oracle's, and FP32-Kahan PageRank is within 1e-6 L1 of the
  oracle's FP64 vector after 20 iterations.
* Triangle counting with two hubs of degree 70,000, above the 65,535 clamp
  of the ranking's 16-bit degree table, so the clamped tie order runs.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import byo_amd
    assert byo_amd.device_count() > 0, "no GPU visible"
    return byo_amd


def oracle_keys(n, count, first=0):
    s, d = O.update_rmat(n, count, first)
    return (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)


@pytest.mark.parametrize("n,count,first", [(1 << 16, 1_000_000, 0), (1 << 16, 50_000, 123_457),
                                           (1 << 24, 2_000_000, 0), (1 << 24, 100_000, 9_999_991),
                                           ((1 << 20) + 5, 300_000, 17)])
def test_update_stream_matches_oracle(B, n, count, first):
    dev = B.update_rmat_keys(n, count, first)
    ref = oracle_keys(n, count, first)
    assert dev.dtype == np.uint64 and len(dev) == count
    assert np.array_equal(dev, ref)


def _cpma_keys(G):
    off, ids = G.export_csr()
    return O.graph_keys(O.Graph(off, ids)), off


@pytest.mark.parametrize("scale,sizes", [(18, (1_000_000, 10_000_000)), (20, (1_000_000, 10_000_000))])
def test_cpma_large_batches_set_semantics(B, scale, sizes):
    """The benchmark's update path at its batch sizes: keys generated on the
    device, resident in HBM, inserted and removed (remove in a shuffled
    order, as the runner shuffles) with the edge set checked after each."""
    print("cpma large batches: RMAT-%d, sizes %s" % (scale, sizes), flush=True)
    H = B.CsrGraph.rmat(scale, 16)
    G = B.CpmaGraph.from_graph(H)
    off, ids = H.export_csr()
    H.close()
    E = O.graph_keys(O.Graph(off, ids))
    del off, ids
    n = G.N()
    rng = np.random.default_rng(scale)
    try:
        for b in sizes:
            keys = B.update_rmat_keys(n, b)
            d = B.DeviceBuffer(keys)
            try:
                print("  insert b=%d" % b, flush=True)
                G.insert_batch_dev(d.ptr, b)
                E = np.union1d(E, keys)
                k, o = _cpma_keys(G)
                assert np.array_equal(k, E), ("insert", b)
                assert G.M() == len(E)
                assert np.array_equal(G.degrees(), np.diff(o).astype(np.uint32))
            finally:
                d.free()
            sh = keys.copy()
            rng.shuffle(sh)
            d = B.DeviceBuffer(sh)
            try:
                print("  remove b=%d" % b, flush=True)
                G.remove_batch_dev(d.ptr, b)
                E = np.setdiff1d(E, keys, assume_unique=False)
                k, o = _cpma_keys(G)
                assert np.array_equal(k, E), ("remove", b)
                assert G.M() == len(E)
                assert np.array_equal(G.degrees(), np.diff(o).astype(np.uint32))
            finally:
                d.free()
    finally:
        G.close()


def test_cpma_rmat24_batches_and_rerun(B):
    """C3 at the benchmark's own scale, RMAT-24 ef 16 as a CPMA.  First, as
    the benchmark reruns it, a symmetric 1e7-key batch is inserted and the
    8-source shared BFS and FP32 PageRank run on the PMA itself (tuning
    applied, per-XCD slices cut from the slot view) against the oracle on
    the updated graph, and the batch is removed again.  Then device
    update-stream batches of 1e6 and 1e7 directed keys are inserted and
    removed (remove shuffled), with the edge set equal to the oracle's
    sorted-array E u B / E minus B after each (gbbs/graph.h:3071-3124).  The
    rerun comes first because removing a directed batch also removes the
    batch's keys that were edges before, which leaves their reverse edges
    alone: the graph is no longer symmetric, and a pull reads out-lists as
    in-lists."""
    print("cpma RMAT-24: build", flush=True)
    H = B.CsrGraph.rmat(24, 16)
    G = B.CpmaGraph.from_graph(H)
    off, ids = H.export_csr()
    H.close()
    E = O.csr_keys(O.Graph(off, ids))
    del off, ids
    n = G.N()
    rng = np.random.default_rng(24)

    def check(step):
        o, i = G.export_csr()
        g = O.Graph(o, i)
        k = O.csr_keys(g)
        assert len(k) == len(E) and np.array_equal(k, E), step
        assert G.M() == len(E), step
        assert np.array_equal(G.degrees(), np.diff(o).astype(np.uint32)), step
        return g

    def update(keys, insert):
        d = B.DeviceBuffer(keys)
        try:
            if insert:
                G.insert_batch_dev(d.ptr, len(keys))
            else:
                G.remove_batch_dev(d.ptr, len(keys))
        finally:
            d.free()

    try:
        half = B.update_rmat_keys(n, 5_000_000)
        sym = np.concatenate([half, ((half & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (half >> np.uint64(32))])
        print("  symmetric insert %d keys, rerun" % len(sym), flush=True)
        update(sym, True)
        E = O.batch_insert(E, sym)
        g = check("symmetric insert")
        G.apply_tuning()
        srcs = O.bfs_sources(g, 8, seed=4)
        par, st = G.bfs_multi(srcs)
        for j, s in enumerate(srcs):
            lv = O.bfs_levels(g, s)
            assert O.check_bfs_tree(g, s, par[j], lv) == 0, s
            assert st[j]["reached"] == int((lv != O.NONE).sum())
        ref, _ = O.pagerank(g, 0.85, 20)
        p, _, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP32)
        err = float(np.abs(p - ref).sum())
        print("  PMA FP32 PageRank L1 vs FP64 oracle", err, flush=True)
        assert err <= 1e-6
        del g, par, p, ref
        update(sym, False)
        E = O.batch_remove(E, sym)
        check("symmetric remove")
        for b in (1_000_000, 10_000_000):
            keys = B.update_rmat_keys(n, b)
            print("  insert b=%d" % b, flush=True)
            update(keys, True)
            E = O.batch_insert(E, keys)
            check(("insert", b))
            sh = keys.copy()
            rng.shuffle(sh)
            print("  remove b=%d" % b, flush=True)
            update(sh, False)
            E = O.batch_remove(E, keys)
            check(("remove", b))
    finally:
        G.close()


# ---- the benchmark graph: RMAT scale 22, edge factor 16 --------------------

@pytest.fixture(scope="module")
def s22(B):
    print("RMAT-22: device graph and oracle graph", flush=True)
    G = B.CsrGraph.rmat(22, 16)
    g = O.rmat_graph(22, 16)
    yield G, g
    G.close()


def test_s22_generator_matches_oracle(B, s22):
    G, g = s22
    off, ids = G.export_csr()
    assert np.array_equal(off, g.offsets)
    assert np.array_equal(ids, g.ids)


def test_s22_bench_bfs_trees_and_levels(B, s22):
    """The benchmark's step as it runs (tuning applied, learned selector
    loaded): every source's parents form a valid BFS tree with the oracle's
    levels, reach and edge counts equal, under the selector and the rule."""
    from byo_amd import selector as S
    G, g = s22
    G.apply_tuning()
    S.load(G)
    srcs = O.bfs_sources(g, 8, seed=1)
    deg = np.diff(g.offsets)
    levels = {s: O.bfs_levels(g, s) for s in srcs}
    for pol in (B.DIR_MLP, B.DIR_THRESHOLD):
        G.set_direction_policy(pol)
        par, st = G.bfs_multi(srcs)
        for j, s in enumerate(srcs):
            lv = levels[s]
            assert O.check_bfs_tree(g, s, par[j], lv) == 0, (pol, s)
            reached = lv != O.NONE
            assert st[j]["reached"] == int(reached.sum())
            assert st[j]["edges_reached"] == int(deg[reached].sum())
            assert st[j]["levels"] == int(lv[reached].max()) + 1
    G.set_direction_policy(B.DIR_THRESHOLD)


def test_s22_cc_tc_pagerank(B, s22):
    G, g = s22
    lab, _ = G.cc()
    assert np.array_equal(lab, O.cc(g))
    cnt, _ = G.tc()
    assert cnt == O.tc(g)
    print("RMAT-22 triangles", cnt, "components", len(np.unique(lab)), flush=True)
    pr, _ = O.pagerank(g, 0.85, 20, 0.0)
    for relabel, pb in ((0, 0), (1, 0), (1, 1)):
        G.set_tuning("pr_relabel", relabel)
        G.set_tuning("pr_pb", pb)
        p32, _, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP32)
        err = float(np.abs(p32 - pr).sum())
        print("relabel", relabel, "propagation blocking", pb, "FP32 L1 vs FP64 oracle", err, flush=True)
        assert err <= 1e-6  # the stated FP32 tolerance
        p64, _, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP64)
        assert float(np.abs(p64 - pr).sum()) <= 1e-12
    G.apply_tuning()


def test_tc_hubs_above_degree_clamp(B):
    """Two hubs adjacent to each other and to the same 70,000 leaves (degree
    70,001 each, both clamped to 65,535 in the ranking's degree table, so
    their order falls to the id tie-break), plus leaf-leaf edges and a
    separate K5: every triangle counted once."""
    leaves = 70_000
    n = leaves + 2 + 5
    L = np.arange(2, 2 + leaves, dtype=np.uint32)
    src = [np.zeros(leaves, np.uint32), np.ones(leaves, np.uint32), np.array([0], np.uint32)]
    dst = [L, L, np.array([1], np.uint32)]
    rng = np.random.default_rng(5)
    a = rng.integers(2, 2 + leaves, 3000).astype(np.uint32)
    b = rng.integers(2, 2 + leaves, 3000).astype(np.uint32)
    src.append(a)
    dst.append(b)
    k5 = np.arange(n - 5, n, dtype=np.uint32)
    for i in range(5):
        for j in range(i + 1, 5):
            src.append(k5[i:i + 1])
            dst.append(k5[j:j + 1])
    g = O.sym_csr(n, np.concatenate(src), np.concatenate(dst))
    assert g.degrees()[0] == g.degrees()[1] and g.degrees()[0] > 65535
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    cnt, _ = G.tc()
    # independent count: hubs 0 and 1 close a triangle with each leaf, each
    # leaf-leaf edge closes one with each hub, K5 holds 10, plus triangles
    # among the leaf-leaf edges themselves (brute force over their sets)
    adj = {}
    for u, v in zip(a.tolist(), b.tolist()):
        if u != v:
            adj.setdefault(u, set()).add(v)
            adj.setdefault(v, set()).add(u)
    e = sum(len(s) for s in adj.values()) // 2
    t_ll = sum(len(adj[u] & adj[v]) for u in adj for v in adj[u] if u < v) // 3
    assert g.m == 4 * leaves + 2 + 20 + 2 * e
    expect = leaves + 2 * e + 10 + t_ll
    assert O.tc(g) == expect
    assert cnt == expect
    G.close()
