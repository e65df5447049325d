"""Every tuning knob value the autotuner may pick keeps results exact: BFS
trees valid with the oracle's levels for every several-source engine,
PageRank within 1e-6 L1 of the oracle's FP64 vector, personalized PageRank
within 1e-6 L1 of the oracle's; bad keys and values are rejected."""
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
def s15(B):
    g = O.rmat_graph(15)
    return g, B.CsrGraph.from_csr(g.offsets, g.ids)


def check_bfs(g, G, srcs):
    par, _ = G.bfs_multi(srcs)
    for j, s in enumerate(srcs):
        assert O.check_bfs_tree(g, s, par[j], O.bfs_levels(g, s)) == 0


@pytest.mark.parametrize("key,values", [("coop_blocks_per_cu", [1, 2, 3, 4]), ("pieces", [1, 4, 16, 32, 64]),
                                        ("pull_dyn", [0, 1, 4, 16, 64]), ("coop_grid_sqrt_div", [0, 1, 14, 1000]),
                                        ("bfs_engine", [0, 1, 2])])
def test_bfs_knobs(B, s15, key, values):
    g, G = s15
    srcs = O.bfs_sources(g, 6, seed=4)
    try:
        for v in values:
            print(key, v, flush=True)
            G.set_tuning(key, v)
            assert G.get_tuning(key) == v
            check_bfs(g, G, srcs)
            par, _ = G.bfs(srcs[0])
            assert O.check_bfs_tree(g, srcs[0], par, O.bfs_levels(g, srcs[0])) == 0
    finally:
        G.set_tuning(key, {"coop_blocks_per_cu": 4, "pieces": 8, "pull_dyn": 8, "coop_grid_sqrt_div": 14, "bfs_engine": 0}[key])


@pytest.mark.parametrize("key,values", [("pr_lanes", [4, 8, 16, 32, 0]), ("pr_grid", [256, 2048, 16384]),
                                        ("pr_seg_log2", [12, 40]), ("pr_xcd_vw", [0, 4, 24, 1024]),
                                        ("pr_pb", [1, 0])])
def test_pagerank_knobs(B, s15, key, values):
    """pr_xcd_vw moves the per-XCD slice bounds (1024: nearly equal vertex
    counts, the hubs' slices holding most edges); the slices are rebuilt
    when it changes."""
    g, G = s15
    ref, _ = O.pagerank(g, 0.85, 20)
    if key == "pr_xcd_vw":
        G.set_tuning("pr_relabel", 1)
        G.set_tuning("pr_xcd", 1)
    try:
        for v in values:
            print(key, v, flush=True)
            G.set_tuning(key, v)
            for prec in (B.PR_FP32, B.PR_FP64):
                p, _, _ = G.pagerank(0.85, 20, 0.0, prec)
                assert np.abs(p - ref).sum() <= 1e-6
    finally:
        G.set_tuning(key, {"pr_lanes": 0, "pr_grid": 4096, "pr_seg_log2": 40, "pr_xcd_vw": 0, "pr_pb": 0}[key])
        G.set_tuning("pr_relabel", 0)
        G.set_tuning("pr_xcd", 0)


@pytest.mark.parametrize("precision", ["PR_FP32", "PR_FP64"])
@pytest.mark.parametrize("relabel,xcd,pb", [(1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1)])
def test_pagerank_layouts(B, s15, precision, relabel, xcd, pb):
    """PageRank on the degree-relabelled copy (ranks mapped back to the
    original ids), with the pull split into per-XCD source slices, and by
    propagation blocking (which takes the place of the XCD split); every
    layout bit-identical between calls, and propagation blocking also for
    grids that do not matter to it."""
    g, G = s15
    ref, _ = O.pagerank(g, 0.85, 20)
    G.set_tuning("pr_relabel", relabel)
    G.set_tuning("pr_xcd", xcd)
    G.set_tuning("pr_pb", pb)
    try:
        p, _, _ = G.pagerank(0.85, 20, 0.0, getattr(B, precision))
        p2, _, _ = G.pagerank(0.85, 20, 0.0, getattr(B, precision))  # cached layout
        G.set_tuning("pr_grid", 300)  # not a multiple of the 8 XCDs
        p3, _, _ = G.pagerank(0.85, 20, 0.0, getattr(B, precision))
    finally:
        G.set_tuning("pr_relabel", 0)
        G.set_tuning("pr_xcd", 0)
        G.set_tuning("pr_pb", 0)
        G.set_tuning("pr_grid", 4096)
    assert np.abs(p - ref).sum() <= (1e-6 if precision == "PR_FP32" else 1e-12)
    assert np.array_equal(p, p2)
    assert np.abs(p3 - ref).sum() <= 1e-6
    if pb:
        assert np.array_equal(p3, p)  # its grids come from the layout


def _star_rmat():
    # RMAT scale 18 ef 4 plus a hub joined to every other vertex: the hub's
    # 262,143 in-edges exceed one gather item (its bin is split), and its
    # out-edges make a source block of one vertex
    g = O.rmat_graph(18, 4)
    n = g.n
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(g.offsets).astype(np.int64))
    keep = src < g.ids  # each undirected edge once
    leaves = np.arange(1, n, dtype=np.uint32)
    return O.sym_csr(n, np.concatenate([src[keep], np.zeros(n - 1, np.uint32)]),
                     np.concatenate([g.ids[keep], leaves]))


def _pagerank_longdouble(g, damping, iters):
    # the Jacobi iteration with every in-sum in 80-bit extended precision
    # (x86 long double, 64-bit mantissa): a reference finer than double
    n = g.n
    deg = np.diff(g.offsets).astype(np.longdouble)
    p = np.full(n, np.longdouble(1) / n)
    starts = g.offsets[:-1].astype(np.int64)
    has = deg > 0
    for _ in range(iters):
        c = np.where(has, p / np.where(has, deg, 1), 0)
        s = np.zeros(n, np.longdouble)
        s[has] = np.add.reduceat(c[g.ids.astype(np.int64)], starts[has])
        p = damping * s + (1 - np.longdouble(damping)) / n
    return p


@pytest.mark.parametrize("shape", ["rmat12_ef1", "rmat14_ef48", "rmat15_ef16", "star"])
def test_pagerank_pb_shapes(B, shape):
    """Propagation blocking where bins and source blocks take odd shapes:
    a sparse graph with many isolated vertices (empty bins), a dense one, the
    bench-like ef 16, and a hub with more in-edges than a gather item takes
    (a split bin: several items adding atomically into one destination);
    FP32 within 1e-6 and FP64 within 1e-12 of the oracle, the FP32 L1 also
    at most the Kahan pull's."""
    if shape == "star":
        g = _star_rmat()
        assert int(np.diff(g.offsets).max()) > (1 << 17)
    else:
        scale, ef = (int(t) for t in shape[4:].split("_ef"))
        g = O.rmat_graph(scale, ef)
    G = B.CsrGraph.from_csr(g.offsets, g.ids)
    ref, l1_ref = O.pagerank(g, 0.85, 25)
    try:
        for relabel in (0, 1):
            G.set_tuning("pr_relabel", relabel)
            G.set_tuning("pr_pb", 0)
            pk, _, _ = G.pagerank(0.85, 25, 0.0, B.PR_FP32)
            G.set_tuning("pr_pb", 1)
            p32, _, _ = G.pagerank(0.85, 25, 0.0, B.PR_FP32)
            p64, l1, _ = G.pagerank(0.85, 25, 0.0, B.PR_FP64)
            e32, ek = float(np.abs(p32 - ref).sum()), float(np.abs(pk - ref).sum())
            print("%s relabel %d: pb FP32 L1 %.2e (Kahan pull %.2e), FP64 %.2e" %
                  (shape, relabel, e32, ek, float(np.abs(p64 - ref).sum())), flush=True)
            assert e32 <= 1e-6 and e32 <= ek
            assert float(np.abs(p64 - ref).sum()) <= 1e-12
            # the residuals to the FP64 vector's own tolerance: the hub's
            # 262,143-term in-sum rounds ~1e-13 differently in the oracle's
            # sequential double sum than in the exact fixed-point sum
            assert np.allclose(l1, l1_ref, rtol=1e-9, atol=1e-12)
            if shape == "star" and relabel == 0:
                # and the fixed-point sums are the closer of the two to an
                # extended-precision reference
                ld = _pagerank_longdouble(g, 0.85, 25)
                e_pb = float(np.abs(p64.astype(np.longdouble) - ld).sum())
                e_or = float(np.abs(ref.astype(np.longdouble) - ld).sum())
                print("star vs long double: pb FP64 %.2e, oracle %.2e" % (e_pb, e_or), flush=True)
                assert e_pb <= e_or
    finally:
        G.close()


@pytest.mark.parametrize("pb", [0, 1])
def test_pagerank_layouts_isolated_and_updates(B, pb):
    """Isolated vertices and ties in degree; the relabelled copy, the XCD
    slices and the propagation-blocking layout follow the graph through
    insert and remove batches."""
    g = O.rmat_graph(12)
    G = B.CpmaGraph.from_csr(g.offsets, g.ids)
    G.set_tuning("pr_relabel", 1)
    G.set_tuning("pr_xcd", 1)
    G.set_tuning("pr_pb", pb)
    s, d = O.update_rmat(g.n, 5000, 3)
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
        ref, _ = O.pagerank(h, 0.85, 15)
        p, _, _ = G.pagerank(0.85, 15, 0.0, B.PR_FP32)
        print(step, h.m, flush=True)
        assert np.abs(p - ref).sum() <= 1e-6, step
    G.close()


def test_ppr_blocks_knob(B, s15):
    g, G = s15
    seeds = O.bfs_sources(g, 4, seed=8)
    base, _ = G.ppr(seeds, 0.85, 15)
    for v in (1, 8):
        G.set_tuning("ppr_blocks_per_cu", v)
        p, _ = G.ppr(seeds, 0.85, 15)
        assert np.array_equal(p, base)  # one lane sums each list in order: identical
    G.set_tuning("ppr_blocks_per_cu", 4)
    for v in (1, 16):
        G.set_tuning("ppr_pieces", v)
        p, _ = G.ppr(seeds, 0.85, 15)
        assert np.array_equal(p, base)
    G.set_tuning("ppr_pieces", 4)
    for j, s in enumerate(seeds):
        assert np.abs(base[j] - O.ppr(g, s, 0.85, 15)).sum() <= 1e-6


def test_bad_knobs(B, s15):
    _, G = s15
    with pytest.raises(B.ByoError):
        G.set_tuning("no_such_knob", 1)
    with pytest.raises(B.ByoError):
        G.set_tuning("pieces", 0)
    with pytest.raises(B.ByoError):
        G.set_tuning("pr_lanes", 6)
