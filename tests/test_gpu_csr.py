"""GPU parity of the static-CSR hot path against the CPU oracle: BFS levels
bit-exact with valid parent trees, direction decisions equal to the GBBS
rule, PageRank within 1e-6 L1 of FP64, component labels and triangle
counts bit-exact, edge sums exact.  Runs through the C ABI (byo_amd)."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")
FIXTURES = ["rMatGraph_J_5_100", "triangles.txt", "star.txt"]


@pytest.fixture(scope="module")
def B():
    import byo_amd
    assert byo_amd.device_count() > 0, "no GPU visible"
    return byo_amd


def dev(B, g):
    return B.CsrGraph.from_csr(g.offsets, g.ids)


def edges_graph(n, edges):
    s = np.array([e[0] for e in edges], np.uint32) if edges else np.zeros(0, np.uint32)
    d = np.array([e[1] for e in edges], np.uint32) if edges else np.zeros(0, np.uint32)
    return O.sym_csr(n, s, d)


def rule_log(g, src, threshold=-1):
    _, dirlog, fs, fd = O.bfs_diropt(g, src, threshold)
    return list(dirlog), list(fs), list(fd)


def check_bfs(B, g, G, src):
    lv = O.bfs_levels(g, src)
    par, st = G.bfs(src)
    assert O.check_bfs_tree(g, src, par, lv) == 0, "invalid parent tree"
    assert np.array_equal(G.bfs_levels(), lv)
    dirlog, fs, fd = rule_log(g, src)
    assert st["fsize"] == fs and st["fdeg"] == fd
    assert st["dir"] == dirlog
    assert st["reached"] == int((lv != O.NONE).sum())
    return st


# ---- the reference's own BFS / CC unit cases (test_bfs.cc, test_cc.cc) ----

def test_reference_unit_cases(B):
    g = edges_graph(3, [])
    G = dev(B, g)
    par, _ = G.bfs(1)
    assert list(par) == [O.NONE, 1, O.NONE]
    assert list(G.cc()[0]) == [0, 1, 2]
    g = edges_graph(7, [(0, 1), (2, 3), (3, 4), (3, 5), (4, 5), (5, 6)])
    G = dev(B, g)
    par, _ = G.bfs(1)
    assert list(par) == [1, 1] + [O.NONE] * 5
    par, _ = G.bfs(2)
    assert par[0] == O.NONE and par[1] == O.NONE and par[2] == 2 and par[3] == 2
    assert par[4] in (3, 5) and par[5] in (3, 4) and par[6] == 5
    assert list(G.cc()[0]) == [0, 0, 2, 2, 2, 2, 2]


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_graphs(B, golden, name):
    g = O.read_adjacency_graph(os.path.join(INPUTS, name))
    G = dev(B, g)
    f = golden[name]
    assert G.N() == f["n"] and G.M() == f["m"]
    st = check_bfs(B, g, G, f["src"])
    assert st["dir"] == f["dirlog"]
    assert list(G.cc()[0]) == f["cc_labels"]
    assert list(G.cc_labelprop()[0]) == f["cc_labels"]
    assert G.tc()[0] == f["triangles"]
    assert G.map_sum()[0] == f["map_sum"]
    off, ids = G.export_csr()
    assert np.array_equal(off, g.offsets) and np.array_equal(ids, g.ids)
    assert np.array_equal(G.degrees(), g.degrees())
    for v in range(min(g.n, 20)):
        assert np.array_equal(G.neighbors(v), g.neighbors(v))


@pytest.mark.parametrize("binary", [False, True])
def test_load_and_save_files(B, golden, tmp_path, binary):
    """Containers loaded from the reference's AdjacencyGraph inputs give the
    golden results; a device graph saved and loaded again is the same."""
    name = "rMatGraph_J_5_100"
    src = os.path.join(INPUTS, name)
    if binary:
        p = str(tmp_path / "g.bin")
        B.write_adjacency(p, *B.read_adjacency(src), binary=True)
    else:
        p = src
    f = golden[name]
    for cls in (B.CsrGraph, B.CpmaGraph):
        G = cls.load(p, binary)
        assert G.N() == f["n"] and G.M() == f["m"]
        assert G.tc()[0] == f["triangles"]
        assert list(G.cc()[0]) == f["cc_labels"]
        G.close()
    R = B.CsrGraph.rmat(12, 8)
    q = str(tmp_path / "r.out")
    R.save(q, binary)
    R2 = B.CsrGraph.load(q, binary)
    a, b = R.export_csr(), R2.export_csr()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    R.close()
    R2.close()


def test_device_rmat_equals_oracle_generator(B, golden):
    for scale, ef in [(8, 4), (12, 16), (16, 16)]:
        g = O.rmat_graph(scale, ef)
        G = B.CsrGraph.rmat(scale, ef)
        off, ids = G.export_csr()
        assert np.array_equal(off, g.offsets) and np.array_equal(ids, g.ids), scale
    assert G.M() == golden["rmat_s16"]["m"]
    assert G.map_sum()[0] == golden["rmat_s16"]["map_sum"]


@pytest.fixture(scope="module")
def s16(B):
    g = O.rmat_graph(16)
    return g, B.CsrGraph.from_csr(g.offsets, g.ids)


def test_bfs_rmat16_many_sources(B, s16, golden):
    g, G = s16
    srcs = golden["rmat_s16"]["sources8"] + [10]
    for s in srcs:
        check_bfs(B, g, G, s)
    st = check_bfs(B, g, G, 10)
    lv = O.bfs_levels(g, 10)
    assert st["fsize"] == golden["rmat_s16"]["bfs10_level_sizes"]
    assert st["dir"] == golden["rmat_s16"]["bfs10_dirlog"]


def test_bfs_multi_matches_single(B, s16):
    g, G = s16
    srcs = O.bfs_sources(g, 4, seed=3)
    par, stats = G.bfs_multi(srcs)
    for k, s in enumerate(srcs):
        lv = O.bfs_levels(g, s)
        assert O.check_bfs_tree(g, s, par[k], lv) == 0
        assert stats[k]["reached"] == int((lv != O.NONE).sum())


@pytest.mark.parametrize("policy", [2, 3])
def test_forced_directions_keep_levels(B, s16, policy):
    """edgeMap with no_dense / dense_only: direction never changes levels."""
    g, G = s16
    G.set_direction_policy(policy)
    try:
        for s in O.bfs_sources(g, 3, seed=9):
            lv = O.bfs_levels(g, s)
            par, st = G.bfs(s)
            assert O.check_bfs_tree(g, s, par, lv) == 0
            assert np.array_equal(G.bfs_levels(), lv)
    finally:
        G.set_direction_policy(0)


def test_bfs_threshold_variants(B, s16):
    g, G = s16
    for thr in (0, 1000, 10 ** 9):
        G.set_direction_policy(0, thr)
        lv = O.bfs_levels(g, 10)
        par, st = G.bfs(10)
        assert O.check_bfs_tree(g, 10, par, lv) == 0
        assert st["dir"] == rule_log(g, 10, thr)[0]
    G.set_direction_policy(0, -1)


def test_bfs_disconnected_and_isolated_sources(B):
    g = O.rmat_graph(12, edge_factor=1)
    G = dev(B, g)
    deg = g.degrees()
    iso = int(np.nonzero(deg == 0)[0][0])
    par, st = G.bfs(iso)
    assert par[iso] == iso and (par != O.NONE).sum() == 1
    assert st["dir"] == [2]
    lab = O.cc(g)
    for s in O.bfs_sources(g, 5, seed=4):
        lv = O.bfs_levels(g, s)
        par, _ = G.bfs(s)
        assert O.check_bfs_tree(g, s, par, lv) == 0
        assert np.array_equal(par != O.NONE, lab == lab[s])


# ---- PageRank ----------------------------------------------------------------

def test_pagerank_fp64_matches_oracle(B, s16):
    g, G = s16
    p_ref, l1_ref = O.pagerank(g, 0.85, 20, 0.0)
    p, l1, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP64)
    assert np.abs(p - p_ref).sum() < 1e-12
    assert np.allclose(l1, l1_ref, rtol=1e-9, atol=1e-15)


def test_pagerank_fp32_kahan_within_1e6(B, s16, fixture_graphs):
    g, G = s16
    p_ref, _ = O.pagerank(g, 0.85, 20, 0.0)
    p, l1, _ = G.pagerank(0.85, 20, 0.0, B.PR_FP32)
    err = np.abs(p - p_ref).sum()
    assert err <= 1e-6, err
    for name, h in fixture_graphs.items():
        H = dev(B, h)
        q_ref, _ = O.pagerank(h, 0.85, 50, 0.0)
        q, _, _ = H.pagerank(0.85, 50, 0.0, B.PR_FP32)
        assert np.abs(q - q_ref).sum() <= 1e-6, name


def test_pagerank_eps_stop_matches_oracle(B, s16):
    g, G = s16
    p_ref, l1_ref = O.pagerank(g, 0.85, 100, 1e-6)
    p, l1, _ = G.pagerank(0.85, 100, 1e-6, B.PR_FP64)
    assert len(l1) == len(l1_ref)
    assert np.abs(p - p_ref).sum() < 1e-12


@pytest.mark.parametrize("layout", ["plain", "relabel_xcd", "relabel_pb", "plain_pb", "cpma", "cpma_pb"])
@pytest.mark.parametrize("eps,iters", [(1e-6, 100), ("at16", 100), (1e-4, 9), (1e-6, 8), (1e-12, 17)])
def test_pagerank_eps_batches_stop_where_the_oracle_does(B, s16, layout, eps, iters):
    """Iterations to a residual bound are queued 8 at a time with a device
    stop word: the iteration count, the residual log and the vector are
    those of stopping at the first residual below eps (or at iters) -- for
    stops inside a batch, at a batch end (at16: an eps between the 15th and
    16th residuals), and at the iteration limit -- in every layout: the
    plain pull, the relabelled copy with per-XCD slices, propagation
    blocking on either, and the CPMA."""
    g, G = s16
    if eps == "at16":
        _, l1_all = O.pagerank(g, 0.85, 30, 0.0)
        eps = float((l1_all[14] + l1_all[15]) / 2)
    p_ref, l1_ref = O.pagerank(g, 0.85, iters, eps)
    if iters == 100 and eps > 1e-6:
        assert len(l1_ref) == 16
    if layout.startswith("cpma"):
        H = B.CpmaGraph.from_csr(g.offsets, g.ids)
        H.set_tuning("pr_pb", 1 if layout.endswith("_pb") else 0)
    else:
        H = G
        H.set_tuning("pr_relabel", 1 if layout.startswith("relabel") else 0)
        H.set_tuning("pr_xcd", 1 if layout == "relabel_xcd" else 0)
        H.set_tuning("pr_pb", 1 if layout.endswith("_pb") else 0)
    try:
        for prec, tol in ((B.PR_FP64, 1e-12), (B.PR_FP32, 1e-6)):
            p, l1, _ = H.pagerank(0.85, iters, eps, prec)
            assert len(l1) == len(l1_ref), (layout, prec, len(l1), len(l1_ref))
            assert np.abs(p - p_ref).sum() <= tol
            if prec == B.PR_FP64:
                assert np.allclose(l1, l1_ref, rtol=1e-9, atol=1e-15)
            p2, l2, _ = H.pagerank(0.85, iters, eps, prec)  # the stop word is cleared per call
            assert len(l2) == len(l1) and np.array_equal(p2, p)
    finally:
        if layout.startswith("cpma"):
            H.close()
        else:
            H.set_tuning("pr_relabel", 0)
            H.set_tuning("pr_xcd", 0)
            H.set_tuning("pr_pb", 0)


def test_ppr_matches_oracle(B, s16):
    g, G = s16
    seeds = O.bfs_sources(g, 5, seed=11)
    p, _ = G.ppr(seeds, 0.85, 30)
    for k, s in enumerate(seeds):
        ref = O.ppr(g, s, 0.85, 30)
        assert np.abs(p[k] - ref).sum() <= 1e-6


# ---- Connectivity / TC / Map ----------------------------------------------------

@pytest.mark.parametrize("scale,ef", [(12, 1), (14, 2), (16, 16)])
def test_cc_min_labels(B, scale, ef):
    g = O.rmat_graph(scale, ef)
    G = dev(B, g)
    ref = O.cc(g)
    assert np.array_equal(G.cc()[0], ref)
    assert np.array_equal(G.cc_labelprop()[0], ref)


@pytest.mark.parametrize("scale,ef", [(10, 16), (14, 8), (16, 16)])
def test_tc_counts(B, scale, ef):
    g = O.rmat_graph(scale, ef)
    G = dev(B, g)
    assert G.tc()[0] == O.tc(g)


def test_tc_dense_hub_lists(B):
    # complete graph K_k: one vertex's oriented list exceeds the LDS stage
    k = 2200
    src = np.repeat(np.arange(k, dtype=np.uint32), k)
    dst = np.tile(np.arange(k, dtype=np.uint32), k)
    g = O.sym_csr(k, src, dst)
    G = dev(B, g)
    assert G.tc()[0] == k * (k - 1) * (k - 2) // 6


def test_map_sum(B, s16, golden):
    g, G = s16
    assert G.map_sum()[0] == O.map_sum(g) == golden["rmat_s16"]["map_sum"]


@pytest.mark.parametrize("rs", [0, 1, 2, 4, 8, 16, 32, 64])
def test_map_remote(B, s16, rs):
    """Map_withRemote_<rs>: random element values (full width, so 64-bit
    sums wrap as size_t does), the reference runner's one-hot input, and
    noise in the padding of wide elements, against the oracle."""
    g, G = s16
    n = g.n
    rng = np.random.default_rng(rs)
    if rs == 0:
        arr = rng.integers(0, 256, (n + 7) // 8, dtype=np.uint8)
    elif rs <= 8:
        arr = rng.integers(0, 2 ** 63, n, dtype=np.uint64).view(np.uint8)[:n * rs].copy()
    else:
        arr = rng.integers(0, 2 ** 63, (n, rs // 8), dtype=np.uint64)
    s, ms = G.map_remote(arr, rs)
    assert s == O.map_remote(g, arr, rs) and ms > 0
    deg = g.degrees()
    v = int(np.argmax(deg))
    one = np.zeros(max(n * rs, (n + 7) // 8), np.uint8)
    if rs == 0:
        one[v >> 3] = 1 << (v & 7)
    else:
        one[v * rs] = 1
    assert G.map_remote(one, rs)[0] == int(deg[v])


def test_errors_are_reported(B, s16):
    g, G = s16
    with pytest.raises(B.ByoError) as e:
        G.bfs(g.n)
    assert e.value.code == B.E_INVAL
    with pytest.raises(B.ByoError) as e:
        G.insert_batch(np.array([1], np.uint64))
    assert e.value.code == B.E_STATE
    bad = g.ids.copy()
    bad[0] = g.n
    with pytest.raises(B.ByoError):
        B.CsrGraph.from_csr(g.offsets, bad)


def test_empty_graph(B):
    G = B.CsrGraph.from_csr(np.zeros(1, np.uint64), np.zeros(0, np.uint32))
    assert G.N() == 0 and G.M() == 0
    assert G.tc()[0] == 0
    assert G.map_sum()[0] == 0


def test_device_er_and_grid_equal_oracle(B):
    e = O.er_graph(12, 4 * 4096, seed=3)
    E = B.CsrGraph.erdos_renyi(12, 4 * 4096, seed=3)
    off, ids = E.export_csr()
    assert np.array_equal(off, e.offsets) and np.array_equal(ids, e.ids)
    g = O.grid3d_graph(9)
    Gd = B.CsrGraph.grid3d(9)
    off, ids = Gd.export_csr()
    assert np.array_equal(off, g.offsets) and np.array_equal(ids, g.ids)
    assert g.m == 2 * 3 * 9 * 9 * 8  # 3 axes x 8 edges per line x 81 lines, both directions
    # BFS on the grid: level of (x,y,z) from the corner is x+y+z
    par, st = Gd.bfs(0)
    lv = Gd.bfs_levels()
    x = np.arange(g.n) % 9
    y = (np.arange(g.n) // 9) % 9
    z = np.arange(g.n) // 81
    assert np.array_equal(lv, (x + y + z).astype(np.uint32))
    assert O.check_bfs_tree(g, 0, par, lv) == 0


def test_direction_schedule_keeps_levels(B, s16):
    g, G = s16
    lv = O.bfs_levels(g, 10)
    for sched in ([0, 0, 1, 1, 3, 0], [3, 3, 3, 3], [1, 0, 1, 0, 1, 0, 1], [255, 255, 3]):
        G.set_direction_schedule(sched)
        par, st = G.bfs(10)
        assert O.check_bfs_tree(g, 10, par, lv) == 0
        for i, d in enumerate(sched):
            if i < st["levels"] and d != 255 and st["dir"][i] != 2:
                assert st["dir"][i] == d
    G.set_direction_policy(0)
