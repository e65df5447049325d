"""Pins the CPU oracle against the reference's own tests and fixtures and
against independent mathematics (brute force, closed forms, scipy).  No GPU.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components, shortest_path

import oracle as O

M32 = 0xFFFFFFFF
M64 = (1 << 64) - 1


def edges_graph(n, edges):
    """MakeUnweightedSymmetricGraph (gbbs/unit_tests/graph_test_utils.h:36-59)."""
    if edges:
        s = np.array([e[0] for e in edges], np.uint32)
        d = np.array([e[1] for e in edges], np.uint32)
    else:
        s = d = np.zeros(0, np.uint32)
    return O.sym_csr(n, s, d)


def scipy_levels(g, src):
    A = sp.csr_matrix((np.ones(g.m), g.ids.astype(np.int64), g.offsets.astype(np.int64)),
                      shape=(g.n, g.n))
    dist = shortest_path(A, unweighted=True, indices=src)
    lv = np.full(g.n, O.NONE, np.uint32)
    fin = np.isfinite(dist)
    lv[fin] = dist[fin].astype(np.uint32)
    return lv


# ---- hashes and the update generator (run_unweighted.h:689-806) -----------

def py_hash32(a):
    a &= M32
    a = ((a + 0x7ed55d16) + (a << 12)) & M32
    a = ((a ^ 0xc761c23c) ^ (a >> 19)) & M32
    a = ((a + 0x165667b1) + (a << 5)) & M32
    a = ((a + 0xd3a2646c) ^ (a << 9)) & M32
    a = ((a + 0xfd7046c5) + (a << 3)) & M32
    a = ((a ^ 0xb55a4f09) ^ (a >> 16)) & M32
    return a


def py_hash64(u):
    v = (u * 3935559000370003845 + 2691343689449507681) & M64
    v ^= v >> 21
    v ^= (v << 37) & M64
    v ^= v >> 4
    v = (v * 4768777513237032717) & M64
    v ^= (v << 20) & M64
    v ^= v >> 41
    v ^= (v << 5) & M64
    return v


def py_update_pair(graph_n, i):
    nn = 1 << ((graph_n - 1).bit_length() - 1)
    h = py_hash32(py_hash64(0) & M32)
    start = py_hash32(((2 * i) & M32) * h & M32)
    stride = py_hash32(((2 * i + 1) & M32) * h & M32)
    row = col = 0
    size = nn
    while size > 1:
        r = py_hash32(start) / 2147483647.0
        if r < 0.5:
            pass
        elif r < 0.5 + 0.1:
            col += size // 2
        elif r < 0.5 + 0.1 + 0.1:
            row += size // 2
        else:
            row += size // 2
            col += size // 2
        start = (start + stride) & M32
        size //= 2
    return row, col


def test_hash_known_answers():
    L = O.lib()
    for x in (0, 1, 2, 12345, 0xFFFFFFFF):
        assert L.orc_ref_hash32(x) == py_hash32(x)
    for x in (0, 1, 2, 1 << 40, M64):
        assert L.orc_ref_hash64(x) == py_hash64(x)
    # literals evaluated once from the Python restatement above
    assert L.orc_ref_hash32(0) == 0x6B4ED927 and L.orc_ref_hash32(1) == 0xB48681B6
    assert L.orc_ref_hash64(0) == 0x7B439D0C1FD00DE3
    assert L.orc_ref_hash64(1) == 0xBEA952A971BA8E83


def test_update_rmat_matches_python_restatement(golden):
    s, d = O.update_rmat(1 << 16, 64)
    for i in range(64):
        assert (int(s[i]), int(d[i])) == py_update_pair(1 << 16, i)
    assert [[int(a), int(b)] for a, b in zip(s, d)] == golden["update_rmat_n65536_first64"]
    s, d = O.update_rmat(1 << 24, 16, first=1000)
    for k in range(16):
        assert (int(s[k]), int(d[k])) == py_update_pair(1 << 24, 1000 + k)
    # H4: endpoints confined to [0, 2^(ceil(log2 n) - 1))
    s, d = O.update_rmat(1 << 16, 100000)
    assert s.max() < (1 << 15) and d.max() < (1 << 15)


def test_rmat_generator_prefix_and_distribution(golden):
    s, d = O.rmat_edges(16, 64, seed=1)
    assert [[int(a), int(b)] for a, b in zip(s, d)] == golden["rmat_s16_seed1_first64"]
    # quadrant frequencies at the top level follow (a, b, c, d)
    s, d = O.rmat_edges(12, 200000, seed=7)
    top_r = (s >> 11) & 1
    top_c = (d >> 11) & 1
    freq = [np.mean((top_r == r) & (top_c == c)) for r, c in ((0, 0), (0, 1), (1, 0), (1, 1))]
    assert np.allclose(freq, [0.5, 0.1, 0.1, 0.3], atol=0.01)


# ---- graph construction (graph_io_test.cc:137-230, graph_test.cc) -----------

def test_sym_graph_normalisation():
    g = edges_graph(2, [(0, 1), (1, 0), (0, 1)])
    assert g.n == 2 and g.m == 2
    assert list(g.neighbors(0)) == [1] and list(g.neighbors(1)) == [0]
    g = edges_graph(3, [(1, 2)])
    assert g.m == 2 and g.degrees()[0] == 0
    g = edges_graph(7, [(1, 5), (0, 5), (6, 3), (2, 0), (1, 0)])
    assert g.m == 10
    assert list(g.neighbors(0)) == [1, 2, 5]
    assert list(g.neighbors(1)) == [0, 5]
    assert list(g.neighbors(2)) == [0]
    assert list(g.neighbors(3)) == [6]
    assert list(g.neighbors(4)) == []
    g = edges_graph(4, [(2, 2), (1, 1)])  # self loops dropped
    assert g.m == 0
    g = edges_graph(0, [])
    assert g.n == 0 and g.m == 0


def test_fixture_facts(fixture_graphs, golden):
    # brute-force facts of the reference's inputs/ graphs
    expect = {"rMatGraph_J_5_100": (128, 708, 107, 4, 125, 7),
              "triangles.txt": (5, 12, 2, 1, 5, 3),
              "star.txt": (10, 18, 0, 1, 10, 2)}
    for name, (n, m, tri, comps, reached, levels) in expect.items():
        g = fixture_graphs[name]
        f = golden[name]
        src = f["src"]
        assert (g.n, g.m) == (n, m) == (f["n"], f["m"])
        assert O.tc_brute(g) == O.tc(g) == tri == f["triangles"]
        lab = O.cc(g)
        assert len(np.unique(lab)) == comps == f["components"]
        lv = O.bfs_levels(g, src)
        assert np.array_equal(lv, scipy_levels(g, src))
        assert (lv != O.NONE).sum() == reached
        assert lv[lv != O.NONE].max() + 1 == levels
        assert list(lv) == f["levels"]


# ---- BFS (BFS.h:30-64, test_bfs.cc:16-51) ---------------------------------

def test_bfs_reference_unit_cases():
    g = edges_graph(3, [])
    par, *_ = O.bfs_diropt(g, 1)
    assert list(par) == [O.NONE, 1, O.NONE]
    g = edges_graph(7, [(0, 1), (2, 3), (3, 4), (3, 5), (4, 5), (5, 6)])
    par, *_ = O.bfs_diropt(g, 1)
    assert list(par) == [1, 1] + [O.NONE] * 5
    par, *_ = O.bfs_diropt(g, 2)
    assert par[0] == O.NONE and par[1] == O.NONE and par[2] == 2 and par[3] == 2
    assert par[4] in (3, 5) and par[5] in (3, 4) and par[6] == 5


def rule(prev_dense, fs, fd, n, m):
    """edge_map_data.h:257-306 closed form."""
    if prev_dense and fs > n // 10:
        return 1
    if fd == 0:
        return 2
    return 1 if fs + fd > m // 20 else 0


@pytest.mark.parametrize("scale,src", [(10, 3), (14, 10), (16, 10)])
def test_bfs_diropt_levels_and_rule(scale, src):
    g = O.rmat_graph(scale)
    lv = O.bfs_levels(g, src)
    assert np.array_equal(lv, scipy_levels(g, src)) if scale <= 14 else True
    par, dirlog, fs, fd = O.bfs_diropt(g, src)
    assert O.check_bfs_tree(g, src, par, lv) == 0
    # frontier k is exactly level k; its degree sum is fixed by levels
    deg = np.diff(g.offsets)
    nl = lv[lv != O.NONE].max() + 1
    for k in range(len(dirlog)):
        at = lv == k
        assert fs[k] == at.sum()
        assert fd[k] == deg[at].sum()
    prev = 0
    for k in range(len(dirlog)):
        assert dirlog[k] == rule(prev, fs[k], fd[k], g.n, g.m)
        prev = 1 if dirlog[k] == 1 else 0
    assert len(dirlog) in (nl, nl + 1)


def test_bfs_tree_checker_catches_planted_errors():
    g = O.rmat_graph(10)
    src = 3
    lv = O.bfs_levels(g, src)
    par, *_ = O.bfs_diropt(g, src)
    assert O.check_bfs_tree(g, src, par, lv) == 0
    reached = np.nonzero((lv != O.NONE) & (lv > 1))[0]
    bad = par.copy()
    v = reached[0]
    bad[v] = src  # src is not at level lv[v]-1 (lv[v] > 1)
    assert O.check_bfs_tree(g, src, bad, lv) == v + 1
    unreached = np.nonzero(lv == O.NONE)[0]
    if len(unreached):
        bad = par.copy()
        bad[unreached[0]] = src
        assert O.check_bfs_tree(g, src, bad, lv) != 0


# ---- PageRank (PageRank.h:80-193) ------------------------------------------

def test_pagerank_star_closed_form(fixture_graphs):
    g = fixture_graphs["star.txt"]  # centre 0, 9 leaves
    n, k, d = g.n, 9, 0.85
    t = (1 - d) / n
    pl = t * (1 + d / k) / (1 - d * d)
    pc = d * k * pl + t
    p, _ = O.pagerank(g, d, 300, 0.0)
    assert abs(p[0] - pc) < 1e-14
    assert np.allclose(p[1:], pl, rtol=0, atol=1e-14)


def test_pagerank_dense_power_iteration(fixture_graphs):
    g = fixture_graphs["rMatGraph_J_5_100"]
    n, d = g.n, 0.85
    deg = np.diff(g.offsets).astype(np.float64)
    A = np.zeros((n, n))
    for i in range(n):
        for s in g.neighbors(i):
            A[i, s] = 1.0 / deg[s]
    p = np.full(n, 1.0 / n)
    for _ in range(20):
        p = d * (A @ p) + (1 - d) / n
    q, l1 = O.pagerank(g, d, 20, 0.0)
    assert np.abs(p - q).max() < 1e-15
    assert len(l1) == 20


def test_pagerank_mass_law_and_eps_stop():
    g = O.rmat_graph(12)
    d = 0.85
    deg = np.diff(g.offsets)
    prev = np.full(g.n, 1.0 / g.n)
    for it in range(1, 6):
        p, _ = O.pagerank(g, d, it, 0.0)
        assert abs(p.sum() - ((1 - d) + d * prev[deg > 0].sum())) < 1e-12
        prev = p
    p, l1 = O.pagerank(g, d, 100, 1e-6)
    assert l1[-1] < 1e-6 and (len(l1) == 1 or l1[-2] >= 1e-6)


def test_pagerank_reference_order_same_fixed_point():
    g = O.rmat_graph(11)
    p1, _ = O.pagerank(g, 0.85, 200, 0.0)
    p2, it = O.pagerank_ref_order(g, 0.85, 200, 1e-13)
    assert np.abs(p1 - p2).sum() < 1e-9


def test_ppr_mass_and_seed():
    g = O.rmat_graph(10)
    seed = O.bfs_sources(g, 1)[0]
    p = O.ppr(g, seed, 0.85, 60)
    comp = O.cc(g) == O.cc(g)[seed]
    assert p[~comp].sum() == 0
    assert abs(p.sum() - 1.0) < 1e-9  # no isolated vertex can hold mass here
    assert p[seed] >= 0.15


def test_ppr_dense_solve(fixture_graphs):
    """The PPR fixed point solves (I - d A D^-1) p = (1 - d) e_s, with
    D^-1 = 0 for isolated vertices (they spread nothing); after 250 Jacobi
    iterations the error is below d^250.  A version with damping and teleport
    swapped, or contributions divided by the target's degree, fails this."""
    g = fixture_graphs["rMatGraph_J_5_100"]
    n, d = g.n, 0.85
    A = np.zeros((n, n))
    for v in range(n):
        for u in g.neighbors(v):
            A[v, u] = 1.0  # p'[v] gathers from its neighbours u
    deg = g.degrees().astype(np.float64)
    dinv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    for seed in (10, 0, 77):
        e = np.zeros(n)
        e[seed] = 1.0
        exact = np.linalg.solve(np.eye(n) - d * A * dinv[None, :], (1 - d) * e)
        p = O.ppr(g, seed, d, 250)
        assert np.abs(p - exact).sum() < 1e-12, seed


def test_ppr_linearity_gives_global_pagerank():
    """PPR is affine in (p0, teleport): the mean over all seeds of PPR after
    t iterations equals global PageRank (p0 = 1/n, teleport (1-d)/n) after t
    iterations, isolated vertices included."""
    g = O.rmat_graph(7)
    for t in (1, 5, 30):
        mean = np.zeros(g.n)
        for s in range(g.n):
            mean += O.ppr(g, s, 0.85, t)
        mean /= g.n
        p, _ = O.pagerank(g, 0.85, t, 0.0)
        assert np.abs(mean - p).sum() < 1e-13, t


# ---- Connectivity (SimpleUnionAsync/Connectivity.h:32-136) -----------------

def test_cc_reference_unit_cases():
    g = edges_graph(3, [])
    assert list(O.cc(g)) == [0, 1, 2]
    g = edges_graph(7, [(0, 1), (2, 3), (3, 4), (3, 5), (4, 5), (5, 6)])
    assert list(O.cc(g)) == [0, 0, 2, 2, 2, 2, 2]


@pytest.mark.parametrize("scale", [8, 12, 16])
def test_cc_min_label_vs_scipy(scale):
    g = O.rmat_graph(scale, edge_factor=2)
    A = sp.csr_matrix((np.ones(g.m), g.ids.astype(np.int64), g.offsets.astype(np.int64)),
                      shape=(g.n, g.n))
    nc, comp = connected_components(A, directed=False)
    minid = np.full(nc, g.n, np.int64)
    np.minimum.at(minid, comp, np.arange(g.n))
    expect = minid[comp].astype(np.uint32)
    assert np.array_equal(O.cc(g), expect)
    assert np.array_equal(O.cc_labelprop(g), expect)


# ---- Triangle counting (Triangle.h:57-204) ---------------------------------

@pytest.mark.parametrize("scale,ef", [(8, 8), (10, 16), (12, 4)])
def test_tc_vs_brute(scale, ef):
    g = O.rmat_graph(scale, edge_factor=ef)
    assert O.tc(g) == O.tc_brute(g)


def test_tc_closed_forms():
    # complete graph K_k has C(k,3) triangles
    k = 9
    g = edges_graph(k, [(i, j) for i in range(k) for j in range(i + 1, k)])
    assert O.tc(g) == k * (k - 1) * (k - 2) // 6
    # trace(A^3)/6 on an RMAT graph
    g = O.rmat_graph(9, edge_factor=8)
    A = np.zeros((g.n, g.n), np.int64)
    for i in range(g.n):
        A[i, g.neighbors(i)] = 1
    assert O.tc(g) == int(np.trace(A @ A @ A)) // 6


def test_orientation_properties():
    g = O.rmat_graph(10)
    dg = O.orient(g)
    assert dg.m * 2 == g.m
    deg = np.diff(g.offsets)
    for u in range(0, g.n, 7):
        for v in dg.neighbors(u):
            assert (deg[u], u) < (deg[v], v)
        nb = dg.neighbors(u)
        assert np.all(nb[1:] > nb[:-1])


# ---- containers / Map -------------------------------------------------------

def test_keys_csr_roundtrip_and_map_sum():
    g = O.rmat_graph(10)
    keys = O.graph_keys(g)
    assert np.all(keys[1:] > keys[:-1])
    g2 = O.keys_to_csr(g.n, keys)
    assert np.array_equal(g2.offsets, g.offsets) and np.array_equal(g2.ids, g.ids)
    assert O.map_sum(g) == int(g.ids.astype(np.uint64).sum())
    rng = np.random.default_rng(5)
    k = rng.integers(0, 1 << 40, 10000, dtype=np.uint64)
    k = np.concatenate([k, k[:3000]])
    assert np.array_equal(O.sort_unique(k), np.unique(k))


# ---- MLP reference math -----------------------------------------------------

def test_mlp_gradients_finite_difference():
    rng = np.random.default_rng(0)
    P = (rng.standard_normal(O.MLP_NPARAM) * 0.3).astype(np.float32)
    x = rng.standard_normal((16, 6)).astype(np.float32)
    y = rng.integers(0, 3, 16).astype(np.int32)
    loss, g, _ = O.mlp_forward_backward(P, x, y)
    h = 1e-2
    idx = rng.choice(O.MLP_NPARAM, 40, replace=False)
    for k in idx:
        Pp = P.astype(np.float64).copy()
        Pm = Pp.copy()
        Pp[k] += h
        Pm[k] -= h
        lp, _, _ = O.mlp_forward_backward(Pp.astype(np.float32), x, y)
        lm, _, _ = O.mlp_forward_backward(Pm.astype(np.float32), x, y)
        fd = (lp - lm) / (float(np.float32(Pp[k])) - float(np.float32(Pm[k])))
        assert abs(fd - g[k]) < 2e-3 + 2e-2 * abs(g[k]), (k, fd, g[k])


def _remote_array(vals, rs, n, pad=0):
    """n values as a Map_withRemote array of read_size rs (pad fills the
    bytes of wide elements after their value)."""
    if rs == 0:
        return np.packbits((vals & 1).astype(np.uint8), bitorder="little")
    if rs <= 8:
        return vals.astype({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[rs])
    a = np.full((n, rs // 8), pad, np.uint64)
    a[:, 0] = vals
    return a


@pytest.mark.parametrize("rs", [0, 1, 2, 4, 8, 16, 32, 64])
def test_map_remote_closed_forms(rs):
    """Map_Remote_runner's input -- one vertex set to 1 -- sums to that
    vertex's in-degree (its degree, the graph being symmetric); all ones sum
    to m; value v at vertex v sums to Test_Map's dst sum."""
    g = O.rmat_graph(10)
    n, deg = g.n, g.degrees()
    for v in (0, int(np.argmax(deg)), n - 1):
        one = np.zeros(n, np.uint64)
        one[v] = 1
        assert O.map_remote(g, _remote_array(one, rs, n, pad=7), rs) == int(deg[v])
    assert O.map_remote(g, _remote_array(np.ones(n, np.uint64), rs, n, pad=3), rs) == g.m
    if rs >= 2:
        ident = np.arange(n, dtype=np.uint64)
        assert O.map_remote(g, _remote_array(ident, rs, n, pad=1), rs) == O.map_sum(g)


def _bc_close(a, b):
    nan = np.isnan(a)
    assert np.array_equal(nan, np.isnan(b))
    assert np.allclose(a[~nan], b[~nan], rtol=1e-9, atol=1e-9)


def test_bc_closed_forms():
    """Star: from a leaf, the centre carries every other leaf (n - 2), the
    leaves nothing; from the centre, the centre = reached - 1.  Path 0-1-..-k
    from 0: vertex i carries the k - i vertices beyond it.  Unreached: NaN."""
    k = 9
    star = O.sym_csr(k + 1, np.zeros(k, np.uint32), np.arange(1, k + 1, dtype=np.uint32))
    b = O.bc(star, 3)
    assert b[0] == k - 1 and b[3] == k and all(b[i] == 0 for i in range(1, k + 1) if i != 3)
    b = O.bc(star, 0)
    assert b[0] == k and all(b[1:] == 0)
    path = O.sym_csr(k + 2, np.arange(k, dtype=np.uint32), np.arange(1, k + 1, dtype=np.uint32))
    b = O.bc(path, 0)
    assert list(b[:k + 1]) == [float(k)] + [float(k - i) for i in range(1, k + 1)]
    assert np.isnan(b[k + 1])  # the isolated vertex


@pytest.mark.parametrize("which", ["rmat", "grid", "er", "triangles"])
def test_bc_matches_brute_force(which):
    if which == "rmat":
        g = O.rmat_graph(7, 4)
    elif which == "grid":
        g = O.grid3d_graph(4)
    elif which == "er":
        g = O.er_graph(6, 90, seed=3)
    else:
        g = O.read_adjacency_graph(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs",
                                                "triangles.txt"))
    for src in (0, g.n // 2):
        _bc_close(O.bc(g, src), O.bc_brute(g, src))


def test_kcore_reference_unit_cases():
    """test_kcore.cc: the edgeless graph and the 8-vertex BasicUsage graph."""
    assert list(O.kcore(O.sym_csr(3, np.zeros(0, np.uint32), np.zeros(0, np.uint32)))) == [0, 0, 0]
    e = np.array([[0, 1], [2, 3], [3, 4], [3, 5], [4, 5], [5, 6]], np.uint32)
    g = O.sym_csr(8, e[:, 0], e[:, 1])
    assert list(O.kcore(g)) == [1, 1, 1, 2, 2, 2, 1, 0]


@pytest.mark.parametrize("which", ["rmat", "grid", "er", "inputs"])
def test_kcore_matches_definition(which):
    if which == "rmat":
        g = O.rmat_graph(8, 8)
    elif which == "grid":
        g = O.grid3d_graph(5)
    elif which == "er":
        g = O.er_graph(7, 400, seed=5)
    else:
        g = O.read_adjacency_graph(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs",
                                                "rMatGraph_J_5_100"))
    assert np.array_equal(O.kcore(g), O.kcore_brute(g))


def test_mis_reference_unit_cases():
    """test_mis.cc: every vertex of the edgeless graph is in the set; the
    8-vertex BasicUsage graph gets a valid MIS (CheckMIS: independent, and
    every vertex outside has a neighbour inside)."""
    assert list(O.mis(O.sym_csr(3, np.zeros(0, np.uint32), np.zeros(0, np.uint32)))) == [1, 1, 1]
    e = np.array([[0, 1], [2, 3], [3, 4], [3, 5], [4, 5], [5, 6]], np.uint32)
    g = O.sym_csr(8, e[:, 0], e[:, 1])
    m = O.mis(g)
    adj = [set(int(x) for x in g.ids[g.offsets[v]:g.offsets[v + 1]]) for v in range(8)]
    assert all(not (m[u] and m[v]) for u in range(8) for v in adj[u])
    assert all(m[v] or any(m[u] for u in adj[v]) for v in range(8))
    assert m[7] == 1  # isolated


@pytest.mark.parametrize("which", ["rmat", "grid", "er", "inputs"])
def test_mis_is_lexicographically_first(which):
    """The set satisfies the characterisation of the lexicographically first
    MIS of the key order, which determines it uniquely; a vertex moved in or
    out is caught."""
    if which == "rmat":
        g = O.rmat_graph(10, 8)
    elif which == "grid":
        g = O.grid3d_graph(8)
    elif which == "er":
        g = O.er_graph(9, 1200, seed=6)
    else:
        g = O.read_adjacency_graph(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs",
                                                "rMatGraph_J_5_100"))
    m = O.mis(g)
    assert O.mis_check(g, m) == 0
    deg = np.diff(g.offsets.astype(np.int64))
    v = int(np.nonzero((m == 0) & (deg > 0))[0][0])
    bad = m.copy()
    bad[v] = 1
    assert O.mis_check(g, bad) > 0
    bad = m.copy()
    bad[int(np.nonzero(m)[0][0])] = 0
    assert O.mis_check(g, bad) > 0


def test_coloring_hand_cases():
    """coloring::color (GraphColoring.h:30-55) on graphs worked by hand: on a
    triangle (all of one degree level) the vertices in key order take 0, 1
    and then deg + 1 = 3, as all of [0, 2) is taken; on a star the centre
    (higher log-degree, first) takes 0 and every leaf, whose one colour slot
    [0, 1) is taken, takes 2; isolated vertices take 0."""
    t = O.sym_csr(4, np.array([0, 1, 0], np.uint32), np.array([1, 2, 2], np.uint32))
    order = sorted(range(3), key=O.mis_key)
    c = O.coloring(t)
    assert [int(c[v]) for v in order] == [0, 1, 3] and c[3] == 0
    star = O.sym_csr(6, np.zeros(5, np.uint32), np.arange(1, 6, dtype=np.uint32))
    assert list(O.coloring(star)) == [0, 2, 2, 2, 2, 2]


@pytest.mark.parametrize("which", ["rmat", "grid", "er", "inputs"])
def test_coloring_matches_definition(which):
    if which == "rmat":
        g = O.rmat_graph(9, 8)
    elif which == "grid":
        g = O.grid3d_graph(6)
    elif which == "er":
        g = O.er_graph(8, 700, seed=2)
    else:
        g = O.read_adjacency_graph(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs",
                                                "rMatGraph_J_5_100"))
    c = O.coloring(g)
    assert O.coloring_check(g, c) == 0
    bad = c.copy()
    v = int(np.argmax(np.diff(g.offsets.astype(np.int64))))
    bad[v] += 1
    assert O.coloring_check(g, bad) > 0


@pytest.mark.parametrize("which", ["rmat", "grid", "er", "inputs"])
def test_spanner_definition_and_connectivity(which):
    """The spanner equals its definition over the LDD (tree edges to the
    smallest neighbour one round nearer, the smallest edge of every adjacent
    cluster pair), is a subgraph of G, and keeps G's components."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    if which == "rmat":
        g = O.rmat_graph(10, 8)
    elif which == "grid":
        g = O.grid3d_graph(7)
    elif which == "er":
        g = O.er_graph(9, 900, seed=7)
    else:
        g = O.read_adjacency_graph(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs",
                                                "rMatGraph_J_5_100"))
    keys = O.spanner(g)
    assert O.spanner_check(g, keys) == 0
    assert O.spanner_check(g, keys[1:]) > 0
    n = g.n
    u = (keys >> np.uint64(32)).astype(np.int64)
    v = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    S = sp.csr_matrix((np.ones(len(keys)), (u, v)), shape=(n, n))
    A = sp.csr_matrix((np.ones(g.m), g.ids.astype(np.int64), g.offsets.astype(np.int64)), shape=(n, n))
    ns, ls = cg.connected_components(S, directed=False)
    na, la = cg.connected_components(A, directed=False)
    assert ns == na
    assert len(set(zip(ls.tolist(), la.tolist()))) == na
    assert len(keys) < g.m // 2 + n


def test_densest_closed_forms():
    """WorkEfficientDensestSubgraph: a k-clique has density k - 1 (every
    vertex removed in round 1); a star with s leaves has 2s/(s+1) in round 1
    (leaves go, then the centre alone at 0)."""
    k = 9
    src = np.repeat(np.arange(k, dtype=np.uint32), k)
    dst = np.tile(np.arange(k, dtype=np.uint32), k)
    d, r = O.densest(O.sym_csr(k, src, dst))
    assert d == k - 1 and r == 1
    s = 7
    d, r = O.densest(O.sym_csr(s + 1, np.zeros(s, np.uint32), np.arange(1, s + 1, dtype=np.uint32)))
    assert d == 2 * s / (s + 1) and r == 2


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_densest_vs_brute_force(seed):
    """On small random graphs the result is the density of some subset (at
    most the optimum) and within the 2(1 + eps) guarantee of it."""
    rng = np.random.default_rng(seed)
    n = 12
    e = rng.integers(0, n, (30, 2)).astype(np.uint32)
    g = O.sym_csr(n, e[:, 0], e[:, 1])
    best = O.densest_brute(g)
    for eps in (0.001, 0.5):
        d, _ = O.densest(g, eps)
        assert d <= best + 1e-12
        assert d >= best / (2 * (1 + eps)) - 1e-12


def test_mis_key_is_splitmix():
    # known answer: splitmix64(0) = 0xE220A8397B1DCDAF, so key(0) = 0xE220A839 << 32
    assert O.mis_key(0) == 0xE220A839 << 32
    assert O.mis_key(5) & 0xFFFFFFFF == 5


def test_pagerank_delta_pins():
    """PageRankDelta (PageRank.h:335-396).  One round is the first Jacobi
    iterate; with every vertex active (local eps -inf) r rounds are r Jacobi
    iterations (each round adds d * A * (p_k - p_{k-1})); and two rounds on
    the path 0-1-2 worked by hand, where the first round's signed test
    (delta > local_eps * p, :289) keeps only vertex 1 active."""
    g = O.rmat_graph(11)
    p, r = O.pagerank_delta(g, 0.0, 0.01, 1)
    assert r == 1 and np.abs(p - O.pagerank(g, 0.85, 1)[0]).sum() < 1e-12
    for iters in (2, 5, 20):
        p, r = O.pagerank_delta(g, 0.0, -1e300, iters)
        assert r == iters and np.abs(p - O.pagerank(g, 0.85, iters)[0]).sum() < 1e-10
    path = O.sym_csr(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    p, r = O.pagerank_delta(path, 0.0, 0.01, 2)
    d, third = 0.85, 1.0 / 3.0
    pre0 = d * (third / 2) + (1 - d) * third
    pre1 = d * (2 * third) + (1 - d) * third
    dod1 = (pre1 - third) / 2
    assert r == 2
    assert np.allclose(p, [pre0 + d * dod1, pre1, pre0 + d * dod1], rtol=0, atol=1e-15)
    # stops once sum |delta| < eps
    p, r = O.pagerank_delta(g, 1e-4, 0.01, 100)
    assert 1 < r < 100


@pytest.mark.parametrize("which", ["rmat", "grid", "er", "inputs"])
def test_ldd_properties(which):
    """LDD (MPX13, permute false, beta 0.2): centres label themselves at the
    round they become eligible and were not reached earlier; every other
    vertex is reached one round after its nearest frontier neighbour, no
    later than its own eligibility round, and joins the smallest cluster
    among the neighbours one round ahead of it; clusters are connected."""
    if which == "rmat":
        g = O.rmat_graph(11)
    elif which == "grid":
        g = O.grid3d_graph(9)
    elif which == "er":
        g = O.er_graph(10, 1500, seed=4)
    else:
        g = O.read_adjacency_graph(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs",
                                                "rMatGraph_J_5_100"))
    cl, rd = O.ldd(g)
    n = g.n
    s = O.ldd_start_round(n)
    assert (cl < n).all() and (rd != O.NONE).all()
    for v in range(n):
        nb = g.ids[g.offsets[v]:g.offsets[v + 1]].astype(np.int64)
        near = rd[nb].min() + 1 if len(nb) else 1 << 40
        if cl[v] == v:
            assert rd[v] == s[v] and near > rd[v]
        else:
            assert rd[v] == near and s[v] >= rd[v]
            ahead = nb[rd[nb] == rd[v] - 1]
            assert cl[v] == cl[ahead].min()
    # clusters are connected: from each centre, a walk inside the cluster reaches all of it
    for c in np.unique(cl):
        members = set(np.nonzero(cl == c)[0].tolist())
        seen, st = {int(c)}, [int(c)]
        while st:
            u = st.pop()
            for w in g.ids[g.offsets[u]:g.offsets[u + 1]]:
                w = int(w)
                if w in members and w not in seen:
                    seen.add(w)
                    st.append(w)
        assert seen == members


# ---- sorted-array batch updates (the CPMA's CPU baseline) -----------------

@pytest.mark.parametrize("m,b", [(0, 50), (1, 3), (5000, 0), (20000, 7000), (300000, 120000)])
def test_batch_insert_remove_match_set_algebra(m, b):
    rng = np.random.default_rng(m + b)
    E = np.unique(rng.integers(0, 1 << 40, m, dtype=np.uint64))
    batch = rng.integers(0, 1 << 40, b, dtype=np.uint64)
    if len(E) and b:
        batch[: b // 3] = E[rng.integers(0, len(E), b // 3)]  # present keys
        batch[b // 3: b // 2] = batch[: b // 2 - b // 3]         # repeats
    assert np.array_equal(O.batch_insert(E, batch), np.union1d(E, batch))
    assert np.array_equal(O.batch_remove(E, batch), np.setdiff1d(E, batch))


def test_csr_keys_equal_python_keys():
    g = O.rmat_graph(12)
    assert np.array_equal(O.csr_keys(g), O.graph_keys(g))
