"""CPU oracle for the BYO hot path -- TEST INFRASTRUCTURE ONLY.

Thin numpy/ctypes front end over ``oracle/liboracle.so`` (built from
``oracle/oracle.c`` by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package, and only as the checker / CPU baseline.  The product
(``byo_amd``) never imports it.

Each wrapper names the reference file:line its C function restates; see
``oracle/oracle.h`` for the exact semantics and how each is pinned.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NONE = 0xFFFFFFFF
MLP_NPARAM = 32 * 6 + 32 + 32 * 32 + 32 + 3 * 32 + 3

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        u64, u32, i64, dbl, i32 = C.c_uint64, C.c_uint32, C.c_int64, C.c_double, C.c_int
        sig = {
            "orc_splitmix64": (u64, [u64]),
            "orc_ref_hash64": (u64, [u64]),
            "orc_ref_hash32": (u32, [u32]),
            "orc_rmat_edges": (None, [i32, u64, dbl, dbl, dbl, u64, u64, _u32p, _u32p]),
            "orc_update_rmat": (None, [u64, u64, u64, _u32p, _u32p]),
            "orc_er_edges": (None, [u64, u64, u64, _u32p, _u32p]),
            "orc_grid3d_edges": (u64, [u64, _u32p, _u32p]),
            "orc_sort_unique_u64": (u64, [_u64p, u64]),
            "orc_build_sym_csr": (u64, [u64, u64, _u32p, _u32p, _u64p, _u32p]),
            "orc_keys_to_csr": (None, [u64, _u64p, u64, _u64p, _u32p]),
            "orc_csr_to_keys": (None, [u64, _u64p, _u32p, _u64p]),
            "orc_batch_insert": (u64, [_u64p, u64, _u64p, u64, _u64p]),
            "orc_batch_remove": (u64, [_u64p, u64, _u64p, u64, _u64p]),
            "orc_bfs_levels": (None, [u64, _u64p, _u32p, u32, _u32p]),
            "orc_bfs_diropt": (i32, [u64, u64, _u64p, _u32p, u32, i64, _u32p, _u8p,
                                     _u64p, _u64p, i32]),
            "orc_check_bfs_tree": (u64, [u64, _u64p, _u32p, u32, _u32p, _u32p]),
            "orc_pagerank_f64": (i32, [u64, _u64p, _u32p, dbl, i32, dbl, _f64p, _f64p]),
            "orc_ppr_f64": (i32, [u64, _u64p, _u32p, u32, dbl, i32, _f64p]),
            "orc_pagerank_ref_order": (i32, [u64, _u64p, _u32p, dbl, i32, dbl, _f64p]),
            "orc_cc_union_find": (None, [u64, _u64p, _u32p, _u32p]),
            "orc_cc_labelprop": (i32, [u64, _u64p, _u32p, _u32p]),
            "orc_tc": (u64, [u64, _u64p, _u32p]),
            "orc_tc_brute": (u64, [u64, _u64p, _u32p]),
            "orc_orient": (None, [u64, _u64p, _u32p, _u64p, _u32p]),
            "orc_map_sum": (u64, [u64, _u32p]),
            "orc_bc": (None, [u64, _u64p, _u32p, u32, _f64p]),
            "orc_kcore": (None, [u64, _u64p, _u32p, _u32p]),
            "orc_mis_key": (u64, [u32]),
            "orc_mis": (None, [u64, _u64p, _u32p, _u8p]),
            "orc_coloring": (None, [u64, _u64p, _u32p, _u32p]),
            "orc_llf_before": (i32, [_u64p, u32, u32]),
            "orc_densest": (dbl, [u64, _u64p, _u32p, dbl, C.POINTER(C.c_uint32)]),
            "orc_spanner": (u64, [u64, _u64p, _u32p, u32, _u64p]),
            "orc_pagerank_delta": (i32, [u64, _u64p, _u32p, dbl, dbl, i32, _f64p]),
            "orc_ldd": (None, [u64, _u64p, _u32p, dbl, _u32p, _u32p]),
            "orc_mlp_forward_backward": (dbl, [_f32p, _f32p, _i32p, i32, _f32p, C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class Graph:
    """Host CSR: offsets uint64[n+1], ids uint32[m] (run_csr.cc:28-191)."""

    def __init__(self, offsets, ids):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.ids = np.ascontiguousarray(ids, dtype=np.uint32)
        self.n = len(self.offsets) - 1
        self.m = int(self.offsets[-1]) if self.n >= 0 else 0

    def degrees(self):
        return np.diff(self.offsets).astype(np.uint32)

    def neighbors(self, v):
        return self.ids[self.offsets[v]:self.offsets[v + 1]]


# ---- generators ---------------------------------------------------------

def rmat_edges(scale, num_edges, a=0.5, b=0.1, c=0.1, seed=1, first=0):
    src = np.empty(num_edges, np.uint32)
    dst = np.empty(num_edges, np.uint32)
    lib().orc_rmat_edges(scale, num_edges, a, a + b, a + b + c, seed, first, src, dst)
    return src, dst


def er_graph(scale, num_edges, seed=1):
    n = 1 << scale
    src = np.empty(num_edges, np.uint32)
    dst = np.empty(num_edges, np.uint32)
    lib().orc_er_edges(n, num_edges, seed, src, dst)
    return sym_csr(n, src, dst)


def grid3d_graph(side):
    cap = 3 * side ** 3
    src = np.empty(cap, np.uint32)
    dst = np.empty(cap, np.uint32)
    k = lib().orc_grid3d_edges(side, src, dst)
    return sym_csr(side ** 3, src[:k], dst[:k])


def update_rmat(graph_n, count, first=0):
    """Batch_insert_runner's update stream (run_unweighted.h:797-806)."""
    src = np.empty(count, np.uint32)
    dst = np.empty(count, np.uint32)
    lib().orc_update_rmat(graph_n, first, count, src, dst)
    return src, dst


def sym_csr(n, src, dst):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    off = np.empty(n + 1, np.uint64)
    ids = np.empty(max(2 * len(src), 1), np.uint32)
    m = lib().orc_build_sym_csr(n, len(src), src, dst, off, ids)
    return Graph(off, ids[:m].copy())


def rmat_graph(scale, edge_factor=16, a=0.5, b=0.1, c=0.1, seed=1):
    n = 1 << scale
    src, dst = rmat_edges(scale, edge_factor * n, a, b, c, seed)
    return sym_csr(n, src, dst)


def sort_unique(keys):
    k = np.array(keys, dtype=np.uint64, copy=True)
    w = lib().orc_sort_unique_u64(k, len(k))
    return k[:w]


def keys_to_csr(n, keys):
    keys = np.ascontiguousarray(keys, np.uint64)
    off = np.empty(n + 1, np.uint64)
    ids = np.empty(max(len(keys), 1), np.uint32)
    lib().orc_keys_to_csr(n, keys, len(keys), off, ids)
    return Graph(off, ids[:len(keys)].copy())


def csr_keys(g):
    """(src<<32)|dst keys of a CSR, built by the C oracle (OpenMP)."""
    k = np.empty(max(g.m, 1), np.uint64)
    lib().orc_csr_to_keys(g.n, g.offsets, g.ids, k)
    return k[:g.m]


def batch_insert(E, batch):
    """E u unique(batch) over sorted unique keys E (gbbs/graph.h:3071-3097
    semantics on a sorted array); batch is sorted in place."""
    E = np.ascontiguousarray(E, np.uint64)
    b = np.array(batch, np.uint64, copy=True)
    out = np.empty(len(E) + len(b) + 1, np.uint64)
    k = lib().orc_batch_insert(E, len(E), b, len(b), out)
    return out[:k]


def batch_remove(E, batch):
    """E minus batch over sorted unique keys E (gbbs/graph.h:3099-3124)."""
    E = np.ascontiguousarray(E, np.uint64)
    b = np.array(batch, np.uint64, copy=True)
    out = np.empty(len(E) + 1, np.uint64)
    k = lib().orc_batch_remove(E, len(E), b, len(b), out)
    return out[:k]


def graph_keys(g):
    src = np.repeat(np.arange(g.n, dtype=np.uint64), np.diff(g.offsets).astype(np.int64))
    return (src << np.uint64(32)) | g.ids.astype(np.uint64)


# ---- algorithms ---------------------------------------------------------

def bfs_levels(g, src):
    lv = np.empty(max(g.n, 1), np.uint32)
    lib().orc_bfs_levels(g.n, g.offsets, g.ids, src, lv)
    return lv[:g.n]


def bfs_diropt(g, src, threshold=-1, max_log=256):
    par = np.empty(max(g.n, 1), np.uint32)
    dirlog = np.zeros(max_log, np.uint8)
    fsize = np.zeros(max_log, np.uint64)
    fdeg = np.zeros(max_log, np.uint64)
    nl = lib().orc_bfs_diropt(g.n, g.m, g.offsets, g.ids, src, threshold, par,
                              dirlog, fsize, fdeg, max_log)
    k = min(nl, max_log)
    return par[:g.n], dirlog[:k].copy(), fsize[:k].copy(), fdeg[:k].copy()


def check_bfs_tree(g, src, parents, levels):
    return int(lib().orc_check_bfs_tree(
        g.n, g.offsets, g.ids, src,
        np.ascontiguousarray(parents, np.uint32), np.ascontiguousarray(levels, np.uint32)))


def pagerank(g, damping=0.85, iters=20, eps=0.0):
    p = np.empty(max(g.n, 1), np.float64)
    l1 = np.zeros(max(iters, 1), np.float64)
    it = lib().orc_pagerank_f64(g.n, g.offsets, g.ids, damping, iters, eps, p, l1)
    return p[:g.n], l1[:it].copy()


def ppr(g, seed, damping=0.85, iters=30):
    p = np.empty(max(g.n, 1), np.float64)
    lib().orc_ppr_f64(g.n, g.offsets, g.ids, seed, damping, iters, p)
    return p[:g.n]


def pagerank_ref_order(g, damping=0.85, iters=100, eps=1e-6):
    p = np.empty(max(g.n, 1), np.float64)
    it = lib().orc_pagerank_ref_order(g.n, g.offsets, g.ids, damping, iters, eps, p)
    return p[:g.n], it


def cc(g):
    lab = np.empty(max(g.n, 1), np.uint32)
    lib().orc_cc_union_find(g.n, g.offsets, g.ids, lab)
    return lab[:g.n]


def cc_labelprop(g):
    lab = np.empty(max(g.n, 1), np.uint32)
    lib().orc_cc_labelprop(g.n, g.offsets, g.ids, lab)
    return lab[:g.n]


def tc(g):
    return int(lib().orc_tc(g.n, g.offsets, g.ids))


def tc_brute(g):
    return int(lib().orc_tc_brute(g.n, g.offsets, g.ids))


def orient(g):
    off = np.empty(g.n + 1, np.uint64)
    ids = np.empty(max(g.m // 2, 1), np.uint32)
    lib().orc_orient(g.n, g.offsets, g.ids, off, ids)
    return Graph(off, ids[:int(off[-1])].copy())


def map_sum(g):
    return int(lib().orc_map_sum(g.m, g.ids))


def ldd(g, beta=0.2):
    """LDD without permutation (LowDiameterDecomposition.h:98-186), the
    smallest cluster id winning ties; returns (cluster, round)."""
    cl = np.empty(max(g.n, 1), np.uint32)
    rd = np.empty(max(g.n, 1), np.uint32)
    lib().orc_ldd(g.n, g.offsets, g.ids, beta, cl, rd)
    return cl[:g.n], rd[:g.n]


def ldd_start_round(n, beta=0.2):
    """Round at which vertex v becomes eligible as a centre: the r with
    shifts[r] <= v < shifts[r+1] (generate_shifts, :12-25)."""
    import math
    last = int(min(n + 1, 2 + math.ceil(math.log(n) / beta)))
    sizes = [math.floor(math.exp(i * beta)) for i in range(last)]
    bounds = np.cumsum([0] + sizes)
    return np.searchsorted(bounds, np.arange(n), side="right") - 1


def pagerank_delta(g, eps=1e-4, local_eps=0.01, max_iters=20):
    """delta::PageRankDelta (PageRank.h:335-396); returns (p, rounds)."""
    p = np.empty(max(g.n, 1), np.float64)
    r = lib().orc_pagerank_delta(g.n, g.offsets, g.ids, eps, local_eps, max_iters, p)
    return p[:g.n], r


def mis_key(v):
    """Priority key of vertex v in the MIS order: (splitmix64(v) >> 32, v)."""
    return int(lib().orc_mis_key(int(v)))


def mis(g):
    """Lexicographically first maximal independent set of the mis_key order
    (MaximalIndependentSet_rootset, MaximalIndependentSet.h:119-170);
    uint8[n] membership."""
    out = np.empty(max(g.n, 1), np.uint8)
    lib().orc_mis(g.n, g.offsets, g.ids, out)
    return out[:g.n]


def mis_check(g, in_mis):
    """Number of violations of the lexicographically-first-MIS
    characterisation in the mis_key order (0 for the set orc_mis defines):
    a member with a member neighbour of smaller key, or a non-member with no
    member neighbour of smaller key."""
    n = g.n
    key = np.array([mis_key(v) for v in range(n)], np.uint64)
    src = np.repeat(np.arange(n), np.diff(g.offsets.astype(np.int64)))
    dst = g.ids.astype(np.int64)
    earlier_in = (key[dst] < key[src]) & (in_mis[dst] == 1)
    has = np.zeros(n, bool)
    np.logical_or.at(has, src[earlier_in], True)
    bad_in = (in_mis == 1) & has
    bad_out = (in_mis == 0) & ~has
    return int(bad_in.sum() + bad_out.sum())


def coloring(g):
    """Greedy colouring in the LLF order (Coloring, GraphColoring.h:71-146)
    with coloring::color's rule; uint32[n]."""
    out = np.empty(max(g.n, 1), np.uint32)
    lib().orc_coloring(g.n, g.offsets, g.ids, out)
    return out[:g.n]


def coloring_check(g, color):
    """Violations of the colouring's definition: an edge whose ends share a
    colour, or a vertex whose colour is not the rule (smallest colour in
    [0, deg) unused by the neighbours before it in LLF order, deg + 1 when
    none is free, 0 when isolated) applied to those neighbours."""
    bad = 0
    for v in range(g.n):
        nb = g.ids[g.offsets[v]:g.offsets[v + 1]]
        deg = len(nb)
        if deg == 0:
            bad += color[v] != 0
            continue
        bad += int(np.sum(color[nb] == color[v]))
        used = {int(color[u]) for u in nb if lib().orc_llf_before(g.offsets, int(u), v)}
        c = 0
        while c < deg and c in used:
            c += 1
        bad += color[v] != (c if c < deg else deg + 1)
    return int(bad)


def spanner(g, k=4):
    """Spanner (MPXV15) with beta = ln(n)/(2k), restated deterministically
    (orc_spanner); sorted uint64 edge keys (u << 32 | v)."""
    out = np.empty(max(g.n + g.m, 1), np.uint64)
    cnt = lib().orc_spanner(g.n, g.offsets, g.ids, k, out)
    return out[:cnt].copy()


def spanner_check(g, keys, k=4):
    """Violations of the spanner's definition, computed from the LDD alone:
    every edge in G; one tree edge (v, p) per non-centre v with p the
    smallest neighbour in v's cluster one round nearer; for every adjacent
    cluster pair exactly its smallest (u, v) edge; nothing else."""
    n = g.n
    if n == 0:
        return len(keys)
    cl, rd = ldd(g, math.log(n) / (2 * k))
    want = set()
    edges = set()
    for u in range(n):
        for v in g.ids[g.offsets[u]:g.offsets[u + 1]]:
            edges.add((u << 32) | int(v))
    best = {}
    for v in range(n):
        nb = [int(x) for x in g.ids[g.offsets[v]:g.offsets[v + 1]]]
        if cl[v] != v:
            p = min(u for u in nb if cl[u] == cl[v] and rd[u] + 1 == rd[v])
            want.add((v << 32) | p)
        for u in nb:
            if cl[v] < cl[u]:
                key = (int(cl[v]), int(cl[u]))
                e = (v << 32) | u
                if key not in best or e < best[key]:
                    best[key] = e
    want |= set(best.values())
    got = set(int(x) for x in keys)
    return len(got ^ want) + len(got - edges) + (len(got) != len(keys))


def densest(g, eps=0.001):
    """WorkEfficientDensestSubgraph (DensestSubgraph.h:8-145); returns
    (largest density over the peeling rounds, rounds)."""
    r = C.c_uint32()
    d = lib().orc_densest(g.n, g.offsets, g.ids, eps, C.byref(r))
    return d, r.value


def densest_brute(g):
    """Largest 2|E(S)|/|S| over all nonempty vertex subsets S (n <= 16)."""
    n = g.n
    best = 0.0
    src = np.repeat(np.arange(n), np.diff(g.offsets.astype(np.int64)))
    dst = g.ids.astype(np.int64)
    for mask in range(1, 1 << n):
        ins = np.array([(mask >> v) & 1 for v in range(n)], bool)
        e = int(np.sum(ins[src] & ins[dst]))
        best = max(best, e / int(ins.sum()))
    return best


def kcore(g):
    """Coreness of every vertex (KCore, KCore.h:8-56)."""
    out = np.empty(max(g.n, 1), np.uint32)
    lib().orc_kcore(g.n, g.offsets, g.ids, out)
    return out[:g.n]


def kcore_brute(g):
    """Coreness by the definition: for k = 1, 2, ... peel vertices of degree
    < k until none is left; a vertex's coreness is the last k whose core
    still holds it (small graphs)."""
    n = g.n
    adj = [set(int(x) for x in g.ids[g.offsets[v]:g.offsets[v + 1]]) for v in range(n)]
    core = np.zeros(n, np.uint32)
    k = 1
    while True:
        alive = set(range(n))
        changed = True
        while changed:
            changed = False
            for v in list(alive):
                if len(adj[v] & alive) < k:
                    alive.discard(v)
                    changed = True
        if not alive:
            return core
        for v in alive:
            core[v] = k
        k += 1


def bc(g, src):
    """Single-source betweenness dependencies, bc_bfs::
    SSBetweennessCentrality_BFS (SSBetweennessCentrality.h:332-454), FP64;
    NaN at unreached vertices."""
    out = np.empty(max(g.n, 1), np.float64)
    lib().orc_bc(g.n, g.offsets, g.ids, src, out)
    return out[:g.n]


def bc_brute(g, src):
    """Textbook single-source dependency by brute force (small graphs):
    delta(v) = sum over t != src, v of sigma(src, v) sigma(v, t) / sigma(src, t)
    for v on a shortest src-t path; the source's own value, which the
    level recursion gives as reached - 1, is set to that."""
    n = g.n
    INF = -1

    def bfs_counts(s):
        dist = np.full(n, INF, np.int64)
        sig = np.zeros(n, np.float64)
        dist[s], sig[s] = 0, 1.0
        q = [s]
        for u in q:
            for v in g.ids[g.offsets[u]:g.offsets[u + 1]]:
                v = int(v)
                if dist[v] == INF:
                    dist[v] = dist[u] + 1
                    q.append(v)
                if dist[v] == dist[u] + 1:
                    sig[v] += sig[u]
        return dist, sig

    ds, ss = bfs_counts(src)
    out = np.full(n, np.nan)
    reached = ds != INF
    for v in range(n):
        if not reached[v]:
            continue
        if v == src:
            out[v] = float(reached.sum() - 1)
            continue
        dv, sv = bfs_counts(v)
        acc = 0.0
        for t in range(n):
            if t in (src, v) or not reached[t]:
                continue
            if dv[t] != INF and ds[v] + dv[t] == ds[t]:
                acc += ss[v] * sv[t] / ss[t]
        out[v] = acc
    return out


def remote_values(remote, read_size, n):
    """Element values of a Map_withRemote array: n elements of read_size
    bytes (0: bits, little bit order; wider than 8: the uint64 at each
    element's start, the value of wide_int<size_t, W>)."""
    r = np.ascontiguousarray(remote).view(np.uint8).reshape(-1)
    if read_size == 0:
        return np.unpackbits(r, bitorder="little")[:n].astype(np.uint64)
    if read_size <= 8:
        dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[read_size]
        return r[:n * read_size].view(dt).astype(np.uint64)
    return r[:n * read_size].view(np.uint64).reshape(n, read_size // 8)[:, 0].copy()


def map_remote(g, remote, read_size):
    """Test_Map_Remote (benchmarks/Map/RemoteRead/Map.h:22-34): sum over all
    edges of remote[dst], modulo 2^64 like a size_t."""
    vals = remote_values(remote, read_size, g.n)
    return int(np.sum(vals[g.ids.astype(np.int64)], dtype=np.uint64)) if g.m else 0


def mlp_forward_backward(params, x, labels):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    labels = np.ascontiguousarray(labels, np.int32)
    grads = np.empty(MLP_NPARAM, np.float32)
    logits = np.empty(3 * len(labels), np.float32)
    loss = lib().orc_mlp_forward_backward(params, x, labels, len(labels), grads,
                                          logits.ctypes.data)
    return loss, grads, logits.reshape(-1, 3)


# ---- fixtures -----------------------------------------------------------

def read_adjacency_graph(path):
    """AdjacencyGraph text format (graph_io.cc:35-100): header, n, m,
    n offsets, m ids."""
    with open(path) as f:
        toks = f.read().split()
    assert toks[0] == "AdjacencyGraph", toks[0]
    n, m = int(toks[1]), int(toks[2])
    off = np.empty(n + 1, np.uint64)
    off[:n] = np.array(toks[3:3 + n], dtype=np.uint64)
    off[n] = m
    ids = np.array(toks[3 + n:3 + n + m], dtype=np.uint32)
    return Graph(off, ids)


def bfs_sources(g, k, seed=1):
    """k distinct sources with degree > 0 drawn by SplitMix64(seed)."""
    deg = np.diff(g.offsets)
    out, x = [], seed
    L = lib()
    tries = 0
    while len(out) < k and tries < 100 * k + 1000:
        x = L.orc_splitmix64(x)
        v = x % g.n
        tries += 1
        if deg[v] > 0 and v not in out:
            out.append(int(v))
    return out
