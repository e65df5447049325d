"""byo_amd -- MI355X-native BYO hot path (GBBS edgeMap + BFS / PageRank /
Connectivity / TriangleCounting over a device CSR and a device CPMA).

Python mirror of the reference's container API (``Graph<Representation>``
in gbbs/graph.h:2154-3150 over ``CSR`` / ``PMA_graph``,
benchmarks/run_structures/run_csr.cc:28-191 and run_single_pma.cc:29-140)
over the C ABI of ``libbyo.so`` (declared in include/byo.h).  Every
computation runs in the gfx950 HIP kernels of the library; importing this
package fails loudly when the library has not been built, and every call
raises ``ByoError`` with the library's message when the library reports an
error (no device included).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbyo.so")

OK, E_INVAL, E_OOM, E_HIP, E_STATE, E_NODEV = 0, -1, -2, -3, -4, -5
NONE = 0xFFFFFFFF
DIR_THRESHOLD, DIR_MLP, DIR_SPARSE, DIR_DENSE, DIR_SCHEDULE = 0, 1, 2, 3, 4
DIR_RULE = 255  # schedule entry meaning "use the threshold rule" 
PR_FP64, PR_FP32 = 0, 1
SEL_SINGLE, SEL_BATCH = 0, 1  # selector weights per traversal engine
MLP_NPARAM = 1379


class ByoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("byo error %d: %s" % (code, msg))
        self.code = code


class BfsStats(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("reached", C.c_uint32),
                ("edges_reached", C.c_uint64), ("device_ms", C.c_double),
                ("fsize", C.c_uint64 * 64), ("fdeg", C.c_uint64 * 64),
                ("dir", C.c_uint8 * 64), ("level_ms", C.c_float * 64)]

    def as_dict(self):
        k = min(self.levels, 64)
        return {"levels": self.levels, "reached": self.reached,
                "edges_reached": self.edges_reached, "device_ms": self.device_ms,
                "fsize": list(self.fsize[:k]), "fdeg": list(self.fdeg[:k]),
                "dir": list(self.dir[:k]), "level_ms": list(self.level_ms[:k])}


_vp = C.c_void_p
_u64 = C.c_uint64
_u32 = C.c_uint32
_i32 = C.c_int
_dbl = C.c_double
_P = C.POINTER

# name -> (restype, argtypes); every symbol include/byo.h declares
SIGNATURES = {
    "byo_last_error": (C.c_char_p, []),
    "byo_version": (C.c_char_p, []),
    "byo_device_count": (_i32, []),
    "byo_csr_create": (_i32, [_vp, _vp, _u64, _u64, _P(_vp)]),
    "byo_csr_create_rmat": (_i32, [_i32, _u64, _dbl, _dbl, _dbl, _u64, _P(_vp)]),
    "byo_csr_create_er": (_i32, [_i32, _u64, _u64, _P(_vp)]),
    "byo_csr_create_grid3d": (_i32, [_u32, _P(_vp)]),
    "byo_cpma_create": (_i32, [_vp, _vp, _u64, _u64, _P(_vp)]),
    "byo_cpma_create_from": (_i32, [_vp, _P(_vp)]),
    "byo_graph_destroy": (_i32, [_vp]),
    "byo_graph_kind": (_i32, [_vp, _P(_i32)]),
    "byo_num_vertices": (_i32, [_vp, _P(_u64)]),
    "byo_num_edges": (_i32, [_vp, _P(_u64)]),
    "byo_degrees": (_i32, [_vp, _vp]),
    "byo_degree": (_i32, [_vp, _u32, _P(_u32)]),
    "byo_neighbors": (_i32, [_vp, _u32, _vp, _u64, _P(_u64)]),
    "byo_export_csr": (_i32, [_vp, _vp, _vp]),
    "byo_memory_size": (_i32, [_vp, _P(_u64)]),
    "byo_adjacency_info": (_i32, [C.c_char_p, _i32, _P(_u64), _P(_u64)]),
    "byo_adjacency_read": (_i32, [C.c_char_p, _i32, _u64, _u64, _vp, _vp]),
    "byo_adjacency_write": (_i32, [C.c_char_p, _i32, _vp, _vp, _u64, _u64]),
    "byo_insert_batch": (_i32, [_vp, _vp, _u64]),
    "byo_remove_batch": (_i32, [_vp, _vp, _u64]),
    "byo_insert_batch_dev": (_i32, [_vp, _vp, _u64]),
    "byo_remove_batch_dev": (_i32, [_vp, _vp, _u64]),
    "byo_insert_batch_dev_noalloc": (_i32, [_vp, _vp, _u64]),
    "byo_cpma_stats": (_i32, [_vp, _P(_u64), _P(_u64), _P(_u64)]),
    "byo_update_rmat_keys": (_i32, [_u64, _u64, _u64, _vp]),
    "byo_device_alloc": (_i32, [_u64, _P(_vp)]),
    "byo_device_free": (_i32, [_vp]),
    "byo_copy_to_device": (_i32, [_vp, _vp, _u64]),
    "byo_set_direction_policy": (_i32, [_vp, _i32, C.c_int64]),
    "byo_set_direction_schedule": (_i32, [_vp, _vp, _u32]),
    "byo_bfs": (_i32, [_vp, _u32, _vp, _P(BfsStats)]),
    "byo_bfs_multi": (_i32, [_vp, _vp, _u32, _vp, _vp]),
    "byo_bfs_levels": (_i32, [_vp, _vp]),
    "byo_pagerank": (_i32, [_vp, _dbl, _i32, _dbl, _i32, _vp, _vp, _P(_i32), _P(_dbl)]),
    "byo_ppr_multi": (_i32, [_vp, _vp, _u32, _dbl, _i32, _vp, _P(_dbl)]),
    "byo_cc": (_i32, [_vp, _vp, _P(_dbl)]),
    "byo_cc_labelprop": (_i32, [_vp, _vp, _P(_u32), _P(_dbl)]),
    "byo_tc": (_i32, [_vp, _P(_u64), _P(_dbl)]),
    "byo_bc": (_i32, [_vp, _u32, _vp, _P(_dbl)]),
    "byo_kcore": (_i32, [_vp, _vp, _P(_u32), _P(_u64), _P(_dbl)]),
    "byo_mis": (_i32, [_vp, _vp, _P(_u32), _P(_dbl)]),
    "byo_coloring": (_i32, [_vp, _vp, _P(_u32), _P(_dbl)]),
    "byo_densest_subgraph": (_i32, [_vp, _dbl, _P(_dbl), _P(_u32), _P(_dbl)]),
    "byo_spanner": (_i32, [_vp, _u32, _vp, _u64, _P(_u64), _P(_dbl)]),
    "byo_ldd": (_i32, [_vp, _dbl, _vp, _P(_u32), _P(_dbl)]),
    "byo_pagerank_delta": (_i32, [_vp, _dbl, _dbl, _i32, _vp, _P(_i32), _P(_dbl)]),
    "byo_map_sum": (_i32, [_vp, _P(_u64), _P(_dbl)]),
    "byo_map_remote": (_i32, [_vp, _vp, _i32, _P(_u64), _P(_dbl)]),
    "byo_profile_enable": (_i32, [_vp, _i32]),
    "byo_profile_get": (_i32, [_vp, C.c_char_p, _P(_u64), _P(_dbl)]),
    "byo_profile_reset": (_i32, [_vp]),
    "byo_insert_batches_dev": (_i32, [_vp, _vp, _vp, _u32, _P(_u32), _P(_dbl)]),
    "byo_set_tuning": (_i32, [_vp, C.c_char_p, C.c_int64]),
    "byo_get_tuning": (_i32, [_vp, C.c_char_p, _P(C.c_int64)]),
    "byo_batch_job": (_i32, [_vp, _vp, _u32, _vp, _u32, _dbl, _i32, _vp, _vp, _vp, _P(_dbl)]),
    "byo_selector_load": (_i32, [_vp, _vp, _vp, _vp]),
    "byo_selector_load_engine": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "byo_trainer_create": (_i32, [_vp, C.c_float, _P(_vp)]),
    "byo_trainer_destroy": (_i32, [_vp]),
    "byo_trainer_step": (_i32, [_vp, _vp, _vp, _i32, _P(C.c_float)]),
    "byo_trainer_grad": (_i32, [_vp, _vp, _vp, _i32, _P(C.c_float), _vp, _vp]),
    "byo_trainer_params": (_i32, [_vp, _vp]),
    "byo_trainer_predict": (_i32, [_vp, _vp, _i32, _vp]),
}

_LIB = None


def lib():
    """The loaded libbyo.so (raises ImportError when it was never built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libbyo.so not built: run `python -m byo_amd.build` "
                              "(or __graft_entry__.build()) first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != OK:
        raise ByoError(rc, lib().byo_last_error().decode())
    return rc


def read_adjacency(path, binary=False):
    """(offsets, ids) of an AdjacencyGraph file (gbbs/graph_io.cc:35-100);
    host only."""
    p = os.fsencode(path)
    n, m = _u64(), _u64()
    check(lib().byo_adjacency_info(p, int(binary), C.byref(n), C.byref(m)))
    off = np.empty(n.value + 1, np.uint64)
    ids = np.empty(max(m.value, 1), np.uint32)
    check(lib().byo_adjacency_read(p, int(binary), n.value, m.value, _ptr(off), _ptr(ids)))
    return off, ids[:m.value]


def write_adjacency(path, offsets, ids, binary=False):
    """Writes a host CSR as an AdjacencyGraph file; host only."""
    off = np.ascontiguousarray(offsets, np.uint64)
    ids = np.ascontiguousarray(ids, np.uint32)
    check(lib().byo_adjacency_write(os.fsencode(path), int(binary), _ptr(off), _ptr(ids), len(off) - 1, len(ids)))


def device_count():
    return lib().byo_device_count()


def update_rmat_keys(graph_n, count, first=0):
    """Batch_insert_runner's update stream (run_unweighted.h:797-806) as
    uint64 keys (src<<32)|dst, generated on the device."""
    k = np.empty(max(count, 1), np.uint64)
    check(lib().byo_update_rmat_keys(graph_n, first, count, k.ctypes.data))
    return k[:count]


class DeviceBuffer:
    """A device allocation holding a host array's bytes (for the *_dev
    batch entry points)."""

    def __init__(self, host_array):
        a = np.ascontiguousarray(host_array)
        self.nbytes = a.nbytes
        self.count = a.size
        p = C.c_void_p()
        check(lib().byo_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value
        check(lib().byo_copy_to_device(C.c_void_p(self.ptr), a.ctypes.data, self.nbytes))

    def free(self):
        if self.ptr:
            check(lib().byo_device_free(C.c_void_p(self.ptr)))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _LIB is not None:
                _LIB.byo_device_free(C.c_void_p(self.ptr))
        except Exception:
            pass


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class Graph:
    """A device-resident container with the reference's Graph<> API:
    N(), M(), degree(i), map_neighbors(i, f), insert_batch / remove_batch,
    plus the whole-algorithm entry points."""

    KIND = None

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    # -- lifetime ------------------------------------------------------
    def close(self):
        if self._h:
            check(lib().byo_graph_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if self._h and _LIB is not None:
                _LIB.byo_graph_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- container API (gbbs/graph.h:2179-2308) --------------------------
    def N(self):
        n = _u64()
        check(lib().byo_num_vertices(self._h, C.byref(n)))
        return n.value

    def M(self):
        m = _u64()
        check(lib().byo_num_edges(self._h, C.byref(m)))
        return m.value

    def degree(self, i):
        d = _u32()
        check(lib().byo_degree(self._h, i, C.byref(d)))
        return d.value

    def degrees(self):
        out = np.empty(max(self.N(), 1), np.uint32)
        check(lib().byo_degrees(self._h, _ptr(out)))
        return out[:self.N()]

    def neighbors(self, i):
        d = self.degree(i)
        out = np.empty(max(d, 1), np.uint32)
        cnt = _u64()
        check(lib().byo_neighbors(self._h, i, _ptr(out), d, C.byref(cnt)))
        return out[:cnt.value]

    def map_neighbors(self, i, f):
        """Calls f(i, dst, None) for every neighbour in ascending order
        (run_csr.cc:38-43)."""
        for d in self.neighbors(i):
            f(i, int(d), None)

    def map_neighbors_early_exit(self, i, f):
        for d in self.neighbors(i):
            if f(i, int(d), None):
                break

    def export_csr(self):
        n, m = self.N(), self.M()
        off = np.empty(n + 1, np.uint64)
        ids = np.empty(max(m, 1), np.uint32)
        check(lib().byo_export_csr(self._h, _ptr(off), _ptr(ids)))
        return off, ids[:m]

    def save(self, path, binary=False):
        """Writes the graph as an AdjacencyGraph file (text, or the -b binary
        layout)."""
        off, ids = self.export_csr()
        write_adjacency(path, off, ids, binary)

    def get_memory_size(self):
        b = _u64()
        check(lib().byo_memory_size(self._h, C.byref(b)))
        return b.value

    def insert_batch(self, keys):
        """Graph::insert_batch (gbbs/graph.h:3071-3097); keys are uint64
        (src<<32)|dst, or an (k, 2) array of (src, dst)."""
        k = as_keys(keys)
        check(lib().byo_insert_batch(self._h, _ptr(k), len(k)))

    def remove_batch(self, keys):
        k = as_keys(keys)
        check(lib().byo_remove_batch(self._h, _ptr(k), len(k)))

    def insert_batch_dev(self, dev_ptr, count):
        """insert_batch with the uint64 keys already in device memory."""
        check(lib().byo_insert_batch_dev(self._h, C.c_void_p(dev_ptr), count))

    def remove_batch_dev(self, dev_ptr, count):
        check(lib().byo_remove_batch_dev(self._h, C.c_void_p(dev_ptr), count))

    def insert_batch_dev_noalloc(self, dev_ptr, count):
        """insert_batch_dev for at most 4096 keys without any allocation by
        the library (a batch that would need a rebalance raises E_STATE)."""
        check(lib().byo_insert_batch_dev_noalloc(self._h, C.c_void_p(dev_ptr), count))

    def insert_batches_dev(self, dev_ptr, offsets):
        """Batches keys[offsets[i]:offsets[i+1]] (device memory) inserted in
        order as one captured HIP graph; returns (graph launches, device ms)."""
        offsets = np.ascontiguousarray(offsets, np.uint64)
        nb = len(offsets) - 1
        graphs = _u32()
        ms = _dbl()
        check(lib().byo_insert_batches_dev(self._h, C.c_void_p(dev_ptr), _ptr(offsets), nb, C.byref(graphs),
                                           C.byref(ms)))
        return graphs.value, ms.value

    # -- algorithms -------------------------------------------------------
    def set_direction_policy(self, policy, threshold=-1):
        check(lib().byo_set_direction_policy(self._h, policy, threshold))

    def set_direction_schedule(self, dirs):
        """Per-level directions (0 push, 1 pull, 3 forward, 255 rule)."""
        d = np.ascontiguousarray(dirs, np.uint8)
        check(lib().byo_set_direction_schedule(self._h, _ptr(d), len(d)))

    def bfs(self, src, want_parents=True):
        n = self.N()
        par = np.empty(max(n, 1), np.uint32) if want_parents else None
        st = BfsStats()
        check(lib().byo_bfs(self._h, src, _ptr(par), C.byref(st)))
        return (par[:n] if want_parents else None), st.as_dict()

    def bfs_multi(self, srcs, want_parents=True, want_stats=True):
        """k BFS; returns (parents (k, n) or None, per-source stats dicts, or
        None with want_stats=False: then no statistics are read back)."""
        srcs = np.ascontiguousarray(srcs, np.uint32)
        n, k = self.N(), len(srcs)
        par = np.empty(max(n * k, 1), np.uint32) if want_parents else None
        st = (BfsStats * max(k, 1))() if want_stats else None
        check(lib().byo_bfs_multi(self._h, _ptr(srcs), k, _ptr(par), C.cast(st, C.c_void_p) if want_stats else None))
        return ((par[:n * k].reshape(k, n) if want_parents else None),
                ([st[i].as_dict() for i in range(k)] if want_stats else None))

    def bfs_levels(self):
        n = self.N()
        lv = np.empty(max(n, 1), np.uint32)
        check(lib().byo_bfs_levels(self._h, _ptr(lv)))
        return lv[:n]

    def pagerank(self, damping=0.85, iters=20, eps=0.0, precision=PR_FP32, want_p=True):
        n = self.N()
        p = np.empty(max(n, 1), np.float64) if want_p else None
        l1 = np.empty(max(iters, 1), np.float64)
        it = _i32()
        ms = _dbl()
        check(lib().byo_pagerank(self._h, damping, iters, eps, precision, _ptr(p), _ptr(l1),
                                 C.byref(it), C.byref(ms)))
        return (p[:n] if want_p else None), l1[:it.value].copy(), ms.value

    def ppr(self, seeds, damping=0.85, iters=30, want_p=True):
        seeds = np.ascontiguousarray(seeds, np.uint32)
        n, k = self.N(), len(seeds)
        p = np.empty(max(n * k, 1), np.float32) if want_p else None
        ms = _dbl()
        check(lib().byo_ppr_multi(self._h, _ptr(seeds), k, damping, iters, _ptr(p), C.byref(ms)))
        return (p[:n * k].reshape(k, n) if want_p else None), ms.value

    def set_tuning(self, key, value):
        check(lib().byo_set_tuning(self._h, key.encode(), int(value)))

    def get_tuning(self, key):
        v = C.c_int64()
        check(lib().byo_get_tuning(self._h, key.encode(), C.byref(v)))
        return v.value

    def apply_tuning(self, path=None):
        """Applies the autotuned knobs stored for the closest graph size
        (byo_amd/tuning.json); returns the knobs applied."""
        from . import autotune
        knobs = autotune.lookup(self.N(), self.M(), path)
        for k, v in knobs.items():
            self.set_tuning(k, v)
        return knobs

    def batch_job(self, bfs_srcs, ppr_seeds, damping=0.85, ppr_iters=30, want_parents=True, want_ppr=True):
        """k_bfs BFS sources (one shared traversal) and k_ppr personalized
        PageRank seeds in one persistent launch; returns (parents[k_bfs, n],
        BFS stats, ppr[k_ppr, n], device ms)."""
        srcs = np.ascontiguousarray(bfs_srcs, np.uint32)
        seeds = np.ascontiguousarray(ppr_seeds, np.uint32)
        n, kb, kp = self.N(), len(srcs), len(seeds)
        par = np.empty(max(n * kb, 1), np.uint32) if want_parents else None
        p = np.empty(max(n * kp, 1), np.float32) if want_ppr else None
        st = (BfsStats * max(kb, 1))()
        ms = _dbl()
        check(lib().byo_batch_job(self._h, _ptr(srcs), kb, _ptr(seeds), kp, damping, ppr_iters, _ptr(par),
                                  C.cast(st, C.c_void_p), _ptr(p), C.byref(ms)))
        return ((par[:n * kb].reshape(kb, n) if want_parents else None), [st[i].as_dict() for i in range(kb)],
                (p[:n * kp].reshape(kp, n) if want_ppr else None), ms.value)

    def cc(self):
        n = self.N()
        lab = np.empty(max(n, 1), np.uint32)
        ms = _dbl()
        check(lib().byo_cc(self._h, _ptr(lab), C.byref(ms)))
        return lab[:n], ms.value

    def cc_labelprop(self):
        n = self.N()
        lab = np.empty(max(n, 1), np.uint32)
        r = _u32()
        ms = _dbl()
        check(lib().byo_cc_labelprop(self._h, _ptr(lab), C.byref(r), C.byref(ms)))
        return lab[:n], r.value, ms.value

    def tc(self):
        c = _u64()
        ms = _dbl()
        check(lib().byo_tc(self._h, C.byref(c), C.byref(ms)))
        return c.value, ms.value

    def pagerank_delta(self, eps=1e-4, local_eps=0.01, max_iters=20, want_p=True):
        """PageRankDelta (PageRank.h:335-396); returns (p, rounds, device ms)."""
        n = self.N()
        p = np.empty(max(n, 1), np.float64) if want_p else None
        r, ms = _i32(), _dbl()
        check(lib().byo_pagerank_delta(self._h, eps, local_eps, int(max_iters), _ptr(p), C.byref(r), C.byref(ms)))
        return (p[:n] if want_p else None), r.value, ms.value

    def ldd(self, beta=0.2):
        """Low-diameter decomposition (LDD, permute false); returns
        (cluster centre per vertex, rounds, device ms)."""
        n = self.N()
        cl = np.empty(max(n, 1), np.uint32)
        r, ms = _u32(), _dbl()
        check(lib().byo_ldd(self._h, float(beta), _ptr(cl), C.byref(r), C.byref(ms)))
        return cl[:n], r.value, ms.value

    def mis(self):
        """Maximal independent set (MaximalIndependentSet_rootset): the
        lexicographically first MIS of the (splitmix64(v) >> 32, v) order;
        returns (uint8[n] membership, rounds, device ms)."""
        n = self.N()
        out = np.empty(max(n, 1), np.uint8)
        rounds, ms = _u32(), _dbl()
        check(lib().byo_mis(self._h, _ptr(out), C.byref(rounds), C.byref(ms)))
        return out[:n], rounds.value, ms.value

    def spanner(self, k=4):
        """MPX spanner (Spanner, beta = ln(n)/(2k)) with fixed choices;
        returns (sorted uint64 edge keys u << 32 | v, device ms)."""
        cnt, ms = _u64(), _dbl()
        check(lib().byo_spanner(self._h, int(k), None, 0, C.byref(cnt), C.byref(ms)))
        out = np.empty(max(cnt.value, 1), np.uint64)
        check(lib().byo_spanner(self._h, int(k), _ptr(out), cnt.value, C.byref(cnt), C.byref(ms)))
        return out[:cnt.value], ms.value

    def densest_subgraph(self, eps=0.001):
        """WorkEfficientDensestSubgraph (ApproxPeelingBKV12); returns
        (largest density over the peeling rounds, rounds, device ms)."""
        d, rounds, ms = _dbl(), _u32(), _dbl()
        check(lib().byo_densest_subgraph(self._h, float(eps), C.byref(d), C.byref(rounds), C.byref(ms)))
        return d.value, rounds.value, ms.value

    def coloring(self):
        """Greedy colouring in LLF order (Coloring, lf false); returns
        (uint32[n] colours, rounds, device ms)."""
        n = self.N()
        out = np.empty(max(n, 1), np.uint32)
        rounds, ms = _u32(), _dbl()
        check(lib().byo_coloring(self._h, _ptr(out), C.byref(rounds), C.byref(ms)))
        return out[:n], rounds.value, ms.value

    def kcore(self):
        """Coreness of every vertex (KCore); returns (uint32[n], k_max,
        rounds, device ms)."""
        n = self.N()
        core = np.empty(max(n, 1), np.uint32)
        kmax, rounds, ms = _u32(), _u64(), _dbl()
        check(lib().byo_kcore(self._h, _ptr(core), C.byref(kmax), C.byref(rounds), C.byref(ms)))
        return core[:n], kmax.value, rounds.value, ms.value

    def bc(self, src):
        """Single-source betweenness dependencies (SSBetweennessCentrality_BFS);
        returns (float64[n], device ms), NaN where src does not reach."""
        n = self.N()
        out = np.empty(max(n, 1), np.float64)
        ms = _dbl()
        check(lib().byo_bc(self._h, int(src), _ptr(out), C.byref(ms)))
        return out[:n], ms.value

    def map_remote(self, remote, read_size):
        """Map_withRemote_<read_size>: sum over edges of remote[dst]; remote
        is a byte buffer of n elements of read_size bytes (0: packed bits,
        little bit order).  Returns (sum, device ms)."""
        buf = np.ascontiguousarray(remote).view(np.uint8).reshape(-1)
        if buf.size == 0:
            buf = np.zeros(1, np.uint8)
        n = self.N()
        need = (n + 7) // 8 if read_size == 0 else n * read_size
        if buf.size < need:
            raise ValueError("remote holds %d bytes, %d needed" % (buf.size, need))
        s, ms = _u64(), _dbl()
        check(lib().byo_map_remote(self._h, _ptr(buf), int(read_size), C.byref(s), C.byref(ms)))
        return s.value, ms.value

    def map_sum(self):
        s = _u64()
        ms = _dbl()
        check(lib().byo_map_sum(self._h, C.byref(s), C.byref(ms)))
        return s.value, ms.value

    # -- profiling --------------------------------------------------------
    def profile(self, on=True):
        check(lib().byo_profile_enable(self._h, 1 if on else 0))

    def profile_get(self, phase):
        n = _u64()
        ms = _dbl()
        check(lib().byo_profile_get(self._h, phase.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def profile_reset(self):
        check(lib().byo_profile_reset(self._h))

    def load_selector(self, params, mean, inv_std, engine=0):
        """Selector weights for engine SEL_SINGLE (byo_bfs) or SEL_BATCH
        (the shared multi-source traversal of bfs_multi)."""
        params = np.ascontiguousarray(params, np.float32)
        mean = np.ascontiguousarray(mean, np.float32)
        inv_std = np.ascontiguousarray(inv_std, np.float32)
        assert params.size == MLP_NPARAM and mean.size == 6 and inv_std.size == 6
        check(lib().byo_selector_load_engine(self._h, int(engine), _ptr(params), _ptr(mean), _ptr(inv_std)))


class CsrGraph(Graph):
    """Static CSR (run_csr.cc:28-191)."""

    @classmethod
    def from_csr(cls, offsets, ids):
        off = np.ascontiguousarray(offsets, np.uint64)
        ids = np.ascontiguousarray(ids, np.uint32)
        h = C.c_void_p()
        check(lib().byo_csr_create(_ptr(off), _ptr(ids), len(off) - 1, len(ids), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path, binary=False):
        """CSR from an AdjacencyGraph file (text, or the -b binary layout)."""
        return cls.from_csr(*read_adjacency(path, binary))

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.5, b=0.1, c=0.1, seed=1):
        h = C.c_void_p()
        check(lib().byo_csr_create_rmat(scale, edge_factor, a, b, c, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def erdos_renyi(cls, scale, num_edges, seed=1):
        h = C.c_void_p()
        check(lib().byo_csr_create_er(scale, num_edges, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def grid3d(cls, side):
        h = C.c_void_p()
        check(lib().byo_csr_create_grid3d(side, C.byref(h)))
        return cls(h.value)


class CpmaGraph(Graph):
    """CPMA-style packed-memory-array graph (run_single_pma.cc:29-140)."""

    def stats(self):
        """{"touched_bytes", "rebalances", "leaves"}: PMA slot bytes rewritten
        by all updates so far (S_touched), rebalances so far, leaf count."""
        t, r, l = _u64(), _u64(), _u64()
        check(lib().byo_cpma_stats(self._h, C.byref(t), C.byref(r), C.byref(l)))
        return {"touched_bytes": t.value, "rebalances": r.value, "leaves": l.value}

    @classmethod
    def load(cls, path, binary=False):
        return cls.from_csr(*read_adjacency(path, binary))

    @classmethod
    def from_csr(cls, offsets, ids):
        off = np.ascontiguousarray(offsets, np.uint64)
        ids = np.ascontiguousarray(ids, np.uint32)
        h = C.c_void_p()
        check(lib().byo_cpma_create(_ptr(off), _ptr(ids), len(off) - 1, len(ids), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_graph(cls, g):
        h = C.c_void_p()
        check(lib().byo_cpma_create_from(g._h, C.byref(h)))
        return cls(h.value)


def as_keys(keys):
    a = np.asarray(keys)
    if a.ndim == 2:
        a = (a[:, 0].astype(np.uint64) << np.uint64(32)) | a[:, 1].astype(np.uint64)
    return np.ascontiguousarray(a, np.uint64)


class Trainer:
    """Device trainer of the direction-selector MLP (own fwd/bwd/Adam kernels)."""

    def __init__(self, params, lr=1e-3):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == MLP_NPARAM
        h = C.c_void_p()
        check(lib().byo_trainer_create(_ptr(p), lr, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            check(lib().byo_trainer_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if self._h and _LIB is not None:
                _LIB.byo_trainer_destroy(self._h)
        except Exception:
            pass

    def step(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        loss = C.c_float()
        check(lib().byo_trainer_step(self._h, _ptr(x), _ptr(y), len(y), C.byref(loss)))
        return loss.value

    def grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        loss = C.c_float()
        g = np.empty(MLP_NPARAM, np.float32)
        lg = np.empty(3 * len(y), np.float32)
        check(lib().byo_trainer_grad(self._h, _ptr(x), _ptr(y), len(y), C.byref(loss), _ptr(g), _ptr(lg)))
        return loss.value, g, lg.reshape(-1, 3)

    def params(self):
        p = np.empty(MLP_NPARAM, np.float32)
        check(lib().byo_trainer_params(self._h, _ptr(p)))
        return p

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(len(x), np.int32)
        check(lib().byo_trainer_predict(self._h, _ptr(x), len(x), _ptr(out)))
        return out
