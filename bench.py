"""Benchmark of the BYO hot path on one MI355X.

Workload (one "step"): direction-optimizing BFS from 8 seeded sources over a
device-resident RMAT scale-22 static CSR (a=.5 b=.1 c=.1, edge factor 16,
symmetrised, deduplicated, uint32 ids), all 8 in one shared traversal (one
persistent launch) whose per-level direction is chosen by the learned
selector.  `value` is BFS GTEPS over the timed steps (Graph500 convention:
undirected edges reached / time).  The same JSON line carries the same
workload under the M/20 threshold rule (learned vs threshold), the
one-source-at-a-time engine, FP32-Kahan and FP64 PageRank it/s (20
iterations) with their L1 difference, CC and TC times, CPMA insert/remove
edges/s at batch 1e7 with their roofline (24 b + 2 S_touched bytes), BFS
and PageRank rerun on the updated CPMA (read from the PMA itself), the
roofline of the BFS launch, the reference's CPU-runnable case (RMAT-16, BFS
from 10, PageRank to 1e-6, CC, TC) on the GPU and the host, and the CPU
baselines: the oracle's OpenMP direction-optimizing BFS on the host on a
bounded sample of sources (cpu_baseline), and the oracle's PageRank, CC, TC
and sorted-array batch updates (cpu_baselines_other).

    python bench.py [--steps K] [--warmup W] [--scale 22] [--cpma-scale 24]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def splitmix64(x):
    M = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def pick_sources(deg, k, seed=1):
    """k distinct sources with degree > 0 from the SplitMix64(seed) stream."""
    out, x, n = [], seed, len(deg)
    while len(out) < k:
        x = splitmix64(x)
        v = x % n
        if deg[v] > 0 and v not in out:
            out.append(int(v))
    return out


def bfs_bytes(n, reached, edges_reached):
    """Algorithmic bytes of one top-down-equivalent BFS: offsets of reached
    vertices + their adjacency once + parent init + parent writes."""
    return 8 * reached + 4 * edges_reached + 4 * n + 4 * reached


def levels_from_parents(par, src):
    """BFS levels of one parent tree (parents[src] = src, 0xFFFFFFFF for
    unreached vertices; -1 there) by pointer jumping: depth[v] accumulates
    the hops to the root while anc[v] doubles towards it."""
    n = len(par)
    idx = np.arange(n, dtype=np.int64)
    reached = par != 0xFFFFFFFF
    anc = np.where(reached, par.astype(np.int64), idx)
    depth = (anc != idx).astype(np.int64)
    while True:
        nxt = anc[anc]
        if np.array_equal(nxt, anc):
            break
        depth += depth[anc]
        anc = nxt
    depth[~reached] = -1
    assert anc[src] == src
    return depth


def sync():
    """Every libbyo call returns after its device work has completed (the
    library synchronises its stream), so the timed region needs no extra
    device synchronisation; kept as the single place to add one."""
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def protocol(f, rounds):
    """run_unweighted.h:135-170: rounds + 1 runs, round 0 untimed, mean of
    the rest (seconds); returns (mean, last result)."""
    ts, res = [], None
    for r in range(rounds + 1):
        t0 = time.perf_counter()
        res = f()
        dt = time.perf_counter() - t0
        if r:
            ts.append(dt)
    return float(np.mean(ts)), res


def cpu_baseline(G, sources, budget_s):
    """The oracle's OpenMP direction-optimizing BFS (restating BFS.h over
    edgeMapData) on this host's cores, on as many of the sources as fit in
    budget_s; returns (GTEPS over that sample, the oracle graph)."""
    import oracle as O
    off, ids = G.export_csr()
    g = O.Graph(off, ids)
    cores = os.cpu_count()
    done, t_total, edges = 0, 0.0, 0
    O.bfs_diropt(g, sources[0])  # warm (page in)
    # the sources are traversed one after another, in passes over the list,
    # until budget_s of CPU time has been spent
    while t_total < budget_s and done < 1000 * len(sources):
        s = sources[done % len(sources)]
        t0 = time.perf_counter()
        par, _, _, fd = O.bfs_diropt(g, s)
        dt = time.perf_counter() - t0
        t_total += dt
        edges += int(fd.sum()) // 2
        done += 1
    return {"value": edges / t_total / 1e9, "unit": "GTEPS", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "oracle OpenMP diropt BFS (M/20 rule) on the same RMAT-%d graph: %d traversals "
                      "(passes over the %d sources), %.1f s" % (int(np.log2(g.n)), done, len(sources), t_total)}, g


def cpu_other(g, H, keys, rounds=1):
    """The oracle on the host cores for the rest of the path, by the runner
    protocol with `rounds` timed runs after an untimed one: PageRank FP64
    (20 iterations), CC and TC on the benchmark graph g (RMAT-22), and the
    CPMA update as a sorted-array insert / remove (sort + unique + merge,
    gbbs/graph.h:3071-3124) of the 1e7-key batch into the RMAT-24 edge set."""
    import oracle as O
    out = {"cores": os.cpu_count(), "cpu_model": cpu_model(), "kind": "port", "rounds": rounds}
    t, _ = protocol(lambda: O.pagerank(g, 0.85, 20, 0.0), rounds)
    out["pr_fp64_it_s"] = round(20 / t, 2)
    t, _ = protocol(lambda: O.cc(g), rounds)
    out["cc_ms"] = round(t * 1e3, 1)
    t, _ = protocol(lambda: O.tc(g), rounds)
    out["tc_ms"] = round(t * 1e3, 1)
    if H is not None:
        off, ids = H.export_csr()
        E = O.csr_keys(O.Graph(off, ids))
        del off, ids
        t, E2 = protocol(lambda: O.batch_insert(E, keys), rounds)
        out["cpma_insert_edges_s"] = round(len(keys) / t, 1)
        t, _ = protocol(lambda: O.batch_remove(E2, keys), rounds)
        out["cpma_remove_edges_s"] = round(len(keys) / t, 1)
        out["cpma_sample"] = "sorted-array insert/remove of the %d-key batch into the %d-edge RMAT set" % (
            len(keys), len(E))
    return out


def c1_plain_method(B, rounds=3):
    """Config C1, the reference's CPU-runnable case, on the GPU and on the
    host: RMAT scale 16 ef 16 (seed 1), BFS from vertex 10 (the next vertex
    with edges if 10 has none), PageRank to L1 < 1e-6 (at most 100
    iterations; FP64 and FP32-Kahan), CC and TC, each by the runner protocol
    (rounds + 1 runs, round 0 dropped, mean; run_unweighted.h:135-170)."""
    import oracle as O
    G = B.CsrGraph.rmat(16, 16)
    off, ids = G.export_csr()
    g = O.Graph(off, ids)
    deg = np.diff(off)
    src = 10
    while deg[src] == 0:
        src += 1
    gpu, cpu = {}, {}
    t, (_, st) = protocol(lambda: G.bfs(src, want_parents=False), rounds)
    gpu["bfs_ms"] = round(t * 1e3, 4)
    gpu["bfs_gteps"] = round(st["edges_reached"] / 2 / t / 1e9, 3)
    p_ref, l1_ref = O.pagerank(g, 0.85, 100, 1e-6)
    for prec, key in ((B.PR_FP64, "fp64"), (B.PR_FP32, "fp32")):
        t, (p, l1, _) = protocol(lambda: G.pagerank(0.85, 100, 1e-6, prec), rounds)
        gpu["pr_%s_ms" % key] = round(t * 1e3, 4)
        gpu["pr_%s_iters" % key] = len(l1)
        gpu["pr_%s_l1_vs_cpu_fp64" % key] = float(np.abs(p - p_ref).sum())
    t, (lab, _) = protocol(lambda: G.cc(), rounds)
    gpu["cc_ms"] = round(t * 1e3, 4)
    t, (tri, _) = protocol(lambda: G.tc(), rounds)
    gpu["tc_ms"] = round(t * 1e3, 4)
    t, (_, _, _, fd) = protocol(lambda: O.bfs_diropt(g, src), rounds)
    cpu["bfs_ms"] = round(t * 1e3, 3)
    cpu["bfs_gteps"] = round(int(fd.sum()) / 2 / t / 1e9, 3)
    t, (_, l1c) = protocol(lambda: O.pagerank(g, 0.85, 100, 1e-6), rounds)
    cpu["pr_fp64_ms"] = round(t * 1e3, 3)
    cpu["pr_fp64_iters"] = len(l1c)
    t, lab_c = protocol(lambda: O.cc(g), rounds)
    cpu["cc_ms"] = round(t * 1e3, 3)
    t, tri_c = protocol(lambda: O.tc(g), rounds)
    cpu["tc_ms"] = round(t * 1e3, 3)
    res = {"graph": "RMAT-16 ef16 (n=%d, m=%d)" % (g.n, g.m), "src": src,
           "protocol": "rounds %d + 1 untimed, mean" % rounds, "gpu": gpu,
           "cpu": dict(cpu, cores=os.cpu_count(), kind="port"),
           "triangles": int(tri), "components": int(len(np.unique(lab))),
           "agree": bool(tri == tri_c and np.array_equal(lab, lab_c))}
    G.close()
    return res


def batched_job(B, args):
    """The batched persistent workload: on RMAT scale job_scale ef 16, 64 BFS
    sources (one shared traversal) and 32 personalized-PageRank seeds (FP32,
    damping 0.85, 30 iterations) in one persistent launch; then 16 insert
    batches of log-uniform sizes 1e3..1e6 (seeded) applied in order as
    captured HIP graphs on a CPMA built from the same graph."""
    J = B.CsrGraph.rmat(args.job_scale, 16)
    knobs = J.apply_tuning()
    from byo_amd import selector as S
    learned = bool(S.load(J).get("batch"))
    n = J.N()
    deg = J.degrees()
    srcs = pick_sources(deg, 64, seed=2)
    seeds = pick_sources(deg, 32, seed=3)
    parts = {}
    # the job's traversal under the rule and under the learned selector
    for name, pol in (("rule", B.DIR_THRESHOLD), ("learned", B.DIR_MLP if learned else B.DIR_THRESHOLD)):
        J.set_direction_policy(pol)
        J.batch_job(srcs, seeds, 0.85, 30, want_parents=False, want_ppr=False)
        J.profile(True)
        J.profile_reset()
        _, st, _, ms = J.batch_job(srcs, seeds, 0.85, 30, want_parents=False, want_ppr=False)
        parts[name] = sum(J.profile_get(k)[1] for k in ("bfs_batch_seed", "bfs_batch_levels", "bfs_batch_tally"))
    bfs_ms = parts["learned"]
    edges = sum(s["edges_reached"] // 2 for s in st)
    res = {"graph": "RMAT-%d ef16 (n=%d, m=%d)" % (args.job_scale, n, J.M()), "tuning": knobs,
           "bfs_sources": 64, "ppr_seeds": 32,
           "ppr_iters": 30, "launch_ms": round(ms, 3), "bfs_part_ms": round(bfs_ms, 3),
           "ppr_part_ms": round(ms - bfs_ms, 3), "bfs_gteps_in_launch": round(edges / (bfs_ms * 1e-3) / 1e9, 2),
           "ppr_ms_per_iteration": round((ms - bfs_ms) / 30, 3),
           "direction": "learned selector" if learned else "M/20 rule",
           "bfs_part_ms_rule": round(parts["rule"], 3),
           "bfs_learned_vs_rule": round(parts["rule"] / parts["learned"], 4)}
    rng = np.random.default_rng(16)
    sizes = [int(10 ** x) for x in rng.uniform(3, 6, 16)]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    keys = B.update_rmat_keys(n, int(offs[-1]))
    C = B.CpmaGraph.from_graph(J)
    J.close()
    dkeys = B.DeviceBuffer(keys)
    m0 = C.M()
    graphs, ins_ms = C.insert_batches_dev(dkeys.ptr, offs)
    res.update({"insert_batches": 16, "insert_sizes": sizes, "insert_keys": int(offs[-1]),
                "edges_added": int(C.M() - m0), "insert_graph_launches": graphs, "insert_ms": round(ins_ms, 3),
                "insert_edges_s": round(float(offs[-1]) / (ins_ms * 1e-3), 1)})
    dkeys.free()
    C.close()
    return res


def only_section(B, G, args, srcs, policy):
    """One section of the benchmark, args.steps times after args.warmup
    untimed runs, for profiling a kernel on its own; returns a short record."""
    n, m = G.N(), G.M()
    res = {"section": args.only, "graph": "RMAT-%d ef%d" % (args.scale, args.ef), "n": n, "m": m}
    if args.only == "bfs":
        G.set_direction_policy(policy)
        f = lambda: G.bfs_multi(srcs, want_parents=False)
    elif args.only == "pr":
        def f():
            for prec in (B.PR_FP32, B.PR_FP64):
                G.pagerank(0.85, 20, 0.0, prec, want_p=False)
    elif args.only == "map":
        one = np.zeros(n * 64, np.uint8)
        one[0] = 1

        def f():
            G.map_sum()
            for w in (4, 64):
                G.map_remote(one, w)
    elif args.only == "cc":
        f = G.cc
    elif args.only == "tc":
        f = G.tc
    else:
        C = B.CpmaGraph.from_graph(G)
        for kv in args.set:
            k, v = kv.split("=", 1)
            C.set_tuning(k, int(v))
        keys = B.update_rmat_keys(n, args.cpma_batch)
        d = B.DeviceBuffer(keys)

        def f():
            C.insert_batch_dev(d.ptr, len(keys))
            C.remove_batch_dev(d.ptr, len(keys))
    for _ in range(args.warmup):
        f()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f()
    res["ms_per_run"] = round((time.perf_counter() - t0) / args.steps * 1e3, 4)
    if args.only == "cpma":
        d.free()
        C.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--sources", type=int, default=8)
    ap.add_argument("--cpma-scale", type=int, default=24)
    ap.add_argument("--cpma-batch", type=int, default=10_000_000)
    ap.add_argument("--job-scale", type=int, default=23, help="scale of the batched job graph (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=8.0)
    ap.add_argument("--only", choices=["bfs", "pr", "map", "cc", "tc", "cpma"], default=None,
                    help="run one section a few times and print its own short JSON line "
                         "(for per-kernel profiling runs)")
    ap.add_argument("--set", action="append", default=[], metavar="KNOB=VALUE",
                    help="tuning knob override applied after the stored tuning (repeatable)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (the library has no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import byo_amd as B
    if B.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    t = time.time()
    G = B.CsrGraph.rmat(args.scale, args.ef)
    n, m = G.N(), G.M()
    deg = G.degrees()
    srcs = pick_sources(deg, args.sources)
    print("# graph RMAT-%d ef%d: n=%d m=%d (built on device in %.1fs)" % (args.scale, args.ef, n, m,
                                                                          time.time() - t), file=sys.stderr)

    knobs = G.apply_tuning()  # autotuned block/tile sizes for this graph size, if stored
    for kv in args.set:
        k, v = kv.split("=", 1)
        G.set_tuning(k, int(v))
        knobs[k] = int(v)
    from byo_amd import selector as S
    learned = bool(S.load(G).get("batch"))
    policy = B.DIR_MLP if learned else B.DIR_THRESHOLD
    if args.only:
        print(json.dumps(only_section(B, G, args, srcs, policy)), flush=True)
        G.close()
        return

    def timed_bfs(pol, steps, mode=None):
        """K steps of the 8-source BFS under a direction policy; returns
        (GTEPS by the wall clock, launches, launch ms, algorithmic bytes,
        stats of the last step)."""
        G.set_direction_policy(pol)
        if mode:
            G.set_tuning("bfs_engine", mode)
        try:
            stats = None
            for _ in range(args.warmup):
                _, stats = G.bfs_multi(srcs, want_parents=False)
            G.profile(True)
            G.profile_reset()
            sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                G.bfs_multi(srcs, want_parents=False, want_stats=False)
            sync()
            wall = time.perf_counter() - t0
            nl, ms = G.profile_get("bfs_launch")
            # reached vertices and edges are fixed by the sources: the timed
            # launches read no statistics back; one launch after them counts
            # them (and matches the warmup's)
            _, after = G.bfs_multi(srcs, want_parents=False)
            if stats is not None:
                assert [(a["reached"], a["edges_reached"]) for a in after] == \
                    [(a["reached"], a["edges_reached"]) for a in stats], "BFS results changed between launches"
            stats = after
            edges = steps * sum(st["edges_reached"] // 2 for st in stats)
            nbytes = steps * sum(bfs_bytes(n, st["reached"], st["edges_reached"]) for st in stats)
        finally:
            if mode:
                G.set_tuning("bfs_engine", 0)
            G.set_direction_policy(B.DIR_THRESHOLD)
        return edges / wall / 1e9, wall, nl, ms, nbytes, stats

    gteps, wall, nlaunch, launch_ms, bytes_total, stats = timed_bfs(policy, args.steps)
    thr_gteps, _, thr_nl, thr_ms, _, thr_stats = timed_bfs(B.DIR_THRESHOLD, args.steps)

    def level_arrays(pol):
        """Per-source BFS level arrays of one 8-source launch (parents read
        back, levels derived on the host; outside the timed steps)."""
        G.set_direction_policy(pol)
        par, _ = G.bfs_multi(srcs, want_parents=True)
        G.set_direction_policy(B.DIR_THRESHOLD)
        return [levels_from_parents(par[j], srcs[j]) for j in range(len(srcs))]

    levels_identical = all(np.array_equal(x, y) for x, y in zip(level_arrays(policy),
                                                                level_arrays(B.DIR_THRESHOLD)))
    # the same graph with 64 sources in one shared traversal (the batched
    # job's width), rule against learned selector, runs interleaved
    srcs64 = pick_sources(deg, 64, seed=2)
    best64 = {}
    for r in range(4):
        for name, pol in (("rule", B.DIR_THRESHOLD), ("learned", policy)):
            G.set_direction_policy(pol)
            G.profile(True)
            G.profile_reset()
            _, st64 = G.bfs_multi(srcs64, want_parents=False)
            ms64 = G.profile_get("bfs_launch")[1]
            if r:
                best64[name] = min(best64.get(name, 1e30), ms64)
            if name == "learned":
                dirs64 = st64[0]["dir"][:max(s["levels"] for s in st64)]
    G.set_direction_policy(B.DIR_THRESHOLD)
    edges64 = sum(s["edges_reached"] // 2 for s in st64)
    seq_gteps = timed_bfs(B.DIR_THRESHOLD, 1, 1)[0]  # bfs_engine 1: one source at a time
    # the dominant kernel is the persistent BFS launch (all sources of a step
    # in one k_bfs_batch dispatch), timed with HIP events on its stream
    avg_ms = launch_ms / max(nlaunch, 1)
    avg_bytes = bytes_total / max(nlaunch, 1)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_bfs_batch_rmat22_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": "BFS GTEPS learned vs threshold; FP32 PR it/s @1e-6 L1; insert edges/s",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic RMAT generated on the device (seed 1)",
        "config": {"workload": "BFS x%d sources in one shared traversal, RMAT scale %d ef %d (a=.5,b=.1,c=.1), "
                               "symmetric CSR, direction by the %s" % (
                                   args.sources, args.scale, args.ef,
                                   "learned selector" if learned else "M/20 rule (no selector weights)"),
                   "n": n, "m": m, "sources": srcs, "levels_dirs": stats[0]["dir"][:max(s["levels"] for s in stats)],
                   "tuning": knobs},
        "learned_vs_threshold": {"learned_gteps": round(gteps, 3), "threshold_gteps": round(thr_gteps, 3),
                                 "speedup": round(gteps / thr_gteps, 4),
                                 "launch_ms_learned": round(avg_ms, 4),
                                 "launch_ms_threshold": round(thr_ms / max(thr_nl, 1), 4),
                                 "threshold_dirs": thr_stats[0]["dir"][:max(s["levels"] for s in thr_stats)],
                                 "levels_identical": levels_identical},
        "learned_vs_threshold_64_sources": {
            "sources": 64, "rule_ms": round(best64["rule"], 4), "learned_ms": round(best64["learned"], 4),
            "speedup": round(best64["rule"] / best64["learned"], 4),
            "learned_gteps": round(edges64 / (best64["learned"] * 1e-3) / 1e9, 2), "learned_dirs": dirs64},
        "one_source_engine_gteps": round(seq_gteps, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_bfs_batch<uint8_t> (one launch = %d BFS)" % len(srcs), "avg_ms": round(avg_ms, 4),
                     "algorithmic_bytes": int(avg_bytes)},
    }

    if not args.no_extras:
        ex = {}
        relabel = G.get_tuning("pr_relabel")
        G.profile(True)
        G.profile_reset()
        p32, l1_32, ms32 = G.pagerank(0.85, 20, 0.0, B.PR_FP32)  # builds the relabelled copy if knob on
        ex["pr_relabel"] = relabel
        ex["pr_relabel_build_ms"] = round(G.profile_get("relabel")[1], 3)
        xcd = G.get_tuning("pr_xcd")
        ex["pr_xcd"] = xcd
        ex["pr_xcd_slices_build_ms"] = round(G.profile_get("pr_xcd_slices")[1], 3)
        ex["pr_pb_layout_build_ms"] = round(G.profile_get("pr_pb_layout")[1], 3)
        G.profile(False)
        p32, l1_32, ms32 = G.pagerank(0.85, 20, 0.0, B.PR_FP32)
        p64, l1_64, ms64 = G.pagerank(0.85, 20, 0.0, B.PR_FP64)
        p64, l1_64, ms64 = G.pagerank(0.85, 20, 0.0, B.PR_FP64)
        ex["pr_fp32_it_s"] = round(20 / (ms32 * 1e-3), 1)
        ex["pr_fp64_it_s"] = round(20 / (ms64 * 1e-3), 1)
        ex["pr_fp32_vs_fp64_speedup"] = round(ms64 / ms32, 3)
        ex["pr_l1_fp32_vs_fp64"] = float(np.abs(p32 - p64).sum())
        w = 4
        pr_bytes = 8 * (n + 1) + (4 + w) * m + 3 * w * n
        ex["pr_fp32_GBs"] = round(pr_bytes / (ms32 / 20 * 1e-3) / 1e9, 1)
        G.set_tuning("pr_relabel", 1 - relabel)  # the other layout, for comparison
        for prec, key in ((B.PR_FP32, "fp32"), (B.PR_FP64, "fp64")):
            G.pagerank(0.85, 20, 0.0, prec, want_p=False)
            ex["pr_%s_it_s_relabel_%d" % (key, 1 - relabel)] = round(
                20 / (G.pagerank(0.85, 20, 0.0, prec, want_p=False)[2] * 1e-3), 1)
        G.set_tuning("pr_relabel", relabel)
        G.set_tuning("pr_xcd", 1 - xcd)  # with and without the per-XCD source slices
        for prec, key in ((B.PR_FP32, "fp32"), (B.PR_FP64, "fp64")):
            G.pagerank(0.85, 20, 0.0, prec, want_p=False)
            ex["pr_%s_it_s_xcd_%d" % (key, 1 - xcd)] = round(
                20 / (G.pagerank(0.85, 20, 0.0, prec, want_p=False)[2] * 1e-3), 1)
        G.set_tuning("pr_xcd", xcd)
        pb = G.get_tuning("pr_pb")
        ex["pr_pb"] = pb
        G.set_tuning("pr_pb", 1 - pb)  # propagation blocking against the pull (other knobs as tuned)
        for prec, key in ((B.PR_FP32, "fp32"), (B.PR_FP64, "fp64")):
            G.pagerank(0.85, 20, 0.0, prec, want_p=False)
            ex["pr_%s_it_s_pb_%d" % (key, 1 - pb)] = round(
                20 / (G.pagerank(0.85, 20, 0.0, prec, want_p=False)[2] * 1e-3), 1)
        G.set_tuning("pr_pb", pb)
        G.cc()
        lab, cc_ms = G.cc()
        ex["cc_ms"] = round(cc_ms, 3)
        ex["cc_components"] = int(len(np.unique(lab)))
        G.bc(srcs[0])
        _, bc_ms = G.bc(srcs[0])
        ex["bc_ms"] = round(bc_ms, 3)  # single-source betweenness (BFS + both passes)
        G.pagerank_delta(want_p=False)
        _, pd_rounds, pd_ms = G.pagerank_delta(want_p=False)  # run_all's -delta settings
        ex["pr_delta_ms"] = round(pd_ms, 3)
        ex["pr_delta_rounds"] = int(pd_rounds)
        G.ldd(0.2)
        ldd_cl, ldd_rounds, ldd_ms = G.ldd(0.2)
        ex["ldd_ms"] = round(ldd_ms, 3)
        ex["ldd_clusters"] = int(len(np.unique(ldd_cl)))
        G.mis()
        mis_in, mis_rounds, mis_ms = G.mis()
        ex["mis_ms"] = round(mis_ms, 3)
        ex["mis_size"] = int(mis_in.sum())
        ex["mis_rounds"] = int(mis_rounds)
        G.spanner()
        sp_e, sp_ms = G.spanner()  # run_all's k = 4
        ex["spanner_ms"] = round(sp_ms, 3)
        ex["spanner_edges"] = int(len(sp_e))
        G.densest_subgraph()
        ds_d, ds_rounds, ds_ms = G.densest_subgraph()  # run_all's eps 0.001
        ex["densest_ms"] = round(ds_ms, 3)
        ex["densest_density"] = round(ds_d, 6)
        ex["densest_rounds"] = int(ds_rounds)
        G.coloring()
        col, col_rounds, col_ms = G.coloring()
        ex["coloring_ms"] = round(col_ms, 3)
        ex["coloring_colours"] = int(col.max()) + 1 if len(col) else 0
        ex["coloring_rounds"] = int(col_rounds)
        G.kcore()
        _, kc_kmax, kc_rounds, kc_ms = G.kcore()
        ex["kcore_ms"] = round(kc_ms, 3)
        ex["kcore_kmax"] = int(kc_kmax)
        ex["kcore_rounds"] = int(kc_rounds)
        G.tc()
        tri, tc_ms = G.tc()
        ex["tc_ms"] = round(tc_ms, 3)
        ex["triangles"] = int(tri)
        G.map_sum()
        # the edge scan's rate varies between runs and boxes by ~10 %: seven
        # runs, the median and the spread
        map_rates = sorted((8 * (n + 1) + 4 * m) / (G.map_sum()[1] * 1e-3) / 1e9 for _ in range(7))
        ex["map_GBs"] = round(map_rates[3], 1)
        ex["map_GBs_min_max"] = [round(map_rates[0], 1), round(map_rates[-1], 1)]
        # Map_withRemote_<w>: one gather of a w-byte element per edge, the
        # runner's one-hot input; G gathers/s for each element width
        mr = {}
        for w in (0, 1, 2, 4, 8, 16, 32, 64):
            one = np.zeros(max(n * w, (n + 7) // 8), np.uint8)
            one[0] = 1
            G.map_remote(one, w)
            mr[str(w)] = round(m / (min(G.map_remote(one, w)[1] for _ in range(3)) * 1e-3) / 1e9, 1)
        ex["map_remote_Gedges_s"] = mr
        # CPMA batch updates at batch size cpma_batch on RMAT-cpma_scale
        if args.cpma_scale == args.scale:
            H = G
        else:
            H = B.CsrGraph.rmat(args.cpma_scale, args.ef)
        C = B.CpmaGraph.from_graph(H)
        for kv in args.set:  # --set knobs reach the container's update path too
            k, v = kv.split("=", 1)
            C.set_tuning(k, int(v))
        keys = B.update_rmat_keys(C.N(), args.cpma_batch)
        dkeys = B.DeviceBuffer(keys)
        C.profile(True)
        C.insert_batch_dev(dkeys.ptr, len(keys))  # round 0 (untimed, as Batch_insert_runner)
        C.remove_batch_dev(dkeys.ptr, len(keys))
        C.profile_reset()
        ins_t, rem_t, ins_s, rem_s = [], [], [], []
        for r in range(3):
            sync()
            s0 = C.stats()["touched_bytes"]
            a = time.perf_counter()
            C.insert_batch_dev(dkeys.ptr, len(keys))
            b = time.perf_counter()
            s1 = C.stats()["touched_bytes"]
            C.remove_batch_dev(dkeys.ptr, len(keys))
            c = time.perf_counter()
            s2 = C.stats()["touched_bytes"]
            ins_t.append(b - a)
            rem_t.append(c - b)
            ins_s.append(s1 - s0)
            rem_s.append(s2 - s1)
        ex["cpma_batch"] = args.cpma_batch
        ex["cpma_insert_edges_s"] = round(args.cpma_batch / np.mean(ins_t), 1)
        ex["cpma_remove_edges_s"] = round(args.cpma_batch / np.mean(rem_t), 1)
        ex["cpma_scale"] = args.cpma_scale
        # update byte model 24 b + 2 S_touched (batch read, one sort pass in
        # and out, PMA slots rewritten read + write; S counted by the kernels)
        bb = args.cpma_batch
        ib = 24 * bb + 2 * float(np.mean(ins_s))
        rb = 24 * bb + 2 * float(np.mean(rem_s))
        ex["cpma_roofline"] = {
            "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "insert": {"achieved": round(ib / np.mean(ins_t) / 1e9, 1), "frac": round(ib / np.mean(ins_t) / 1e9 / HBM_PEAK_GBS, 4),
                       "s_touched_bytes": int(np.mean(ins_s)), "algorithmic_bytes": int(ib)},
            "remove": {"achieved": round(rb / np.mean(rem_t) / 1e9, 1), "frac": round(rb / np.mean(rem_t) / 1e9 / HBM_PEAK_GBS, 4),
                       "s_touched_bytes": int(np.mean(rem_s)), "algorithmic_bytes": int(rb)}}
        # Batch_insert_runner's sweep (run_unweighted.h:781-893): b = 10 ..
        # 1e7 in powers of 10, the same update stream prefix for every size,
        # rounds + 1 with round 0 untimed, insert then remove the batch
        sweep = {}
        b = 10
        while b <= args.cpma_batch:
            ti, tr = [], []
            for r in range(4):
                a = time.perf_counter()
                C.insert_batch_dev(dkeys.ptr, b)
                m1 = time.perf_counter()
                C.remove_batch_dev(dkeys.ptr, b)
                m2 = time.perf_counter()
                if r:
                    ti.append(m1 - a)
                    tr.append(m2 - m1)
            sweep[str(b)] = {"insert_edges_s": round(b / np.mean(ti), 1), "remove_edges_s": round(b / np.mean(tr), 1)}
            b *= 10
        ex["cpma_sweep"] = sweep
        # C3's rerun: BFS and PageRank on the updated graph, read from the PMA
        # itself (a symmetric batch of 1e7 keys inserted into a fresh copy of
        # the graph, so it stays undirected: removing the directed batches
        # above also removed the batch keys that were edges before, and left
        # their reverse edges); the slot-index rebuild after the update is
        # timed separately, and the same algorithms on the CSR of the graph
        # before the update are given for comparison
        C.close()
        C = B.CpmaGraph.from_graph(H)
        C.profile(True)
        half = keys[: args.cpma_batch // 2]
        symk = np.concatenate([half, ((half & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (half >> np.uint64(32))])
        dsym = B.DeviceBuffer(symk)
        C.insert_batch_dev(dsym.ptr, len(symk))
        C.profile_reset()
        rr = {"inserted_keys": int(len(symk)), "m": C.M()}
        csrcs = pick_sources(C.degrees(), args.sources, seed=4)
        C.apply_tuning()
        _, cst = C.bfs_multi(csrcs, want_parents=False)  # builds the slot index
        rr["index_build_ms"] = round(C.profile_get("cpma_index")[1], 3)
        C.profile_reset()
        for _ in range(3):
            _, cst = C.bfs_multi(csrcs, want_parents=False)
        nl, lms = C.profile_get("bfs_launch")
        rr["pma_bfs_gteps"] = round(sum(x["edges_reached"] // 2 for x in cst) / (lms / nl * 1e-3) / 1e9, 2)
        rr["csr_export_launches_bfs"] = C.profile_get("cpma_csr_export")[0]
        C.pagerank(0.85, 20, 0.0, B.PR_FP32, want_p=False)  # builds the PageRank layout (slices / blocking)
        rr["pma_pr_pb"] = C.get_tuning("pr_pb")
        rr["pma_pr_xcd"] = C.get_tuning("pr_xcd")
        rr["pma_pr_xcd_slices_build_ms"] = round(C.profile_get("pr_xcd_slices")[1], 3)
        # propagation blocking cuts its layout from the CPMA's CSR export (once
        # per update, timed with it); the iterations read the layout only
        rr["pma_pr_pb_layout_build_ms"] = round(C.profile_get("pr_pb_layout")[1] + C.profile_get("cpma_csr_export")[1], 3)
        rr["pma_pr_fp32_it_s"] = round(20 / (C.pagerank(0.85, 20, 0.0, B.PR_FP32, want_p=False)[2] * 1e-3), 1)
        rr["csr_export_launches"] = C.profile_get("cpma_csr_export")[0]
        if H is not G:
            H.apply_tuning()
            H.profile(True)
            H.profile_reset()
            hsrcs = pick_sources(H.degrees(), args.sources, seed=4)
            for _ in range(3):
                _, hst = H.bfs_multi(hsrcs, want_parents=False)
            nl, lms = H.profile_get("bfs_launch")
            rr["csr_before_update_bfs_gteps"] = round(sum(x["edges_reached"] // 2 for x in hst) / (lms / nl * 1e-3) / 1e9, 2)
            H.pagerank(0.85, 20, 0.0, B.PR_FP32, want_p=False)
            rr["csr_before_update_pr_fp32_it_s"] = round(
                20 / (H.pagerank(0.85, 20, 0.0, B.PR_FP32, want_p=False)[2] * 1e-3), 1)
        C.remove_batch_dev(dsym.ptr, len(symk))
        dsym.free()
        ex["cpma_rerun"] = rr
        dkeys.free()
        C.close()
        cpma_keys = keys
        if args.job_scale:
            ex["batched_job"] = batched_job(B, args)
        out["extras"] = ex
        out["c1_plain_method"] = c1_plain_method(B)

    if not args.no_cpu:
        out["cpu_baseline"], g = cpu_baseline(G, srcs, args.cpu_budget)
        if not args.no_extras:
            out["cpu_baselines_other"] = cpu_other(g, H if H is not G else None, cpma_keys)
    if not args.no_extras and H is not G:
        H.close()
    # every device handle is released here, before the interpreter exits
    G.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
