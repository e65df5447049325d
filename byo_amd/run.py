"""Command-line runner over the device containers: the host side of
run_csr / run_single_cpma (benchmarks/run_structures/run_csr.cc:192-225,
run_single_pma.cc) driving run_all (benchmarks/run_unweighted.h:930-1034)
with the reference's runners and timing protocol.

    python -m byo_amd.run -s [-b] [-rounds R] [-src S] [-i] [-max_batch B]
                          [--cpma] [--fp32] (<AdjacencyGraph file> | -rmat SCALE [-ef EF])

Flags follow run_csr.cc:193-204: -s symmetric input (required: directed
graphs are refused, as there), -b the binary AdjacencyGraph layout, -c
refused ("reads in uncompressed files"), -rounds (default 3), -src (default
0), -i updates only (Batch_insert_runner instead of the algorithms), -max_batch
(default 1000000), -d accepted and ignored (no output arrays are dumped).
Additions: --cpma runs on the CPMA container (run_single_cpma), --fp32 runs
PageRank in FP32-Kahan instead of FP64, -rmat SCALE builds the RMAT graph of
that scale on the device instead of reading a file.

Every runner follows run_unweighted.h:135-170: rounds + 1 runs, round 0
untimed, the mean of the rest printed as "# time per iter"; run_all ends with
one "# # # <algorithm>, <seconds>" line per runner.  Times are wall-clock
seconds of the library call (results left on the device, as every entry
point returns after its device work has completed).  Batch updates take the
update stream on the host, as the reference's do, so their times include
the PCIe copy of the batch.
"""
import argparse
import sys
import time

import numpy as np

from . import (DIR_THRESHOLD, PR_FP32, PR_FP64, CpmaGraph, CsrGraph, device_count, read_adjacency,
               update_rmat_keys)


def _runner(name, f, rounds, params=None, report=None):
    """rounds + 1 runs of f, round 0 untimed; prints the reference's lines
    and returns the mean time of the timed runs."""
    print("### Application: %s" % name)
    if params:
        print("### Params: %s" % params)
    print("### ------------------------------------")
    print("### ------------------------------------")
    total = 0.0
    for r in range(rounds + 1):
        t0 = time.perf_counter()
        res = f()
        tt = time.perf_counter() - t0
        print("### Running Time: %g" % tt)
        if r == 0:
            if report:
                report(res)
        else:
            total += tt
    per = total / rounds if rounds else 0.0
    print("# time per iter: %g" % per)
    sys.stdout.flush()
    return per


def run_all(G, rounds=3, src=0, inserts=False, max_batch=1_000_000, pr_precision=PR_FP64):
    """run_unweighted.h:930-1034 over a device container; returns the
    {algorithm: seconds} map it prints."""
    n, m = G.N(), G.M()
    print("### Threads: device")
    print("### n: %d" % n)
    print("### m: %d" % m)
    tm = {}
    if not inserts:
        tm["Map"] = _runner("Base Map", G.map_sum, rounds, report=lambda r: print("Edgesum was %d" % r[0]))
        rng = np.random.default_rng(0)
        for w in (0, 1, 2, 4, 8, 16, 32, 64):
            remote = np.zeros(max(n * w, (n + 7) // 8), np.uint8)
            v = int(rng.integers(0, max(n, 1)))
            if w == 0:
                remote[v >> 3] = 1 << (v & 7)
            else:
                remote[v * w] = 1
            tm["Map_withRemote_%d" % w] = _runner(
                "Map with Remote: %d" % w, lambda: G.map_remote(remote, w), rounds,
                report=lambda r: print("Edgesum was %d" % r[0]))
        G.set_direction_policy(DIR_THRESHOLD)
        tm["BFS"] = _runner("BFS", lambda: G.bfs(src, want_parents=False), rounds, "-src = %d" % src)
        tm["WorkEfficientDensestSubgraph"] = _runner(
            "WorkEfficientDensestSubgraph", lambda: G.densest_subgraph(0.001), rounds,
            report=lambda r: print("### Density of 2(1+eps)-Densest Subgraph is: %g" % r[0]))
        tm["SimpleUnionCC"] = _runner(
            "SimpleUnionCC", G.cc, rounds,
            report=lambda r: print("num components = %d" % len(np.unique(r[0]))))
        tm["Coloring"] = _runner("Coloring", G.coloring, rounds,
                                 report=lambda r: print("num colors = %d" % (int(r[0].max()) + 1 if n else 0)))
        tm["KCore"] = _runner("KCore", G.kcore, rounds, report=lambda r: print("### Max core: %d" % r[1]))
        tm["LDD"] = _runner("LDD", lambda: G.ldd(0.2), rounds, "-beta = 0.2")
        tm["MIS"] = _runner("MIS", G.mis, rounds, report=lambda r: print("MIS size: %d" % int(r[0].sum())))
        tm["PageRank"] = _runner(
            "PageRank", lambda: G.pagerank(0.85, 20, 1e-4, pr_precision, want_p=False), rounds,
            report=lambda r: print("### iterations: %d, last L1: %g" % (len(r[1]), r[1][-1] if len(r[1]) else 0)))
        tm["SSBetweennessCentrality"] = _runner("SSBetweennessCentrality", lambda: G.bc(src), rounds,
                                                "-src = %d" % src)
        tm["Spanner"] = _runner("Spanner", lambda: G.spanner(4), rounds, "-k = 4",
                                report=lambda r: print("spanner size: %d" % len(r[0])))
        if m < 1_000_000_000:
            tm["TriangleCount"] = _runner("TriangleCount", G.tc, rounds,
                                          report=lambda r: print("### Num triangles = %d" % r[0]))
    else:
        if not isinstance(G, CpmaGraph):
            raise SystemExit("-i needs a container with insert_batch (--cpma)")
        batch_insert_runner(tm, G, rounds, max_batch)
    for alg in sorted(tm):
        print("# # # %s, %g" % (alg, tm[alg]))
    sys.stdout.flush()
    return tm


def batch_insert_runner(tm, G, rounds, max_batch, seed=0):
    """Batch_insert_runner (run_unweighted.h:781-893): b = 1, 10, ...,
    max_batch; each size rounds + 1 times (round 0 untimed): insert the
    update stream's first b keys (the same prefix every round and size),
    shuffle, remove them.  The shuffle is seeded (the reference draws a
    random_device seed)."""
    n = G.N()
    rng = np.random.default_rng(seed)
    b = 1
    while b <= max_batch:
        print("batch size: %d" % b)
        ins = rem = 0.0
        keys = update_rmat_keys(n, b)
        for ts in range(rounds + 1):
            t0 = time.perf_counter()
            G.insert_batch(keys)
            dt = time.perf_counter() - t0
            if ts > 0:
                ins += dt
            print("done inserts in %g" % dt)
            print("sorts took 0")
            sh = keys.copy()
            rng.shuffle(sh)
            t0 = time.perf_counter()
            G.remove_batch(sh)
            dt = time.perf_counter() - t0
            if ts > 0:
                rem += dt
            print("done deletes in %g" % dt)
        tm["batch_insert_%d" % b] = ins / rounds if rounds else 0.0
        tm["batch_remove_%d" % b] = rem / rounds if rounds else 0.0
        tm["batch_sort_%d" % b] = 0.0  # sorting is part of the insert (graph.h:3090-3092)
        sys.stdout.flush()
        b *= 10


def parse(argv):
    ap = argparse.ArgumentParser(prog="python -m byo_amd.run", description=__doc__.split("\n\n")[0])
    ap.add_argument("file", nargs="?")
    ap.add_argument("-s", action="store_true", help="symmetric input (required)")
    ap.add_argument("-c", action="store_true", help="compressed input (refused)")
    ap.add_argument("-b", action="store_true", help="binary AdjacencyGraph layout")
    ap.add_argument("-m", action="store_true", help="mmap (accepted; files are read whole)")
    ap.add_argument("-d", action="store_true", help="dump output arrays (accepted, ignored)")
    ap.add_argument("-i", action="store_true", help="run only the batch updates")
    ap.add_argument("-rounds", type=int, default=3)
    ap.add_argument("-src", type=int, default=0)
    ap.add_argument("-max_batch", type=int, default=1_000_000)
    ap.add_argument("-rmat", type=int, default=None, help="RMAT scale generated on the device instead of a file")
    ap.add_argument("-ef", type=int, default=16, help="RMAT edge factor")
    ap.add_argument("--cpma", action="store_true", help="run on the CPMA container (run_single_cpma)")
    ap.add_argument("--fp32", action="store_true", help="PageRank in FP32-Kahan")
    return ap.parse_args(argv)


def main(argv=None):
    a = parse(sys.argv[1:] if argv is None else argv)
    if a.c:
        print("reads in uncompressed files", file=sys.stderr)
        return -1
    if not a.s:
        print("does not support directed graphs yet", file=sys.stderr)
        return -1
    if (a.file is None) == (a.rmat is None):
        print("give exactly one of <inFile> or -rmat SCALE", file=sys.stderr)
        return -1
    if device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return -1
    print("### Graph: %s" % (a.file if a.file else "RMAT scale %d ef %d (device)" % (a.rmat, a.ef)))
    if a.file:
        off, ids = read_adjacency(a.file, a.b)
        G = (CpmaGraph if a.cpma else CsrGraph).from_csr(off, ids)
    else:
        C = CsrGraph.rmat(a.rmat, a.ef)
        if a.cpma:
            G = CpmaGraph.from_graph(C)
            C.close()
        else:
            G = C
    try:
        print("total bytes used = %d" % G.get_memory_size())
        if not 0 <= a.src < max(G.N(), 1):
            print("-src out of range", file=sys.stderr)
            return -1
        run_all(G, a.rounds, a.src, a.i, a.max_batch, PR_FP32 if a.fp32 else PR_FP64)
    finally:
        G.close()
    return 1  # run_csr.cc:225 returns 1


if __name__ == "__main__":
    rc = main()
    sys.exit(0 if rc == 1 else 1)
