"""Autotuner of the block and tile sizes (the tuning knobs of byo_set_tuning).

For one graph it times every knob's candidates on the device -- the other
knobs held at their best so far (one coordinate-descent pass) -- keeps the
fastest, and stores the winners per graph size in byo_amd/tuning.json.
Graph.apply_tuning() later applies the entry of the closest size.  What each
knob is timed on:

    coop_blocks_per_cu, pieces,  the benchmark's step exactly: its 8 sources
    pull_dyn, coop_grid_sqrt_div (SplitMix64 seed 1) in one shared traversal
                                 under the committed learned selector
                                 (launch ms, best of 7)
    ppr_blocks_per_cu            32-seed personalized PageRank, 5 iterations
    pr_relabel, pr_xcd,          FP32 PageRank, 5 iterations (the one-time
    pr_xcd_vw, pr_lanes, pr_grid layouts are built before the timing)
    pr_pb

    python -m byo_amd.autotune --scale 22 --scale 23
    python -m byo_amd.autotune --scale 22 --only pr_xcd_vw   # one knob, the
                                 # entry's other knobs as stored
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING = os.path.join(HERE, "tuning.json")

CANDIDATES = {
    "coop_blocks_per_cu": [2, 3, 4],
    "pieces": [4, 8, 16, 32],
    "pull_dyn": [0, 4, 8, 16],
    "coop_grid_sqrt_div": [0, 8, 14, 20],
    "ppr_blocks_per_cu": [4, 8],
    "ppr_pieces": [2, 4, 8],
    "pr_relabel": [0, 1],
    "pr_xcd": [0, 1],
    "pr_xcd_vw": [0, 2, 4, 6, 8, 12, 16, 24],
    "pr_lanes": [0, 4, 8, 16, 32],
    "pr_grid": [4096, 8192, 16384, 32768, 65536],
    "pr_pb": [0, 1],
}
WORKLOAD = {"coop_blocks_per_cu": "bfs", "pieces": "bfs", "pull_dyn": "bfs", "coop_grid_sqrt_div": "bfs", "ppr_blocks_per_cu": "ppr", "ppr_pieces": "ppr", "pr_lanes": "pr",
            "pr_grid": "pr", "pr_relabel": "pr", "pr_xcd": "pr",
            "pr_xcd_vw": "pr", "pr_pb": "pr"}


def _sources(deg, k, seed):
    from .selector import sources
    return sources(deg, k, seed)


def measure(G, workload, srcs, seeds, repeats=3):
    """Best-of-repeats device ms of one workload (BFS: best of 7)."""
    from . import PR_FP32
    best = float("inf")
    for _ in range(7 if workload == "bfs" else repeats):
        if workload == "bfs":
            G.profile(True)
            G.profile_reset()
            G.bfs_multi(srcs, want_parents=False)
            ms = G.profile_get("bfs_launch")[1]
        elif workload == "ppr":
            ms = G.ppr(seeds, 0.85, 5, want_p=False)[1]
        else:
            ms = G.pagerank(0.85, 5, 0.0, PR_FP32, want_p=False)[2]
        best = min(best, ms)
    return best


def tune(G, label="", only=None, start=None):
    """Coordinate descent over CANDIDATES (or the keys in only, the others
    set to start); returns (knobs, times).  The BFS knobs are timed on the
    launch bench.py times: the same sources and the same direction policy
    (the learned selector when weights are committed)."""
    from . import DIR_MLP
    from . import selector as S
    deg = G.degrees()
    srcs = _sources(deg, 8, 1)
    seeds = _sources(deg, 32, 12)
    if S.load(G).get("batch"):
        G.set_direction_policy(DIR_MLP)
    knobs, times = {}, {}
    for key, v in (start or {}).items():
        G.set_tuning(key, v)
    for key, cands in CANDIDATES.items():
        if only and key not in only:
            continue
        wl = WORKLOAD[key]
        measure(G, wl, srcs, seeds, 1)  # warm
        res = {}
        for v in cands:
            G.set_tuning(key, v)
            res[v] = measure(G, wl, srcs, seeds)
        best = min(res, key=res.get)
        G.set_tuning(key, best)
        knobs[key] = best
        times[key] = {str(v): round(t, 4) for v, t in res.items()}
        print("%s %s: %s -> %s" % (label, key, times[key], best), flush=True)
    return knobs, times


def load(path=None):
    p = path or TUNING
    if not os.path.exists(p):
        return {"entries": []}
    with open(p) as f:
        return json.load(f)


def lookup(n, m, path=None):
    """Knobs of the stored entry closest in log2 of the edge count (and of the
    vertex count on ties); {} when nothing is stored."""
    entries = load(path)["entries"]
    if not entries or m <= 0:
        return {}
    def dist(e):
        return (abs(math.log2(max(e["m"], 1)) - math.log2(m)), abs(math.log2(max(e["n"], 1)) - math.log2(max(n, 1))))
    return dict(min(entries, key=dist)["knobs"])


def main():
    from . import CsrGraph
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, action="append", default=None)
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--out", default=TUNING)
    ap.add_argument("--only", action="append", default=None, choices=sorted(CANDIDATES),
                    help="tune only this knob (repeatable); the stored entry keeps the rest")
    args = ap.parse_args()
    db = load(args.out)
    for scale in args.scale or [22]:
        t0 = time.time()
        G = CsrGraph.rmat(scale, args.ef)
        label = "rmat%d_ef%d" % (scale, args.ef)
        old = next((e for e in db["entries"] if e["graph"] == label), None)
        if args.only and old is None:
            raise SystemExit("--only needs a stored entry for %s" % label)
        knobs, times = tune(G, label, args.only, old["knobs"] if args.only else None)
        if args.only:
            knobs = dict(old["knobs"], **knobs)
            times = dict(old.get("times_ms", {}), **times)
        entry = {"graph": label, "n": G.N(), "m": G.M(), "knobs": knobs, "times_ms": times}
        db["entries"] = [e for e in db["entries"] if e["graph"] != label] + [entry]
        G.close()
        print("%s tuned in %.0fs" % (label, time.time() - t0), flush=True)
    with open(args.out, "w") as f:
        json.dump(db, f, indent=1)
    print(json.dumps(db), flush=True)


if __name__ == "__main__":
    sys.exit(main())
