"""Writes tests/golden/golden.json: known answers used by the parity tests.

Run from the repository root:  python tests/golden/make_golden.py

Sources of truth, per entry:
  * inputs/*        -- the reference's own fixture graphs (AdjacencyGraph text,
                       copied verbatim into tests/golden/inputs/); facts are
                       computed by brute force (triple enumeration, FIFO BFS,
                       DSU) and cross-checked against the oracle's fast paths.
  * update_rmat_*   -- Batch_insert_runner's update stream
                       (run_unweighted.h:781-806); checked independently in
                       tests/test_oracle.py by a pure-Python restatement.
  * rmat_s16_*      -- this framework's own graph generator (prefix and
                       summary facts), which the device generator must match.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402


def fixture_facts(path, src):
    g = O.read_adjacency_graph(path)
    lv = O.bfs_levels(g, src)
    par, dirlog, fs, fd = O.bfs_diropt(g, src)
    lab = O.cc(g)
    reached = lv != O.NONE
    return {
        "n": g.n, "m": g.m, "src": src,
        "triangles": O.tc_brute(g),
        "components": int(len(np.unique(lab))),
        "isolated": int((np.diff(g.offsets) == 0).sum()),
        "max_degree": int(np.diff(g.offsets).max()),
        "bfs_reached": int(reached.sum()),
        "bfs_levels": int(lv[reached].max()) + 1,
        "levels": [int(x) for x in lv],
        "cc_labels": [int(x) for x in lab],
        "dirlog": [int(x) for x in dirlog],
        "map_sum": O.map_sum(g),
    }


def main():
    out = {}
    ins = os.path.join(HERE, "inputs")
    out["rMatGraph_J_5_100"] = fixture_facts(os.path.join(ins, "rMatGraph_J_5_100"), 10)
    out["triangles.txt"] = fixture_facts(os.path.join(ins, "triangles.txt"), 0)
    out["star.txt"] = fixture_facts(os.path.join(ins, "star.txt"), 0)

    s, d = O.update_rmat(1 << 16, 64)
    out["update_rmat_n65536_first64"] = [[int(a), int(b)] for a, b in zip(s, d)]
    s, d = O.update_rmat(1 << 24, 16, first=1000)
    out["update_rmat_n16777216_from1000_16"] = [[int(a), int(b)] for a, b in zip(s, d)]

    s, d = O.rmat_edges(16, 64, seed=1)
    out["rmat_s16_seed1_first64"] = [[int(a), int(b)] for a, b in zip(s, d)]
    g = O.rmat_graph(16)
    par, dirlog, fs, fd = O.bfs_diropt(g, 10)
    lv = O.bfs_levels(g, 10)
    reached = lv != O.NONE
    p, l1 = O.pagerank(g, 0.85, 20, 0.0)
    out["rmat_s16"] = {
        "n": g.n, "m": g.m, "map_sum": O.map_sum(g), "triangles": O.tc(g),
        "components": int(len(np.unique(O.cc(g)))),
        "bfs10_level_sizes": [int(x) for x in np.bincount(lv[reached])],
        "bfs10_dirlog": [int(x) for x in dirlog],
        "bfs10_fdeg": [int(x) for x in fd],
        "sources8": O.bfs_sources(g, 8),
        "pr20_l1_log": [float(x) for x in l1],
        "pr20_sum": float(p.sum()),
        "pr20_first16": [float(x) for x in p[:16]],
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
