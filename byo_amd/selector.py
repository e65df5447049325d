"""Learned edgeMap direction selector: data collection, training and
evaluation.

The selector replaces GBBS's fixed M/20 rule (gbbs/edge_map_data.h:257-306)
as the per-level choice between sparse push (class 0), dense pull (class 1)
and dense forward (class 2, edgeMapDenseForward).  Labels come from timings
the engine itself measures: for every level of a BFS run under the rule, the
level is re-run with each direction forced (byo_set_direction_schedule) and
the direction with the smallest device time of that level is the label.
Features are exactly the ones the device evaluates (bfs_coop.hip, decide):

    log2(1+|F|), log2(1+sum deg F), log2(1+unvisited edges), level,
    log2(1+avg degree), log2(1+max degree)

Training runs on the GPU through the library's own forward/backward/Adam
kernels (byo_trainer_*).  Direction never changes BFS levels, so the learned
policy is bit-exact whatever it predicts; the rule stays the fallback when no
weights are loaded.

    python -m byo_amd.selector train --out byo_amd/selector_weights.json
    python -m byo_amd.selector eval  --weights byo_amd/selector_weights.json
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import (DIR_MLP, DIR_RULE, DIR_THRESHOLD, MLP_NPARAM, CsrGraph, Trainer)

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "selector_weights.json")
CLASS_TO_DIR = (0, 1, 3)
DIR_TO_CLASS = {0: 0, 1: 1, 3: 2}


def features(fs, fdeg, visited_deg, level, n, m, maxdeg):
    avg = m / n if n else 0.0
    return [math.log2(1.0 + fs), math.log2(1.0 + fdeg), math.log2(1.0 + max(m - visited_deg, 0)),
            float(level), math.log2(1.0 + avg), math.log2(1.0 + maxdeg)]


def splitmix64(x):
    M = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def sources(deg, k, seed):
    out, x, n = [], seed, len(deg)
    while len(out) < k:
        x = splitmix64(x)
        v = x % n
        if deg[v] > 0 and v not in out:
            out.append(int(v))
    return out


def collect(G, srcs, repeats=1, max_levels=64):
    """Per-level samples (features, times of push / pull / forward) of BFS
    runs from srcs.  Each level is re-run with every direction forced."""
    n, m = G.N(), G.M()
    deg = G.degrees()
    maxdeg = int(deg.max()) if n else 0
    X, T = [], []
    for s in srcs:
        G.set_direction_policy(DIR_THRESHOLD)
        _, st = G.bfs(s, want_parents=False)
        L = min(st["levels"], max_levels)
        vdeg = 0
        for lv in range(L):
            if st["dir"][lv] == 2:
                break
            times = []
            for d in CLASS_TO_DIR:
                sched = [DIR_RULE] * lv + [d]
                G.set_direction_schedule(sched)
                best = float("inf")
                for _ in range(repeats):
                    _, s2 = G.bfs(s, want_parents=False)
                    best = min(best, s2["level_ms"][lv])
                times.append(best)
            X.append(features(st["fsize"][lv], st["fdeg"][lv], vdeg, lv, n, m, maxdeg))
            T.append(times)
            vdeg += st["fdeg"][lv]
    G.set_direction_policy(DIR_THRESHOLD)
    return np.array(X, np.float32), np.array(T, np.float64)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_NPARAM, np.float32)
    p[0:192] = rng.standard_normal(192) * math.sqrt(2 / 6)
    p[224:1248] = rng.standard_normal(1024) * math.sqrt(2 / 32)
    p[1280:1376] = rng.standard_normal(96) * math.sqrt(1 / 32)
    return p


def train(X, T, epochs=3000, lr=3e-3, seed=0):
    """Cost-sensitive training: every sample is repeated in proportion to how
    much the wrong choice would cost (relative gap between worst and best
    direction), then fit with softmax cross-entropy on the device."""
    y = np.argmin(T, axis=1).astype(np.int32)
    gap = (T.max(axis=1) - T.min(axis=1)) / np.maximum(T.min(axis=1), 1e-6)
    reps = np.clip(np.round(1 + 3 * np.minimum(gap, 3.0)), 1, 10).astype(int)
    Xr = np.repeat(X, reps, axis=0)
    yr = np.repeat(y, reps)
    mean = Xr.mean(axis=0)
    std = Xr.std(axis=0)
    inv_std = (1.0 / np.where(std > 1e-6, std, 1.0)).astype(np.float32)
    Xn = ((Xr - mean) * inv_std).astype(np.float32)
    tr = Trainer(init_params(seed), lr=lr)
    losses = []
    for e in range(epochs):
        losses.append(tr.step(Xn, yr))
    params = tr.params()
    pred = tr.predict(((X - mean) * inv_std).astype(np.float32))
    tr.close()
    acc = float((pred == y).mean())
    chosen = T[np.arange(len(y)), pred].sum()
    return {"params": params.tolist(), "mean": mean.astype(np.float32).tolist(),
            "inv_std": inv_std.tolist(), "train_accuracy": acc,
            "train_time_ratio_vs_best": float(chosen / T.min(axis=1).sum()),
            "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "samples": int(len(y))}


def load(G, path=WEIGHTS):
    with open(path) as f:
        w = json.load(f)
    G.load_selector(np.array(w["params"], np.float32), np.array(w["mean"], np.float32),
                    np.array(w["inv_std"], np.float32))
    return w


def compare(G, srcs, path=WEIGHTS, repeats=3):
    """Device time of k BFS (one launch) under the rule and the learned
    policy, and whether their levels agree."""
    load(G, path)
    res = {}
    lv = {}
    for name, pol in (("threshold", DIR_THRESHOLD), ("learned", DIR_MLP)):
        G.set_direction_policy(pol)
        G.bfs_multi(srcs, want_parents=False)
        best = float("inf")
        for _ in range(repeats):
            G.profile(True)
            G.profile_reset()
            _, st = G.bfs_multi(srcs, want_parents=False)
            _, ms = G.profile_get("bfs_launch")
            best = min(best, ms)
        res[name] = best
        res[name + "_dirs"] = st[0]["dir"]
        lv[name] = []
        for s in srcs[:2]:
            G.bfs(s, want_parents=False)
            lv[name].append(G.bfs_levels())
    G.set_direction_policy(DIR_THRESHOLD)
    res["speedup"] = res["threshold"] / res["learned"]
    res["levels_identical"] = all(np.array_equal(a, b) for a, b in zip(lv["threshold"], lv["learned"]))
    return res


# RMAT scale 18-22 over the skew range: (0.25, 0.25, 0.25) is the uniform
# (Erdos-Renyi-like) corner, (0.57, 0.19, 0.19) the Graph500 one.
TRAIN_GRAPHS = [
    ("rmat", 18, 16, (0.5, 0.1, 0.1)), ("rmat", 18, 4, (0.25, 0.25, 0.25)),
    ("rmat", 19, 16, (0.57, 0.19, 0.19)), ("rmat", 19, 8, (0.25, 0.25, 0.25)),
    ("rmat", 20, 16, (0.5, 0.1, 0.1)), ("rmat", 20, 8, (0.45, 0.15, 0.15)),
    ("rmat", 20, 4, (0.35, 0.2, 0.2)), ("rmat", 21, 16, (0.57, 0.19, 0.19)),
    ("rmat", 21, 32, (0.5, 0.1, 0.1)), ("rmat", 21, 8, (0.25, 0.25, 0.25)),
    ("rmat", 22, 16, (0.5, 0.1, 0.1)), ("rmat", 22, 4, (0.25, 0.25, 0.25)),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--sources", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=3000)
    args = ap.parse_args()
    if args.cmd == "train":
        Xs, Ts = [], []
        t0 = time.time()
        for kind, scale, ef, (a, b, c) in TRAIN_GRAPHS:
            G = CsrGraph.rmat(scale, ef, a, b, c, seed=7 + scale)
            srcs = sources(G.degrees(), args.sources, seed=100 + scale)
            X, T = collect(G, srcs, repeats=2)
            Xs.append(X)
            Ts.append(T)
            G.close()
            print("collected %s-%d ef%d: %d samples (%.0fs)" % (kind, scale, ef, len(X), time.time() - t0),
                  flush=True)
        X = np.concatenate(Xs)
        T = np.concatenate(Ts)
        w = train(X, T, epochs=args.epochs)
        w["train_graphs"] = [list(map(str, g)) for g in TRAIN_GRAPHS]
        w["label_counts"] = np.bincount(np.argmin(T, axis=1), minlength=3).tolist()
        with open(args.out, "w") as f:
            json.dump(w, f)
        print(json.dumps({k: v for k, v in w.items() if k not in ("params", "mean", "inv_std")}), flush=True)
    else:
        # held-out graphs: RMAT scale 24 ef 32, 256^3 grid, ER n=2^23 degree 8,
        # and the bench graph (RMAT scale 22 ef 16)
        for name, make in (("rmat22_ef16", lambda: CsrGraph.rmat(22, 16)),
                           ("er23_deg8", lambda: CsrGraph.erdos_renyi(23, 4 << 23)),
                           ("grid256", lambda: CsrGraph.grid3d(256)),
                           ("rmat24_ef32", lambda: CsrGraph.rmat(24, 32))):
            G = make()
            srcs = sources(G.degrees(), 4, seed=5)
            res = compare(G, srcs, args.weights)
            res["n"], res["m"] = G.N(), G.M()
            print(name, json.dumps(res), flush=True)
            G.close()


if __name__ == "__main__":
    sys.exit(main())
