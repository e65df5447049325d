"""Learned edgeMap direction selector: data collection, training and
evaluation.

The selector replaces GBBS's fixed M/20 rule (gbbs/edge_map_data.h:257-306)
as the per-level choice between sparse push (class 0), dense pull (class 1)
and a third choice (class 2): for the shared multi-source traversal the
bitmap pull (the frontier's representation: each in-neighbour is tested
against a 1-bit union frontier before its per-source mask is read), for the
one-source traversal dense forward (edgeMapDenseForward; its pull already
tests a bitmap).  Labels come from timings the engine itself measures: for
every level of a BFS run under the rule, the run is repeated with each
direction forced at that level (byo_set_direction_schedule) and the label is
the direction with the least device time from that level to the end.
Features are exactly the ones the device evaluates (bfs_coop.hip, decide):

    log2(1+|F|), log2(1+sum deg F), log2(1+unvisited edges), level,
    log2(1+avg degree), log2(1+max degree)

Training runs on the GPU through the library's own forward/backward/Adam
kernels (byo_trainer_*).  Direction never changes BFS levels, so the learned
policy is bit-exact whatever it predicts; the rule stays the fallback when no
weights are loaded.

Each traversal engine has its own network, trained on its own timings: the
one-source traversal (byo_bfs) and the shared multi-source traversal of
byo_bfs_multi, whose union frontier makes the pull scan far longer (a vertex
stops only once every active source is found) and the push cheaper per
source, so the same frontier statistics call for different directions.

    python -m byo_amd.selector train --engine batch
    python -m byo_amd.selector eval  --engine batch
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import (DIR_MLP, DIR_RULE, DIR_THRESHOLD, MLP_NPARAM, SEL_BATCH, SEL_SINGLE, CsrGraph, Trainer)

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "selector_weights.json")
WEIGHTS_BATCH = os.path.join(HERE, "selector_weights_batch.json")
ENGINES = {"single": (SEL_SINGLE, WEIGHTS), "batch": (SEL_BATCH, WEIGHTS_BATCH)}
# tuning knob holding each engine's decision margin (logits x 100)
MARGIN_KNOB = {"single": "sel_margin_x100", "batch": "sel_margin_batch_x100"}
BATCH = 8  # sources per shared traversal when collecting and evaluating
WIDE = 64  # the batched job's sources per shared traversal
# direction of each selector class, per engine (DirCtl::third on the device)
CLASS_TO_DIR = {"single": (0, 1, 3), "batch": (0, 1, 4)}


def dir_to_class(d):
    return 2 if d in (3, 4) else (1 if d == 1 else 0)


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


def _run(G, srcs, engine):
    """One traversal: a single source (engine "single") or one shared
    traversal of a batch of sources ("batch"); returns the per-level record
    (for a batch, the union frontier's)."""
    if engine == "single":
        return G.bfs(srcs, want_parents=False)[1]
    return G.bfs_multi(srcs, want_parents=False)[1][0]


def collect(G, units, repeats=1, max_levels=64, engine="single"):
    """Per-level samples (features, times of push / pull / forward) of
    traversals over units (sources, or batches of sources for the batched
    engine).  Each level is re-run with every direction of the engine's
    classes forced."""
    n, m = G.N(), G.M()
    deg = G.degrees()
    maxdeg = int(deg.max()) if n else 0
    X, T, R, Z = [], [], [], []
    for u in units:
        G.set_direction_policy(DIR_THRESHOLD)
        st = _run(G, u, engine)
        L = min(max(st["levels"], int(np.count_nonzero(st["fsize"]))), max_levels)
        total = float(np.sum(st["level_ms"][:min(max(st["levels"], int(np.count_nonzero(st["fsize"]))), 64)]))
        vdeg = 0
        for lv in range(L):
            if st["fsize"][lv] == 0 or st["dir"][lv] == 2:
                break
            times = []
            for d in CLASS_TO_DIR[engine]:
                sched = [DIR_RULE] * lv + [d]
                G.set_direction_schedule(sched)
                # the label is the time from this level to the end of the
                # traversal: a direction also changes what the next level
                # starts from (a pull leaves no queue, so a push after it
                # packs one first)
                best = float("inf")
                for _ in range(repeats):
                    s2 = _run(G, u, engine)
                    end = min(max(s2["levels"], int(np.count_nonzero(s2["fsize"]))), 64)
                    best = min(best, float(np.sum(s2["level_ms"][lv:end])))
                times.append(best)
            X.append(features(st["fsize"][lv], st["fdeg"][lv], vdeg, lv, n, m, maxdeg))
            T.append(times)
            R.append(dir_to_class(st["dir"][lv]))
            Z.append(total)
            vdeg += st["fdeg"][lv]
    G.set_direction_policy(DIR_THRESHOLD)
    return np.array(X, np.float32), np.array(T, np.float64), np.array(R, np.int32), np.array(Z, np.float64)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_NPARAM, np.float32)
    p[0:192] = rng.standard_normal(192) * math.sqrt(2 / 6)
    p[224:1248] = rng.standard_normal(1024) * math.sqrt(2 / 32)
    p[1280:1376] = rng.standard_normal(96) * math.sqrt(1 / 32)
    return p


CLAMP = 3.0  # normalised features are clamped to +-CLAMP (kSelClamp on the device)


def normalise(X, mean, inv_std):
    return np.clip((X - mean) * inv_std, -CLAMP, CLAMP).astype(np.float32)


RULE_EDGE = 0.05  # another direction must beat the rule's by 5% to become the label
MIN_SHARE = 0.04  # ... and save at least 4% of the whole traversal's time
DENSIFY = False   # the device's sel_densify default: sparse levels of the rule stay sparse


def labels(T, R, Z=None):
    """Class with the least time from the level to the end of the traversal,
    where the rule's own choice (R) wins unless another direction beats it
    by RULE_EDGE and by MIN_SHARE of the traversal's total time Z under the
    rule: small levels time within noise of each other, and the network
    should learn to leave the rule only where that pays."""
    Ta = T.copy()
    if R is None:
        return np.argmin(Ta, axis=1).astype(np.int32)
    idx = np.arange(len(R))
    Ta[idx, R] *= (1.0 - RULE_EDGE)
    y = np.argmin(Ta, axis=1).astype(np.int32)
    if Z is not None:
        small = (T[idx, R] - T[idx, y]) < MIN_SHARE * Z
        y[small] = R[small]
    if not DENSIFY:
        # the device keeps the rule's sparse levels sparse (knob sel_densify)
        keep = (R == 0) & (y != 0)
        y[keep] = 0
    return y


def train(X, T, R=None, epochs=3000, lr=3e-3, seed=0, Z=None):
    """Cost-sensitive training: every sample is repeated in proportion to how
    much the wrong choice would cost (relative gap between worst and best
    direction), then fit with softmax cross-entropy on the device."""
    y = labels(T, R, Z)
    gap = (T.max(axis=1) - T.min(axis=1)) / np.maximum(T.min(axis=1), 1e-6)
    reps = np.clip(np.round(1 + 3 * np.minimum(gap, 3.0)), 1, 10).astype(int)
    Xr = np.repeat(X, reps, axis=0)
    yr = np.repeat(y, reps)
    mean = Xr.mean(axis=0)
    std = Xr.std(axis=0)
    inv_std = (1.0 / np.where(std > 1e-6, std, 1.0)).astype(np.float32)
    Xn = normalise(Xr, mean, inv_std)
    tr = Trainer(init_params(seed), lr=lr)
    losses = []
    for e in range(epochs):
        losses.append(tr.step(Xn, yr))
    params = tr.params()
    pred = tr.predict(normalise(X, mean, inv_std))
    tr.close()
    acc = float((pred == y).mean())
    chosen = T[np.arange(len(y)), pred].sum()
    return {"params": params.tolist(), "mean": mean.astype(np.float32).tolist(),
            "inv_std": inv_std.tolist(), "train_accuracy": acc,
            "train_time_ratio_vs_best": float(chosen / T.min(axis=1).sum()),
            "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "samples": int(len(y))}


def load(G, path=None, engine=None):
    """Loads the committed weights of every engine (or one file for one
    engine) into G; returns the weight dicts by engine name."""
    out = {}
    for name, (eid, default) in ENGINES.items():
        if engine is not None and name != engine:
            continue
        p = path if path is not None else default
        if not os.path.exists(p):
            continue
        with open(p) as f:
            w = json.load(f)
        G.load_selector(np.array(w["params"], np.float32), np.array(w["mean"], np.float32),
                        np.array(w["inv_std"], np.float32), engine=eid)
        if "margin" in w:
            G.set_tuning(MARGIN_KNOB[name], int(round(100 * float(w["margin"]))))
        out[name] = w
    return out


def _timed(G, srcs, engine):
    """Device ms of the traversals of srcs in one call: the batched engine
    (bfs_multi, shared traversals of up to 64 sources) or the one-source
    engine (every source in turn, still one persistent launch)."""
    G.set_tuning("bfs_engine", 1 if engine == "single" else 0)
    try:
        G.profile(True)
        G.profile_reset()
        _, st = G.bfs_multi(srcs, want_parents=False)
        _, ms = G.profile_get("bfs_launch")
    finally:
        G.set_tuning("bfs_engine", 0)
    return ms, st


def compare(G, srcs, path=None, repeats=5, engine="batch"):
    """Device time of the traversals of srcs under the rule and the learned
    policy (runs alternate between the two, best of `repeats` each), and
    whether their levels agree (levels of every source are derived from its
    parent tree on the device)."""
    load(G, path, engine)
    res = {"engine": engine}
    pols = (("threshold", DIR_THRESHOLD), ("learned", DIR_MLP))
    best = {name: float("inf") for name, _ in pols}
    last = {}
    for r in range(repeats + 1):
        for name, pol in pols:
            G.set_direction_policy(pol)
            ms, st = _timed(G, srcs, engine)
            if r:  # the first round warms both
                best[name] = min(best[name], ms)
            last[name] = st
    lv = {}
    for name, pol in pols:
        G.set_direction_policy(pol)
        st = last[name]
        res[name] = best[name]
        res[name + "_dirs"] = st[0]["dir"][:max(s["levels"] for s in st)]
        lv[name] = []
        for s in srcs[:2]:
            G.bfs(s, want_parents=False)
            lv[name].append(G.bfs_levels())
        lv[name].append([(x["levels"], x["reached"], x["edges_reached"]) for x in st])
    G.set_direction_policy(DIR_THRESHOLD)
    res["speedup"] = res["threshold"] / res["learned"]
    res["same_directions"] = res["threshold_dirs"] == res["learned_dirs"]
    res["levels_identical"] = all(np.array_equal(np.asarray(a), np.asarray(b))
                                  for a, b in zip(lv["threshold"], lv["learned"]))
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
    ("rmat", 22, 8, (0.25, 0.25, 0.25)), ("rmat", 22, 32, (0.5, 0.1, 0.1)),
]


def train_units(G, eng, scale, nunits):
    """Training traversals of one graph: single sources for the one-source
    engine; for the shared traversal, batches of BATCH sources (the
    benchmark's step) and of WIDE sources (the batched job), whose union
    frontiers put the same level at very different sizes."""
    deg = G.degrees()
    if eng == "single":
        return sources(deg, nunits, seed=100 + scale)
    srcs = sources(deg, nunits * BATCH + (nunits // 2) * WIDE, seed=100 + scale)
    units = [srcs[i * BATCH:(i + 1) * BATCH] for i in range(nunits)]
    base = nunits * BATCH
    units += [srcs[base + i * WIDE: base + (i + 1) * WIDE] for i in range(nunits // 2)]
    return units


HELD_OUT = (("rmat22_ef16", lambda: CsrGraph.rmat(22, 16)),
            ("er23_deg8", lambda: CsrGraph.erdos_renyi(23, 4 << 23)),
            ("grid256", lambda: CsrGraph.grid3d(256)),
            ("rmat24_ef32", lambda: CsrGraph.rmat(24, 32)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--engine", choices=sorted(ENGINES), default="batch")
    ap.add_argument("--out", default=None)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--sources", type=int, default=4, help="training units per graph")
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--dump", default=None, help="also write the samples (X, T, R, Z) to this .npz")
    ap.add_argument("--samples", default=None, help="train from the samples of an earlier --dump")
    ap.add_argument("--min-share", type=float, default=None, help="MIN_SHARE for the labels")
    args = ap.parse_args()
    eng = args.engine
    global MIN_SHARE
    if args.min_share is not None:
        MIN_SHARE = args.min_share
    if args.cmd == "train":
        out = args.out or ENGINES[eng][1]
        Xs, Ts, Rs, Zs = [], [], [], []
        t0 = time.time()
        for kind, scale, ef, (a, b, c) in ([] if args.samples else TRAIN_GRAPHS):
            G = CsrGraph.rmat(scale, ef, a, b, c, seed=7 + scale)
            units = train_units(G, eng, scale, args.sources)
            X, T, R, Z = collect(G, units, repeats=3, engine=eng)
            Xs.append(X)
            Ts.append(T)
            Rs.append(R)
            Zs.append(Z)
            G.close()
            print("collected %s-%d ef%d: %d samples (%.0fs)" % (kind, scale, ef, len(X), time.time() - t0),
                  flush=True)
        if args.samples:
            d = np.load(args.samples)
            X, T, R, Z = d["X"], d["T"], d["R"], d["Z"]
        else:
            X = np.concatenate(Xs)
            T = np.concatenate(Ts)
            R = np.concatenate(Rs)
            Z = np.concatenate(Zs)
        if args.dump:
            np.savez(args.dump, X=X, T=T, R=R, Z=Z)
        w = train(X, T, R, epochs=args.epochs, Z=Z)
        w["engine"] = eng
        w["classes_to_dirs"] = list(CLASS_TO_DIR[eng])
        w["train_graphs"] = [list(map(str, g)) for g in TRAIN_GRAPHS]
        w["label_counts"] = np.bincount(labels(T, R, Z), minlength=3).tolist()
        w["min_share"] = MIN_SHARE
        # level time the rule's own choices cost against the per-level best
        w["rule_time_ratio_vs_best"] = float(T[np.arange(len(R)), R].sum() / T.min(axis=1).sum())
        with open(out, "w") as f:
            json.dump(w, f)
        print(json.dumps({k: v for k, v in w.items() if k not in ("params", "mean", "inv_std")}), flush=True)
    else:
        # held-out graphs: RMAT scale 24 ef 32, 256^3 grid, ER n=2^23 degree
        # 8, and the RMAT scale 22 ef 16 graph; the shared traversal at both
        # of its widths
        widths = (BATCH, WIDE) if eng == "batch" else (4,)
        for name, make in HELD_OUT:
            G = make()
            for k in widths:
                srcs = sources(G.degrees(), k, seed=5)
                res = compare(G, srcs, args.weights, engine=eng)
                res["n"], res["m"], res["sources"] = G.N(), G.M(), k
                print(name, json.dumps(res), flush=True)
            G.close()


if __name__ == "__main__":
    sys.exit(main())
