"""Learned pivot selection: training data from FB runs, GPU training, evaluation.

The scorer is an MLP 12 -> 32 -> 32 -> 1 (BF16 weights and activations, FP32
accumulation) whose forward, backward and Adam steps are HIP kernels in
libgpuscc.so (kernels_mlp.hip); this module only drives them.

Targets come from the method itself: a plain FB run gives the SCC partition;
the run is then replayed election by election through the phase-level entry
points, and at each election a live vertex is a positive example when its SCC
is a largest SCC of its colour set (colour sets whose SCCs are all trivial carry
no information and are skipped).  With the learned rule the highest-scoring
vertex of each colour slot becomes the pivot; the partition does not depend on
the choice, only the number of FB steps does.

    python -m gpuscc_code_amd.learned train --graphs 512 --out gpuscc_code_amd/pivot_mlp.npy
    python -m gpuscc_code_amd.learned evaluate --params gpuscc_code_amd/pivot_mlp.npy
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import scc

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS_PATH = os.path.join(HERE, "pivot_mlp.npy")
NF, NH = scc.MLP_FEATURES, scc.MLP_HIDDEN


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def init_params(seed=0):
    """He initialisation; layout W1[32][12] b1 W2[32][32] b2 w3 b3."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / NF), NH * NF), np.zeros(NH),
             rng.normal(0, np.sqrt(2.0 / NH), NH * NH), np.zeros(NH),
             rng.normal(0, np.sqrt(1.0 / NH), NH), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def load_params(path=PARAMS_PATH):
    return np.load(path, allow_pickle=False).astype(np.float32)


def largest_scc_targets(colour, label):
    """For live vertices with colours and SCC labels: (is in a largest SCC of its
    colour set, colour set has a non-trivial SCC)."""
    colour = np.asarray(colour, np.int64)
    label = np.asarray(label, np.int64)
    key = (colour << 32) | label
    _, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    size = cnt[inv]
    _, cinv = np.unique(colour, return_inverse=True)
    best = np.zeros(cinv.max() + 1 if len(cinv) else 0, np.int64)
    np.maximum.at(best, cinv, size)
    b = best[cinv]
    return size == b, b > 1


def collect(ctx, g, rng, per_election=128, max_elections=48, schedule="h"):
    """Replays a plain FB run of g and returns (X, y) samples of its elections."""
    ctx.upload(g)
    truth, _ = ctx.run(schedule)
    n = g.n
    ctx.set_state(scc.make_word(np.ones(n, np.uint32), np.zeros(n, np.uint32)),
                  np.arange(n, dtype=np.int32))
    ctx.trim()
    Xs, ys = [], []
    step = 0
    while step < max_elections:
        words, _ = ctx.get_state()
        live = np.nonzero((words & 4) == 0)[0]
        if len(live) == 0:
            break
        y, informative = largest_scc_targets(words[live] >> 8, truth[live])
        cand, yc = live[informative], y[informative]
        if len(cand):
            pos, neg = np.nonzero(yc)[0], np.nonzero(~yc)[0]
            k = per_election // 2
            sel = np.concatenate([rng.choice(pos, min(k, len(pos)), replace=False),
                                  rng.choice(neg, min(k, len(neg)), replace=False)])
            Xs.append(ctx.features(cand[sel], step))
            ys.append(yc[sel].astype(np.float32))
        if ctx.update(recolour=False) == 0:
            break
        ctx.reach(2, schedule)
        ctx.trim()
        ctx.recolour()
        step += 1
    return Xs, ys


def pivot_hit_rates(ctx, g, params, schedule="h", max_elections=48):
    """Replays a plain FB run of g; at every election and for every colour set
    with a non-trivial SCC, checks whether the plain pick (lowest live id) and
    the scorer's pick (highest score, ties to the lowest id) lie in a largest
    SCC of the set.  Returns (sets, plain hits, learned hits)."""
    ctx.upload(g)
    truth, _ = ctx.run(schedule)
    n = g.n
    ctx.set_state(scc.make_word(np.ones(n, np.uint32), np.zeros(n, np.uint32)),
                  np.arange(n, dtype=np.int32))
    ctx.trim()
    ctx.mlp_set(params)
    sets = hits_plain = hits_learned = 0
    step = 0
    while step < max_elections:
        words, _ = ctx.get_state()
        live = np.nonzero((words & 4) == 0)[0]
        if len(live) == 0:
            break
        colour = (words[live] >> 8).astype(np.int64)
        y, informative = largest_scc_targets(colour, truth[live])
        logits = ctx.mlp_infer(ctx.features(live, step))
        order = np.lexsort((live, -logits, colour))           # per colour: best score, then id
        first = np.ones(len(order), bool)
        first[1:] = colour[order][1:] != colour[order][:-1]
        best = order[first]
        lowest = np.lexsort((live, colour))
        lfirst = np.ones(len(lowest), bool)
        lfirst[1:] = colour[lowest][1:] != colour[lowest][:-1]
        low = lowest[lfirst]
        keep = informative[best]
        sets += int(keep.sum())
        hits_learned += int(y[best][keep].sum())
        hits_plain += int(y[low][informative[low]].sum())
        if ctx.update(recolour=False) == 0:
            break
        ctx.reach(2, schedule)
        ctx.trim()
        ctx.recolour()
        step += 1
    return sets, hits_plain, hits_learned


def training_graphs(count, seed):
    """R-MAT scale 16-20, grids 256-1024 per side, planted-cycle ER graphs
    (sizes drawn towards the small end)."""
    rng = np.random.default_rng(seed)
    for i in range(count):
        kind = i % 3
        u = rng.random() ** 2
        s = int(rng.integers(1, 1 << 30))
        if kind == 0:
            yield f"rmat{16 + int(u * 5)}", scc.gen_rmat(16 + int(u * 5), 8, seed=s)
        elif kind == 1:
            side = int(256 * 4 ** u)
            yield f"grid{side}", scc.gen_mesh(side, 0.6, seed=s)
        else:
            nv = int(np.exp(np.log(1e4) + u * (np.log(1e6) - np.log(1e4))))
            deg = float(rng.uniform(1.5, 6.0))
            yield f"er{nv}", scc.gen_batch(1, nv, nv, deg, deg, 0.05, 0.5, seed=s)


def train(n_graphs=512, epochs=12, batch=8192, lr=3e-3, seed=1, per_election=128,
          max_elections=48):
    rng = np.random.default_rng(seed)
    X_all, y_all = [], []
    t0 = time.time()
    with scc.Context(0) as ctx:
        for i, (name, g) in enumerate(training_graphs(n_graphs, seed)):
            Xs, ys = collect(ctx, g, rng, per_election, max_elections)
            X_all += Xs
            y_all += ys
            if i % 16 == 0:
                log(f"[{i}/{n_graphs}] {name} n={g.n} samples={sum(len(y) for y in y_all)} "
                    f"{time.time() - t0:.0f}s")
        X = np.concatenate(X_all).astype(np.float32)
        y = np.concatenate(y_all).astype(np.float32)
        log(f"training set: {len(y)} samples, {y.mean():.3f} positive")
        ctx.mlp_set(init_params(seed))
        losses = ctx.mlp_train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
        log("epoch losses: " + " ".join(f"{x:.4f}" for x in losses))
        params = ctx.mlp_get()
    return params, dict(samples=int(len(y)), positive=float(y.mean()),
                        losses=[float(x) for x in losses])


def evaluate(params, graphs, schedules=("h",), variants=(0,), reps=3):
    """Plain vs learned pivots: FB steps, milliseconds (median), identical labels."""
    rows = []
    with scc.Context(0) as ctx:
        ctx.mlp_set(params)
        for name, g in graphs:
            ctx.upload(g)
            ctx.set_batch(g.graph_offsets if name.startswith("batch") else None)
            for sched in schedules:
                for var in variants:
                    res = {}
                    for rule, pv in (("plain", scc.PIVOT_PLAIN), ("learned", scc.PIVOT_LEARNED)):
                        lab, st = ctx.run(sched, var, pivot=pv)
                        ms = []
                        for _ in range(reps):
                            _, s2 = ctx.run(sched, var, pivot=pv, labels=False)
                            ms.append(s2["ms_total"])
                        res[rule] = dict(steps=st["steps"], ms=float(np.median(ms)), labels=lab,
                                         pivots=st["pivots_total"])
                    same = bool(np.array_equal(res["plain"]["labels"], res["learned"]["labels"]))
                    row = dict(graph=name, n=g.n, m=g.m, schedule=sched, variant=var,
                               plain_steps=res["plain"]["steps"], learned_steps=res["learned"]["steps"],
                               plain_ms=round(res["plain"]["ms"], 3),
                               learned_ms=round(res["learned"]["ms"], 3),
                               speedup=round(res["plain"]["ms"] / res["learned"]["ms"], 3),
                               identical_labels=same)
                    log(json.dumps(row))
                    rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--graphs", type=int, default=512)
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--out", default=PARAMS_PATH)
    t.add_argument("--seed", type=int, default=1)
    e = sub.add_parser("evaluate")
    e.add_argument("--params", default=PARAMS_PATH)
    e.add_argument("--s24", action="store_true", help="include R-MAT scale 24, edge factor 16, seed 5")
    e.add_argument("--batch", action="store_true", help="include the batch of 4,096 graphs (seed 4)")
    e.add_argument("--variants", nargs="+", type=int, default=[0, 1])
    args = ap.parse_args()
    if args.cmd == "train":
        params, info = train(n_graphs=args.graphs, epochs=args.epochs, seed=args.seed)
        np.save(args.out, params)
        print(json.dumps(info))
    else:
        params = load_params(args.params)
        graphs = [("mesh2048", scc.gen_mesh(2048, 0.6, seed=3)),
                  ("rmat22", scc.gen_rmat(22, 8, seed=2))]
        if args.s24:
            graphs.append(("rmat24ef16", scc.gen_rmat(24, 16, seed=5)))
        if args.batch:
            graphs.append(("batch4096", scc.gen_batch(4096, 64, 65536, 1.5, 6.0, 0.05, 0.5, seed=4)))
        rows = evaluate(params, graphs, schedules=("h", "d"), variants=tuple(args.variants))
        print(json.dumps(rows))


if __name__ == "__main__":
    main()
