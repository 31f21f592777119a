"""Autotuner for the run-time options of the FB engine (scc_opts).

The reference fixes its launch shape and switch points at compile time
(256-work-item blocks, the 1 % switch and alpha/beta of scc.cu:493-563 and
reach_b.cu:540,645).  Here they are tuned per GPU and per graph class:
coordinate descent over the knobs below, each trial being the median wall time
of `reps` uninstrumented decompositions of an already uploaded graph
(scc_set_opts changes the options between runs, so no trial re-uploads).  A
value is kept only when it is faster than the best so far by more than
`min_gain`, and the kept configuration must reproduce the labels of the
default configuration exactly.  Results are cached per gfx target and graph
class in tuned.json, which opts_for() reads (bench.py uses it by default).

    python -m gpuscc_code_amd.autotune --workloads rmat22 mesh2048 batch4096 rmat24
"""
import argparse
import dataclasses
import json
import os
import statistics
import sys
import time

import numpy as np

from . import scc

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED_PATH = os.path.join(HERE, "tuned.json")

# Knob -> candidate values given (n, compute units); 0 is the library default.
KNOBS = {
    "grid_blocks": lambda n, cus: [0, 2 * cus, 6 * cus, 8 * cus],
    "hybrid_arc_div": lambda n, cus: [0, 6, 30, -1],
    "hybrid_threshold": lambda n, cus: [0, max(1024, n // 10), max(1024, n // 40), max(1024, n // 80)],
    "local_rounds": lambda n, cus: [0, 16, 256, -1],
    "huge_row": lambda n, cus: [0, 512, 1024, 2048, 16384],
    "block_size": lambda n, cus: [0, 128, 512],
    "reach_graphs": lambda n, cus: [0, 1],
    "reach_block": lambda n, cus: [0, 512],
    "hybrid_policy": lambda n, cus: [0, 2],
    "tail_live": lambda n, cus: [0, 8192, -1],
}
ORDER = ["hybrid_policy", "grid_blocks", "reach_block", "hybrid_arc_div", "hybrid_threshold", "local_rounds",
         "huge_row", "block_size", "reach_graphs", "tail_live"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def graph_class(g, batch=False):
    """Shape and size class of a graph: 'batch', 'skewed' (max degree above
    64x the mean, R-MAT-like) or 'regular' (meshes), with m rounded to a power of two."""
    if batch:
        shape = "batch"
    else:
        deg = np.diff(np.asarray(g.row, np.int64))
        mean = max(g.m / max(g.n, 1), 1.0)
        shape = "skewed" if len(deg) and deg.max() > 64 * mean else "regular"
    return f"{shape}-m2^{int(round(np.log2(max(g.m, 1))))}"


def class_key(cls, schedule, variant):
    return f"{cls}/{schedule}/{variant}"


def load_table(path=TUNED_PATH):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def opts_for(g, arch, schedule="h", variant=0, batch=False, path=TUNED_PATH, base=None):
    """Tuned options for g: the entry of its class, else the nearest size of the
    same shape (within a factor 4 in m), else the defaults.  Returns
    (alg_opt, entry key or None)."""
    base = base or scc.alg_opt()
    table = load_table(path).get(arch, {})
    cls = graph_class(g, batch)
    shape, size = cls.split("-m2^")
    best, best_d = None, None
    for key, ent in table.items():
        c, sch, var = key.split("/")
        if sch != schedule or int(var) != variant:
            continue
        s2, z2 = c.split("-m2^")
        d = abs(int(z2) - int(size))
        if s2 == shape and d <= 2 and (best_d is None or d < best_d):
            best, best_d = key, d
    if best is None:
        return base, None
    return dataclasses.replace(base, **table[best]["opts"]), best


def _median_ms(ctx, schedule, variant, reps):
    ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.run(schedule, variant, labels=False)
        ms.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(ms)


def tune(ctx, g, schedule="h", variant=0, reps=5, rounds=2, min_gain=0.02, knobs=None, base=None):
    """Coordinate descent on an uploaded graph g (and batch, if set on ctx).
    Returns dict(opts=..., ms=..., default_ms=..., trials=...)."""
    arch, cus = ctx.device_info()
    base = base or scc.alg_opt()
    knobs = knobs or ORDER
    ctx.set_opts(base)
    ref, _ = ctx.run(schedule, variant)
    _median_ms(ctx, schedule, variant, 1)  # warm-up
    best_ms = default_ms = _median_ms(ctx, schedule, variant, reps)
    best = {}
    trials = 1
    for _ in range(rounds):
        changed = False
        for k in knobs:
            for v in KNOBS[k](g.n, cus):
                if best.get(k, 0) == v:
                    continue
                cand = dict(best, **{k: v})
                ctx.set_opts(dataclasses.replace(base, **cand))
                ms = _median_ms(ctx, schedule, variant, reps)
                trials += 1
                if ms < best_ms * (1 - min_gain):
                    labels, _ = ctx.run(schedule, variant)
                    if not np.array_equal(labels, ref):
                        raise RuntimeError(f"options {cand} changed the labels")
                    best, best_ms, changed = cand, ms, True
                    log(f"  {k}={v}: {ms:.3f} ms (kept)")
        if not changed:
            break
    ctx.set_opts(dataclasses.replace(base, **best))
    return {"opts": {k: v for k, v in best.items() if v != 0}, "ms": round(best_ms, 4),
            "default_ms": round(default_ms, 4), "trials": trials, "arch": arch}


WORKLOADS = {
    "rmat22": lambda: (scc.gen_rmat(22, 8, seed=2), False),
    "rmat24": lambda: (scc.gen_rmat(24, 16, seed=5), False),
    "mesh2048": lambda: (scc.gen_mesh(2048, 0.6, seed=3), False),
    "batch4096": lambda: (scc.gen_batch(4096, 64, 65536, 1.5, 6.0, 0.05, 0.5, seed=4), True),
}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workloads", nargs="+", default=list(WORKLOADS), choices=list(WORKLOADS))
    ap.add_argument("--schedule", default="h", choices=list(scc.SCHEDULES))
    ap.add_argument("--variants", nargs="+", type=int, default=[0])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--out", default=TUNED_PATH)
    args = ap.parse_args()
    table = load_table(args.out)
    with scc.Context(0) as ctx:
        for w in args.workloads:
            g, batch = WORKLOADS[w]()
            ctx.upload(g)
            ctx.set_batch(g.graph_offsets if batch else None)
            for var in args.variants:
                log(f"{w} variant {var}:")
                res = tune(ctx, g, args.schedule, var, reps=args.reps, rounds=args.rounds)
                res["graph"] = w
                key = class_key(graph_class(g, batch), args.schedule, var)
                table.setdefault(res["arch"], {})[key] = res
                log(f"{w} -> {key}: {res}")
                with open(args.out, "w") as f:
                    json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps(table))


if __name__ == "__main__":
    main()
