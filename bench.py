#!/usr/bin/env python3
"""Benchmark of the FB-SCC hot path on one MI355X.

A "step" is one whole SCC decomposition (scc_run: trim, pivots, forward and
backward reachability, colour updates, canonical labels) of the workload graph,
which is already resident in HBM when the timed region starts.  Default
workload: R-MAT scale 22 (4,194,304 vertices, 33,554,432 arcs drawn, edge
factor 8, a=0.57 b=0.19 c=0.19, seed 2, duplicates and self-loops removed).

value = MTEPS = arcs after dedup / seconds per decomposition.
roofline: algorithmic bytes of the reachability level kernels (from device
work counters in an untimed counting run) over their summed duration, timed
with HIP events on the library's stream during the timed region.
cpu_baseline: the oracle's single-threaded Tarjan on the same graph.
learned_pivot: the same decompositions with the trained MLP pivot scorer
(time, FB steps, plain time / learned time, labels bit-exact).

Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCC MTEPS + ms/graph, learned-pivot vs plain FB on 1 MI355X (bit-exact)"
HBM_PEAK_GBS = 8000.0
KERNEL_OF = {"q": "k_reach_sweep", "t": "k_reach_topo", "d": "k_reach_queue", "l": "k_reach_lb",
             "h": "k_reach_hybrid", "c": "k_reach_lb+k_reach_sweep", "b": "k_reach_hybrid"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(sched, n, prof):
    """Bytes the reachability levels must move, per the data layout:
    a sweep level reads every 4-byte vertex word; an expanded vertex reads its
    row-offset pair (8) and writes its word (4); a scanned arc reads its column
    id (4) and gathers the target's word (4); a new visit is one 4-byte atomic,
    plus a 4-byte queue push and (on queue levels) a 4-byte queue read."""
    S, X, E, V = prof["sweep_levels"], prof["expanded"], prof["edges"], prof["visited"]
    if sched in ("q", "t"):
        return S * 4 * n + X * 12 + E * 8 + V * 4
    if sched in ("d", "l"):
        return X * 16 + E * 8 + V * 8
    return S * 4 * n + X * 12 + E * 8 + V * 8


def traffic_for(args):
    """HBM bytes per launch of the level kernel measured by rocprofv3 counters
    (profiles/), when they were measured for this exact configuration."""
    if not args.traffic or not os.path.exists(args.traffic):
        return None
    t = json.load(open(args.traffic))
    if (args.workload, args.schedule, args.scale, args.edge_factor, args.seed) != ("rmat", "h", 22, 8, 2):
        return None
    return t["bytes_per_launch"]


def write_levels(path, workload, rows):
    """Per-level table of one decomposition (profile flag 4 of the library)."""
    from gpuscc_code_amd import scc
    with open(path, "w") as f:
        f.write(f"# {workload}\n")
        f.write("\t".join("kernel_us" if k == "ms" else k for k in scc.LEVEL_FIELDS) + "\n")
        for r in rows:
            f.write("\t".join(f"{x:.1f}" if i == 2 else f"{int(x)}"
                              for i, x in enumerate([r[0], r[1], r[2] * 1e3] + list(r[3:]))) + "\n")


def cpu_tarjan_baseline(g, reps=3):
    import oracle
    try:
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
    except (AttributeError, OSError):
        old = None
    times = []
    try:
        for _ in range(reps):
            t, _cnt = oracle.tarjan_time(g.row, g.col)
            times.append(t)
    finally:
        if old is not None:
            os.sched_setaffinity(0, old)
    t = statistics.median(times)
    return {"value": round(g.m / t / 1e6, 2), "unit": "MTEPS", "cores": 1, "kind": "port",
            "sample": f"oracle Tarjan (single thread, pinned) on the whole graph, median of {reps} "
                      f"runs, {t * 1e3:.1f} ms each; host has {os.cpu_count()} CPUs"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--schedule", default="h", choices=list("qtdlhcb"),
                    help="q t d l h as the reference's -a; c = COMBINE, b = BEST")
    ap.add_argument("--variant", type=int, default=0, choices=[0, 1, 2])
    ap.add_argument("--workload", default="rmat", choices=["rmat", "mesh", "batch"],
                    help="rmat: R-MAT (default, scale 22); mesh: 2048x2048 grid; "
                         "batch: 4096 uneven ER graphs with planted cycles")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=8)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-permute", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sequential", action="store_true",
                    help="forward reach to its fixpoint, then backward (the reference's order)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run the calibration copy kernel over MIB MiB (for traffic counters)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "k_reach_hybrid_rmat22_traffic.json"),
                    metavar="JSON",
                    help="measured HBM bytes per launch of the level kernel (from rocprofv3 counters)")
    ap.add_argument("--levels", default=None, metavar="FILE",
                    help="also write the per-level log of one decomposition (kernel time, frontier, "
                         "push/pull, work) to FILE")
    ap.add_argument("--grid-blocks", type=int, default=0,
                    help="workgroups of the persistent kernels (0 = 4 per CU)")
    ap.add_argument("--pull-arc-div", type=int, default=0,
                    help="hybrid: frontier arcs >= m / DIV also pull (0 = 14, -1 = off)")
    ap.add_argument("--pull-threshold", type=int, default=0,
                    help="hybrid: frontier size at which a direction pulls (0 = n/20)")
    ap.add_argument("--no-tuned", action="store_true",
                    help="library defaults instead of the autotuned options (gpuscc_code_amd/tuned.json)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-pivot leg (same workload and options, pivots elected by the "
                         "trained MLP scorer, gpuscc_code_amd/pivot_mlp.npy)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the single-GPU figure is the one reported", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from gpuscc_code_amd import autotune, scc

    t0 = time.time()
    if args.workload == "rmat":
        g = scc.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed,
                         permute=not args.no_permute)
        wl = (f"R-MAT scale {args.scale}, edge factor {args.edge_factor}, a=0.57 b=0.19 c=0.19, "
              f"seed {args.seed}, dedup, no self-loops, {'un' if args.no_permute else ''}permuted ids")
    elif args.workload == "mesh":
        g = scc.gen_mesh(2048, 0.6, seed=3)
        wl = "2048x2048 directed mesh, lattice edge bidirectional w.p. 0.6 else one random direction, seed 3"
    else:
        g = scc.gen_batch(4096, 64, 65536, 1.5, 6.0, 0.05, 0.5, seed=4)
        wl = ("4096 directed ER graphs, n log-uniform in [64, 65536], mean out-degree U(1.5, 6), "
              "planted cycles Geometric(0.05) over half the vertices, seed 4, packed CSR")
    log(f"generated {args.workload}: n={g.n} m={g.m} in {time.time() - t0:.1f}s")
    ctx = scc.Context(0, scc.alg_opt(sequential_reach=1 if args.sequential else 0,
                                     hybrid_threshold=args.pull_threshold,
                                     hybrid_arc_div=args.pull_arc_div,
                                     grid_blocks=args.grid_blocks))
    ctx.upload(g)
    num_graphs = 1
    if args.workload == "batch":
        ctx.set_batch(g.graph_offsets)
        num_graphs = len(g.graph_offsets) - 1
    # autotuned options for this GPU and graph class, unless any knob is given by hand
    tuned_key, opts = None, {}
    manual = args.sequential or args.grid_blocks or args.pull_arc_div or args.pull_threshold
    if not args.no_tuned and not manual:
        arch, _ = ctx.device_info()
        opt, tuned_key = autotune.opts_for(g, arch, args.schedule, args.variant,
                                           batch=args.workload == "batch")
        if tuned_key:
            ctx.set_opts(opt)
            opts = autotune.load_table()[arch][tuned_key]["opts"]

    if args.calibrate_copy:
        ctx.profile_copy(args.calibrate_copy << 20)

    # untimed counting run: algorithmic work of the reachability levels
    ctx.profile(1)
    ctx.run(args.schedule, args.variant, labels=False)
    work = ctx.profile_read(reset=True)
    ctx.profile(0)
    for _ in range(args.warmup):
        ctx.run(args.schedule, args.variant, labels=False)

    # timed region: K decompositions, nothing instrumented
    stats = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        _, st = ctx.run(args.schedule, args.variant, labels=False)
        stats.append(st)
    t_total = time.perf_counter() - t_start

    # kernel durations for the roofline: K more decompositions with every level
    # kernel bracketed by HIP events on its stream
    ctx.profile(2)
    ctx.profile_read(reset=True)
    for _ in range(args.steps):
        ctx.run(args.schedule, args.variant, labels=False)
    prof = ctx.profile_read(reset=True)
    ctx.profile(0)

    ms_per_step = t_total * 1e3 / args.steps
    mteps = g.m / (ms_per_step * 1e3)

    # learned-pivot leg: the same decomposition with pivots elected by the trained
    # scorer (its static features are computed by the warm-up run, once per graph)
    learned = None
    if not args.no_learned:
        from gpuscc_code_amd import learned as learned_mod
        ctx.mlp_set(learned_mod.load_params())
        for _ in range(max(args.warmup, 1)):
            _, st_l = ctx.run(args.schedule, args.variant, pivot=scc.PIVOT_LEARNED, labels=False)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, st_l = ctx.run(args.schedule, args.variant, pivot=scc.PIVOT_LEARNED, labels=False)
        ms_l = (time.perf_counter() - t0) * 1e3 / args.steps
        lab_l, _ = ctx.run(args.schedule, args.variant, pivot=scc.PIVOT_LEARNED)
        learned = {"ms_per_step": round(ms_l, 3), "value": round(g.m / (ms_l * 1e3), 2), "unit": "MTEPS",
                   "ms_per_graph": round(ms_l / num_graphs, 6), "fb_steps": st_l["steps"],
                   "plain_fb_steps": stats[-1]["steps"],
                   "plain_over_learned": round(ms_per_step / ms_l, 3), "labels": lab_l}

    if args.levels:  # one more decomposition, levels launched one at a time
        ctx.profile(4)
        ctx.run(args.schedule, args.variant, labels=False)
        write_levels(args.levels, wl, ctx.profile_levels())
        ctx.profile(0)

    # correctness of the measured configuration
    labels, st_check = ctx.run(args.schedule, args.variant)
    import oracle
    ref, tst = oracle.tarjan(g.row, g.col)
    bit_exact = bool(np.array_equal(labels, ref))
    if learned is not None:
        learned["bit_exact"] = bool(np.array_equal(learned.pop("labels"), ref))

    per_run_bytes = algorithmic_bytes(args.schedule, g.n, work)
    launches = prof["reach_launches"] / args.steps
    kernel_ms = prof["reach_kernel_ms"] / args.steps
    achieved = per_run_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    roofline = {
        "bound": "hbm", "kernel": KERNEL_OF[args.schedule],
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_for(args),
        "bytes_per_launch": int(per_run_bytes / max(launches, 1)),
        "avg_launch_us": round(kernel_ms * 1e3 / max(launches, 1), 2),
        "launches_per_step": launches, "kernel_ms_per_step": round(kernel_ms, 3),
        "work_per_step": {k: work[k] for k in ("sweep_levels", "queue_levels", "expanded", "edges",
                                               "visited")},
    }
    cpu = None if args.no_cpu else cpu_tarjan_baseline(g)
    last = stats[-1]
    out = {
        "metric": METRIC, "value": round(mteps, 2), "unit": "MTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "ms_per_graph": round(ms_per_step / num_graphs, 6), "graphs_per_s": round(num_graphs * 1e3 / ms_per_step, 1),
        "config": {"workload": wl, "graphs": num_graphs,
                   "n": g.n, "m": g.m, "schedule": args.schedule,
                   "variant": ["scc", "scc_w", "scc_2_w"][args.variant], "pivot_rule": "plain",
                   "reach": "sequential" if args.sequential else "fused",
                   "tuned": tuned_key, "opts": opts},
        "bit_exact": bit_exact, "num_scc": st_check["num_scc"], "biggest_scc": tst["biggest"],
        "fb_steps": last["steps"], "fwd_depths": last["fwd_depths"], "bwd_depths": last["bwd_depths"],
        "trim_iters": last["trim_iters"], "host_syncs": last["host_syncs"],
        "phase_ms": {k: round(statistics.mean(s[k] for s in stats), 3)
                     for k in ("ms_fwd", "ms_bwd", "ms_trim", "ms_update")},
        "learned_pivot": learned, "roofline": roofline, "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
