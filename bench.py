#!/usr/bin/env python3
"""Benchmark of the FB-SCC hot path on one MI355X.

A "step" is one whole SCC decomposition (scc_run: trim, pivots, forward and
backward reachability, colour updates, canonical labels) of the workload graph,
which is already resident in HBM when the timed region starts.  Default
workload: R-MAT scale 24, edge factor 16, a=0.57 b=0.19 c=0.19, seed 5
(16,777,216 vertices, 268,435,456 arcs drawn, duplicates and self-loops
removed), the graph the learned-versus-plain pivot comparison is quoted on.

value = MTEPS = arcs after dedup / seconds per decomposition.
cold_ms: the first decomposition after the upload (no chain lengths
remembered from an earlier run on the graph).
roofline: algorithmic bytes of the reachability levels per decomposition
(per unit: one state byte per vertex per sweep level, 13 bytes per expanded
vertex, 9 per scanned arc, 9 per new visit on the queue schedules, 1 on the
sweeps; counted by device counters in an untimed counting run) over the
reach time of the timed decompositions, taken with HIP events on the
library's stream around each reach chain.
traffic: HBM bytes per launch of the level kernel from rocprofv3
FETCH_SIZE / WRITE_SIZE runs of this workload (profiles/), when measured.
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
    """Bytes the reachability levels must move per decomposition, per unit of
    work as the reference lays its data out (a u8 state and a u32 colour per
    vertex): a sweep level reads one state byte per vertex (n + 1); an expanded
    vertex reads its row-offset pair, its colour and writes its state (13); a
    scanned arc reads its column id, the target's state and colour (9); a new
    visit writes one state byte, plus a queue push and pop (8) on the schedules
    with frontier queues."""
    S, X, E, V = prof["sweep_levels"], prof["expanded"], prof["edges"], prof["visited"]
    per_visit = 1 if sched in ("q", "t") else 9
    return S * (n + 1) + X * 13 + E * 9 + V * per_visit


def layout_bytes(sched, n, prof):
    """The same work counted in this build's layout (colour and state in one
    4-byte word per vertex): a sweep reads 4 bytes per vertex, an expanded
    vertex 12 (offsets 8, word 4), a scanned arc 8 (column id, target word), a
    new visit one 4-byte atomic plus 4 for its queue push (and 4 for the pop on
    the queue schedules)."""
    S, X, E, V = prof["sweep_levels"], prof["expanded"], prof["edges"], prof["visited"]
    if sched in ("q", "t"):
        return S * 4 * n + X * 12 + E * 8 + V * 4
    if sched in ("d", "l"):
        return X * 16 + E * 8 + V * 8
    return S * 4 * n + X * 12 + E * 8 + V * 8


def traffic_for(args):
    """HBM bytes per decomposition moved by the level kernel, from rocprofv3
    FETCH_SIZE / WRITE_SIZE runs of this exact configuration (profiles/), or
    None when it was not measured."""
    if not args.traffic or not os.path.exists(args.traffic):
        return None
    key = {"workload": args.workload, "schedule": args.schedule, "variant": args.variant,
           "scale": args.scale, "edge_factor": args.edge_factor, "seed": args.seed}
    if args.workload != "rmat":
        key = {k: key[k] for k in ("workload", "schedule", "variant")}
    for e in json.load(open(args.traffic)).get("entries", []):
        if all(e.get(k) == v for k, v in key.items()):
            return e
    return None


def write_levels(path, workload, rows):
    """Per-level table of one decomposition (profile flag 4 of the library)."""
    from gpuscc_code_amd import scc
    with open(path, "w") as f:
        f.write(f"# {workload}\n")
        f.write("\t".join("kernel_us" if k == "ms" else k for k in scc.LEVEL_FIELDS) + "\n")
        for r in rows:
            f.write("\t".join(f"{x:.1f}" if i == 2 else f"{int(x)}"
                              for i, x in enumerate([r[0], r[1], r[2] * 1e3] + list(r[3:]))) + "\n")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_tarjan(g, reps):
    """The oracle's Tarjan on the graph, pinned to one core: (labels, stats,
    DFS seconds of each repetition; the timed region is the DFS only)."""
    import oracle
    try:
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
    except (AttributeError, OSError):
        old = None
    times = []
    try:
        for _ in range(max(reps, 1)):
            ref, tst = oracle.tarjan(g.row, g.col)
            times.append(tst["seconds"])
    finally:
        if old is not None:
            os.sched_setaffinity(0, old)
    return ref, tst, times


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--schedule", default="h", choices=list("qtdlhcb"),
                    help="q t d l h as the reference's -a; c = COMBINE, b = BEST")
    ap.add_argument("--variant", type=int, default=0, choices=[0, 1, 2])
    ap.add_argument("--workload", default="rmat", choices=["rmat", "mesh", "batch"],
                    help="rmat: R-MAT (default: scale 24, edge factor 16, seed 5); mesh: 2048x2048 grid; "
                         "batch: 4096 uneven ER graphs with planted cycles")
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--no-permute", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=3,
                    help="Tarjan repetitions of the CPU baseline (median reported)")
    ap.add_argument("--sequential", action="store_true",
                    help="forward reach to its fixpoint, then backward (the reference's order)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run the calibration copy kernel over MIB MiB (for traffic counters)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "k_reach_hybrid_traffic.json"),
                    metavar="JSON",
                    help="measured HBM bytes of the level kernel (from rocprofv3 counters)")
    ap.add_argument("--levels", default=None, metavar="FILE",
                    help="also write the per-level log of one decomposition (kernel time, frontier, "
                         "push/pull, work) to FILE")
    ap.add_argument("--grid-blocks", type=int, default=0,
                    help="workgroups of the persistent kernels (0 = 4 per CU)")
    ap.add_argument("--pull-arc-div", type=int, default=0,
                    help="hybrid: frontier arcs >= m / DIV also pull (0 = 14, -1 = off)")
    ap.add_argument("--pull-threshold", type=int, default=0,
                    help="hybrid: frontier size at which a direction pulls (0 = n/20)")
    ap.add_argument("--relabel", type=int, default=0, choices=[-1, 0, 1],
                    help="renumbering at upload: 0 by degree skew (library default), 1 always, -1 never")
    ap.add_argument("--tile-reach", type=int, default=0,
                    help="tile reach of the first FB step (scc_opts.tile_reach): 0 the library default (tiles for "
                         "graphs of high diameter), -1 off, 1 tiles of 4096 vertices, T > 1 tiles of T vertices")
    ap.add_argument("--no-tuned", action="store_true",
                    help="library defaults instead of the autotuned options (gpuscc_code_amd/tuned.json)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-pivot leg (same workload and options, pivots elected by the "
                         "trained MLP scorer, gpuscc_code_amd/pivot_mlp.npy)")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the labels check against the oracle's Tarjan (profiling runs)")
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
                                     grid_blocks=args.grid_blocks, relabel=args.relabel,
                                     tile_reach=args.tile_reach))
    t0 = time.time()
    ctx.upload(g)
    log(f"uploaded in {time.time() - t0:.1f}s")
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
                                           batch=args.workload == "batch",
                                           base=scc.alg_opt(relabel=args.relabel, tile_reach=args.tile_reach))
        if tuned_key:
            ctx.set_opts(opt)
            opts = autotune.load_table()[arch][tuned_key]["opts"]

    if args.calibrate_copy:
        ctx.profile_copy(args.calibrate_copy << 20)
    runs = 0  # decompositions this process runs (for per-decomposition counter totals)

    # cold: the first decomposition after the upload (and option change), with no
    # chain lengths remembered from an earlier run on this graph
    t0 = time.perf_counter()
    _, st_cold = ctx.run(args.schedule, args.variant, labels=False)
    cold_ms = (time.perf_counter() - t0) * 1e3
    runs += 1

    # untimed counting run: algorithmic work of the reachability levels (the
    # counters are cleared first: the cold run's resumed chains add to them)
    ctx.profile(1)
    ctx.profile_read(reset=True)
    ctx.run(args.schedule, args.variant, labels=False)
    work = ctx.profile_read(reset=True)
    ctx.profile(0)
    runs += 1
    for _ in range(args.warmup):
        ctx.run(args.schedule, args.variant, labels=False)
        runs += 1

    # timed region: K decompositions, nothing instrumented (the reach chains of
    # the speculative steps are bracketed by HIP events on the library's stream,
    # which give ms_fwd)
    stats = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        _, st = ctx.run(args.schedule, args.variant, labels=False)
        stats.append(st)
    t_total = time.perf_counter() - t_start
    runs += args.steps

    # cross-check of the level-kernel time: K more decompositions with every level
    # kernel bracketed by its own pair of HIP events
    ctx.profile(2)
    ctx.profile_read(reset=True)
    for _ in range(args.steps):
        ctx.run(args.schedule, args.variant, labels=False)
    prof = ctx.profile_read(reset=True)
    ctx.profile(0)
    runs += args.steps

    # one more decomposition with the phases of the one-workgroup tail's steps timed (on
    # the same path as the timed ones)
    ctx.profile(8)
    ctx.profile_read(reset=True)
    _, st_tail = ctx.run(args.schedule, args.variant, labels=False)
    tailp = ctx.profile_read(reset=True)
    ctx.profile(0)
    runs += 1

    ms_per_step = t_total * 1e3 / args.steps
    mteps = g.m / (ms_per_step * 1e3)

    # learned-pivot leg: the same decomposition with pivots elected by the trained
    # scorer
    learned = None
    if not args.no_learned:
        from gpuscc_code_amd import learned as learned_mod
        ctx.mlp_set(learned_mod.load_params())
        for _ in range(max(args.warmup, 1)):
            _, st_l = ctx.run(args.schedule, args.variant, pivot=scc.PIVOT_LEARNED, labels=False)
            runs += 1
        st_ls = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, st_l = ctx.run(args.schedule, args.variant, pivot=scc.PIVOT_LEARNED, labels=False)
            st_ls.append(st_l)
        ms_l = (time.perf_counter() - t0) * 1e3 / args.steps
        runs += args.steps
        lab_l = None
        if not args.no_check:
            lab_l, _ = ctx.run(args.schedule, args.variant, pivot=scc.PIVOT_LEARNED)
            runs += 1
        learned = {"ms_per_step": round(ms_l, 3), "value": round(g.m / (ms_l * 1e3), 2), "unit": "MTEPS",
                   "ms_per_graph": round(ms_l / num_graphs, 6), "fb_steps": st_l["steps"],
                   "plain_fb_steps": stats[-1]["steps"],
                   "fwd_depths": st_l["fwd_depths"], "plain_fwd_depths": stats[-1]["fwd_depths"],
                   "host_syncs": st_l["host_syncs"],
                   "phase_ms": {k: round(statistics.mean(s[k] for s in st_ls), 3)
                                for k in ("ms_fwd", "ms_bwd", "ms_trim", "ms_update")},
                   "plain_over_learned": round(ms_per_step / ms_l, 3), "labels": lab_l}

    if args.levels:  # one more decomposition, levels launched one at a time
        ctx.profile(4)
        ctx.run(args.schedule, args.variant, labels=False)
        write_levels(args.levels, wl, ctx.profile_levels())
        ctx.profile(0)
        runs += 1

    # correctness of the measured configuration, and the CPU baseline
    bit_exact, st_check, tst, cpu = None, None, None, None
    if not args.no_check:
        labels, st_check = ctx.run(args.schedule, args.variant)
        runs += 1
        reps = 0 if args.no_cpu else args.cpu_reps
        ref, tst, cpu_times = cpu_tarjan(g, reps)
        bit_exact = bool(np.array_equal(labels, ref))
        if learned is not None:
            learned["bit_exact"] = bool(np.array_equal(learned.pop("labels"), ref))
        if not args.no_cpu:
            t = statistics.median(cpu_times)
            cpu = {"value": round(g.m / t / 1e6, 2), "unit": "MTEPS", "cores": 1, "kind": "port",
                   "ms": round(t * 1e3, 1),
                   "sample": f"the oracle's iterative Tarjan (single thread, pinned to one core) on the "
                             f"whole graph, DFS only, median of {len(cpu_times)} runs of "
                             f"{', '.join(f'{x * 1e3:.0f}' for x in cpu_times)} ms",
                   "host_cpu": cpu_model(), "host_threads": os.cpu_count()}
    elif learned is not None:
        learned.pop("labels")

    # roofline of the reachability level kernel
    per_run_bytes = algorithmic_bytes(args.schedule, g.n, work)
    reach_ms = statistics.mean(s["ms_fwd"] + s["ms_bwd"] for s in stats)  # timed region, HIP events
    last = stats[-1]
    launches_live = max(last["fwd_depths"], 1)  # level kernels of the timed decompositions' reach chains
    achieved = per_run_bytes / (reach_ms * 1e-3) / 1e9 if reach_ms > 0 else 0.0
    launches_ev = prof["reach_launches"] / args.steps
    kernel_ms_ev = prof["reach_kernel_ms"] / args.steps
    tr = traffic_for(args)
    traffic = None
    if tr is not None:
        traffic = int(tr["bytes_per_decomposition"] / launches_live)
    roofline = {
        "bound": "hbm", "kernel": KERNEL_OF[args.schedule],
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "bytes_per_launch": int(per_run_bytes / launches_live),
        "avg_launch_us": round(reach_ms * 1e3 / launches_live, 2),
        "launches_per_step": launches_live, "reach_ms_per_step": round(reach_ms, 4),
        "bytes_per_step": int(per_run_bytes),
        "layout_bytes_per_step": int(layout_bytes(args.schedule, g.n, work)),
        "traffic_over_algorithmic": None if traffic is None else round(traffic * launches_live / per_run_bytes, 2),
        "events_per_kernel": {"launches_per_step": launches_ev, "kernel_ms_per_step": round(kernel_ms_ev, 4),
                              "avg_launch_us": round(kernel_ms_ev * 1e3 / max(launches_ev, 1), 2)},
        "work_per_step": {k: work[k] for k in ("sweep_levels", "queue_levels", "expanded", "edges",
                                               "visited", "tile_runs", "tile_rounds", "tile_phase_ms",
                                               "tile_busy_ms", "tile_idle_polls")},
    }
    out = {
        "metric": METRIC, "value": round(mteps, 2), "unit": "MTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "ms_per_graph": round(ms_per_step / num_graphs, 6), "graphs_per_s": round(num_graphs * 1e3 / ms_per_step, 1),
        "cold_ms": round(cold_ms, 3), "cold_host_syncs": st_cold["host_syncs"],
        "config": {"workload": wl, "graphs": num_graphs,
                   "n": g.n, "m": g.m, "schedule": args.schedule,
                   "variant": ["scc", "scc_w", "scc_2_w"][args.variant], "pivot_rule": "plain",
                   "reach": "sequential" if args.sequential else "fused",
                   "tuned": tuned_key, "opts": opts, "renumbered": bool(last["renumbered"]),
                   "tiles": last["renumbered"] == 2},
        "bit_exact": bit_exact, "num_scc": None if st_check is None else st_check["num_scc"],
        "biggest_scc": None if tst is None else tst["biggest"],
        "fb_steps": last["steps"], "fwd_depths": last["fwd_depths"], "bwd_depths": last["bwd_depths"],
        "trim_iters": last["trim_iters"], "host_syncs": last["host_syncs"],
        "phase_ms": {k: round(statistics.mean(s[k] for s in stats), 3)
                     for k in ("ms_fwd", "ms_bwd", "ms_trim", "ms_update")},
        "decompositions_run": runs,
        "tail": {"steps": st_tail["tail_steps"], "live_at_start": st_tail["tail_live"],
                 "phase_ms": {k: round(x, 4) for k, x in zip(("reach", "trim", "update", "confirm"),
                                                               tailp["tail_phase_ms"])},
                 "lds_copy_live": tailp["tail_local_live"], "lds_copy_step": tailp["tail_local_step"]},
        "learned_pivot": learned, "roofline": roofline, "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
