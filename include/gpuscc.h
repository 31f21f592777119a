/*
 * gpuscc.h -- C ABI of the MI355X (gfx950) Forward-Backward SCC library.
 *
 * The reference (chenxuhao/gpuscc_code) has no FFI: its path sits behind the
 * `./scc [-t] [-a q|t|d|l|h] [-b N] -f graph` process interface (scc.cu:57-148)
 * and an internal C++ header API of phase functions on device pointers
 * (reach_q.h:8-21, reach_d.h:14-25, scc_routines.h:8-50, wcc.h:18-30,
 * load.h:5, tarjan.h:5).  Each entry point below names the reference
 * interface it replaces.
 *
 * Conventions: plain C types only; vertex ids are 0-based (the reference's
 * files and arrays are 1-based, load.cpp:128-139); the CSR is int32
 * (row[n+1], col[m]), rows sorted ascending and duplicate-free.  Host buffers
 * are caller-owned; device buffers are context-owned; one HIP stream per
 * context; a context is not thread-safe.  Every function returns an scc_status
 * and never calls exit() (the reference exits on any error, GPUerrchk.h:6-12).
 */
#ifndef GPUSCC_H
#define GPUSCC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SCC_OK = 0,
    SCC_E_HIP = 1,       /* a HIP runtime call failed (no GPU, launch error, ...) */
    SCC_E_OOM = 2,       /* device or host allocation failed */
    SCC_E_INPUT = 3,     /* malformed graph, file or argument */
    SCC_E_OVERFLOW = 4,  /* a size exceeds a documented limit (never a silent drop) */
    SCC_E_STATE = 5      /* call out of order (e.g. run before upload) */
} scc_status;

/* Reachability schedules: reach_strategy (scc.h:7-16), CLI -a q|t|d|l|h (scc.cu:94-104). */
typedef enum {
    SCC_SCHED_Q = 0,  /* topology sweep, one lane per vertex (reach_q.cu) */
    SCC_SCHED_T = 1,  /* topology sweep with workgroup edge balancing (reach_t.cu) */
    SCC_SCHED_D = 2,  /* data-driven frontier queues (reach_d.cu) */
    SCC_SCHED_L = 3,  /* frontier queues with load-balanced edge expansion (reach_l.cu) */
    SCC_SCHED_H = 4,  /* hybrid: hybrid_policy 0 (default) pushes or pulls each level and direction by
                         its frontier's size; hybrid_policy 1 is the reference's HYBRID rule
                         (scc.cu:493-563, switch :877-891): t until n/100 vertices are
                         FB-eliminated, then q */
    SCC_SCHED_C = 5,  /* COMBINE: l until n/100 vertices are FB-eliminated, then q (scc.cu:565-638) */
    SCC_SCHED_B = 6   /* BEST (scc.cu:640-700, reach_b.cu:374-725): in the first FB step a
                         direction-optimising reach (top-down until the frontier's arcs exceed the
                         unexplored arcs / 15, bottom-up while the frontier grows or holds more
                         than n/18 vertices); later steps t until n/100 vertices are
                         FB-eliminated, then q */
} scc_schedule;

/* The three binaries of Makefile:12-19. */
typedef enum {
    SCC_VARIANT_PLAIN = 0,  /* scc */
    SCC_VARIANT_W = 1,      /* scc_w: one WCC pass once >= n/100 vertices are FB-eliminated */
    SCC_VARIANT_2W = 2      /* scc_2_w: scc_w plus one trim-2 pass before the WCC pass */
} scc_variant;

typedef enum {
    SCC_PIVOT_PLAIN = 0,    /* lowest live id per colour slot (the reference rule, made deterministic) */
    SCC_PIVOT_LEARNED = 1   /* highest MLP score per colour slot (needs scc_mlp_set_params) */
} scc_pivot_rule;

/* Options: alg_opt (scc.h:19-27).  Zero means "default". */
typedef struct {
    int32_t block_size;        /* workgroup size for sweeps (-b, scc.cu:82); default 256 */
    int32_t hybrid_policy;     /* 0: a direction pulls when its frontier reaches hybrid_threshold vertices
                                  or m / hybrid_arc_div arcs; 1: the reference's rule, sweep-with-
                                  balancing until >= n/100 FB-eliminated, then q; 2: pulls when its
                                  frontier's arcs reach twice the live vertices it has not reached
                                  (the work of a pull level) */
    int32_t hybrid_threshold;  /* h: frontier size (vertices) at or above which a level sweeps */
    int32_t slot_bits;         /* log2 of the election table size; default from n */
    int32_t grid_blocks;       /* workgroups of the grid-stride kernels; default from the CU count */
    int32_t sequential_reach;  /* 1: forward reach to its fixpoint, then backward (scc.cu:285-704);
                                  0: both directions advance in the same level kernels */
    int32_t huge_row;          /* rows of at least this many arcs are split over all workgroups
                                  (deferred one level); default 1024 for q/d, 4096 for t/l/h */
    int32_t local_rounds;      /* d, l, h: after its share of a level, a workgroup keeps expanding its own
                                  new visits in LDS for up to this many rounds (0 = default 64,
                                  -1 = off) */
    int32_t hybrid_arc_div;    /* h: a direction also pulls when the arcs of its frontier (estimated
                                  from degree classes) reach m / this (0 = default 14, -1 = off) */
    int32_t reach_graphs;      /* 1: the reach chains of speculative steps are captured once as HIP
                                  graphs and replayed (one launch per chain); 0 = launched level by level */
    int32_t reach_block;       /* workgroup size of the level kernels (reach t/d/l/h, trim): 256 or 512
                                  (0 = 256) */
    int32_t tail_live;         /* once at most this many vertices are live, the remaining FB steps run in
                                  one workgroup and one launch (0 = default 65536, -1 = off) */
    int32_t batch_solver;      /* batches (scc_set_batch), plain variant and pivot rule: 0 (default) all
                                  graphs advance together in the level kernels, their elections sharing
                                  one table; 1: every graph is solved on its own in one persistent launch,
                                  a workgroup per graph at a time, largest first, each with its own
                                  election tables (one host sync; slower than 0 once graphs exceed a
                                  few thousand vertices) */
    int32_t relabel;           /* read by scc_upload (and scc_set_batch): renumber the vertices inside the
                                  library by decreasing degree (within each graph of a batch), hubs
                                  first, so the words and rows the scans touch most share cache lines.
                                  Results are unchanged: elections, roots and labels use the caller's
                                  ids.  0 (default): when the degrees are skewed (largest above 64
                                  times the mean) and n >= 65536; 1: always; -1: never */
    int32_t tile_reach;        /* read by scc_upload: number the vertices tile by tile (regions of up to
                                  4,096 vertices grown by BFS, made of cells of 64 grown by BFS inside
                                  them) so that the reach of the first FB step (schedule h,
                                  hybrid_policy 0, single graph) closes cells in registers, tiles in
                                  one workgroup each, and crosses tile boundaries between iterations.
                                  Results are unchanged.  0 (default): tiles of up to 4,096 vertices
                                  when the degrees are not skewed, relabel is 0, n >= 65536 and a BFS
                                  from vertex 0 over the arcs taken both ways is at least 256 levels
                                  deep (high diameter: meshes, road-like graphs); -1: never; 1:
                                  always, tiles of up to 4,096; N > 1: always, tiles of at most N */
} scc_opts;

/* Run statistics (the stdout block of scc.cu:998-1005 plus counters). */
typedef struct {
    int64_t steps;          /* FB steps (->steps) */
    int64_t fwd_depths;     /* sum of forward levels (->sum of fwdDepths) */
    int64_t bwd_depths;     /* sum of backward levels */
    int64_t trim_iters;     /* sum of trim sweeps */
    int64_t num_scc;        /* count of P bits (->count of SCC, scc.cu:914-919) */
    int64_t pivots_total;   /* pivots elected over the run */
    int64_t wcc_step;       /* step at which the WCC pass ran, 0 if never */
    int64_t trim2_pairs;    /* pairs removed by trim-2 */
    int64_t trivial;        /* size-1 SCCs (tarjan.cpp:237) */
    int64_t biggest;        /* largest SCC size */
    int64_t host_syncs;     /* device->host waits the run made */
    int64_t first_pivot;    /* first pivot (lowest live id after trim-1, scc.cu:235-245), -1 if none */
    int64_t tail_steps;     /* of steps, those run in the one-workgroup tail (scc_opts.tail_live) */
    int64_t tail_live;      /* live vertices when the tail started, 0 if it did not run */
    int64_t renumbered;     /* 1: the upload renumbered the vertices by degree (scc_opts.relabel),
                               2: tile by tile (scc_opts.tile_reach); 0: neither */
    double ms_total;        /* upload-excluded wall time, labels ready (scc.cu:217->1011) */
    double ms_fwd, ms_bwd, ms_trim, ms_update; /* per phase, host clock (STOPWATCH, timing.h:14) */
} scc_stats;

typedef struct scc_ctx scc_ctx;

/* Host-side graph owned by the library (scc_graph_free releases it). */
typedef struct {
    int32_t n;
    int64_t m;              /* arcs after dedup */
    int32_t *row, *col;     /* CSR, row[n+1] */
    int32_t *rowT, *colT;   /* transposed CSR */
    int64_t m_reported;     /* load.cpp's countOfEdges (nnz, or nnz + mirrored for symmetric) */
    int32_t pattern, symmetric, mtx;
    int32_t max_in, min_in, max_out, min_out;
    double average;         /* nnz / n, as load.cpp:60 */
    int32_t num_graphs;     /* batched graphs only: number of graphs, else 0 */
    int32_t *graph_offsets; /* batched graphs only: vertex offsets [num_graphs+1] */
} scc_host_graph;

/* ---- context ------------------------------------------------------------ */
/* Replaces the device setup of computeSCC (scc.cu:199-215). */
scc_status scc_create(int device, const scc_opts *opts, scc_ctx **out);
/* Changes the run-time options of a context (workgroup size, grid, hybrid
 * thresholds, huge rows, local rounds, reach order) between runs; slot_bits
 * is fixed once a graph is uploaded (0 keeps it).  No reference counterpart:
 * alg_opt is fixed per process there (scc.cu:76-109); the autotuner needs it. */
scc_status scc_set_opts(scc_ctx *ctx, const scc_opts *opts);
/* Device of the context: gfx target name (e.g. "gfx950") and compute units. */
scc_status scc_device_info(const scc_ctx *ctx, char *arch, int32_t len, int32_t *cu_count);
void scc_destroy(scc_ctx *ctx);
const char *scc_status_string(scc_status s);
const char *scc_last_error(const scc_ctx *ctx);

/* Copies a CSR + CSR^T into HBM (replaces the cudaMalloc/cudaMemcpy of
 * scc.cu:200-211).  rowT/colT may be NULL: the library then builds the
 * transpose (on the host) before copying.  n >= 0, 0 <= m < 2^31. */
scc_status scc_upload(scc_ctx *ctx, int32_t n, int64_t m, const int32_t *row, const int32_t *col,
                      const int32_t *rowT, const int32_t *colT);

/* Whole FB run: computeSCC (scc.cu:150-1022) for one schedule and variant.
 * labels_out (host, n entries, may be NULL) receives canonical labels: the
 * minimum vertex id of each vertex's SCC.  stats may be NULL. */
scc_status scc_run(scc_ctx *ctx, int schedule, int variant, int pivot_rule, int32_t *labels_out,
                   scc_stats *stats);

/* Batched graphs: the uploaded CSR is the disjoint union of num_graphs graphs,
 * graph g owning vertices [graph_offsets[g], graph_offsets[g+1]).  Later
 * scc_run calls give every graph its own initial colour (g + 1), so each
 * graph gets its own pivot per step and all graphs advance in the same
 * launches (the reference's colour mechanism, scc_routines.cu:207-221,
 * applied to the union).  No reference counterpart: the reference solves one
 * graph per process.  Fails with SCC_E_INPUT if an arc crosses graphs.
 * num_graphs = 0 clears the setting.  Labels stay global vertex ids. */
scc_status scc_set_batch(scc_ctx *ctx, const int32_t *graph_offsets, int32_t num_graphs);

/* Device address of the labels of the last run.  Valid until the next run,
 * phase-level call, scc_state_set or upload (after a plain-rule run the
 * labels are the root array itself; a later call that writes roots first moves
 * them to the label buffer, so ask again for the address after such a call). */
scc_status scc_labels_device(scc_ctx *ctx, const int32_t **dptr);

/* ---- phase-level entry points (tests, autotuner) -------------------------
 * They act on the context's vertex words: word = colour << 8 | state, with the
 * state bits of scc.h:48-93 (B=1 F=2 E=4 T=8 P=16 BExt=32 FExt=64) and a
 * 24-bit colour. */
scc_status scc_state_set(scc_ctx *ctx, const uint32_t *words, const int32_t *roots);
scc_status scc_state_get(scc_ctx *ctx, uint32_t *words, int32_t *roots);
/* trim1/trim (scc_routines.cu:584-643): trim within colours to the fixpoint. */
scc_status scc_phase_trim(scc_ctx *ctx, int64_t *iterations);
/* fwd_reach_X / bwd_reach_X (reach_q.h:8-21, reach_t.h, reach_d.h, reach_l.h):
 * dir 0 = forward over the CSR (bit F), 1 = backward over CSR^T (bit B),
 * 2 = both in the same level kernels.  Seeds are the live vertices holding
 * the bit.  *depth = level kernels that did work. */
scc_status scc_phase_reach(scc_ctx *ctx, int dir, int schedule, int64_t *depth);
/* update (scc_routines.cu:526-555): recolour = 1 splits colours and eliminates
 * F&B; 0 only elects (first pivot, scc.cu:235-245). */
scc_status scc_phase_update(scc_ctx *ctx, int recolour, int64_t *pivots);
/* trim2 (scc_routines.cu:645-668): one pass. */
scc_status scc_phase_trim2(scc_ctx *ctx, int64_t *pairs);
/* update_wcc (wcc.cu:146-226): WCC labels, fresh colours, election. */
scc_status scc_phase_wcc(scc_ctx *ctx, int64_t *iterations, int64_t *pivots);
/* countElimmed (scc.cu:42-51). */
scc_status scc_phase_count_elimmed(scc_ctx *ctx, int64_t *count);
/* unsetPivots (scc_routines.cu:10-27): every vertex in the pivot state 19
 * (P|F|B) goes back to state 0, as before the WCC pass (scc.cu:796). */
scc_status scc_phase_unset_pivots(scc_ctx *ctx);
/* The colour split of update without the election (scc_routines.cu:199-210). */
scc_status scc_phase_recolour(scc_ctx *ctx);
/* Scorer features (12 floats per vertex, layout in kernels_mlp.hip) of the
 * vertices idx[0..count) for the current words; step is the FB step index. */
scc_status scc_phase_features(scc_ctx *ctx, int32_t step, const int32_t *idx, int64_t count,
                              float *X_out);

/* ---- learned pivot scorer --------------------------------------------------
 * No reference counterpart (the reference's pivot is whichever vertex wins an
 * atomicCAS race, scc_routines.cu:214-221).  An MLP 12 -> 32 -> 32 -> 1 with
 * BF16 weights and activations and FP32 accumulation scores every live
 * vertex before each election; with SCC_PIVOT_LEARNED the highest score of a
 * colour slot wins (ties: lowest id).  The partition does not depend on the
 * scores.  Parameters: 1505 floats, W1[32][12] b1[32] W2[32][32] b2[32]
 * w3[32] b3, held in FP32 (master) and BF16 (used by every kernel). */
#define SCC_MLP_PARAMS 1505
#define SCC_MLP_FEATURES 12
scc_status scc_mlp_set_params(scc_ctx *ctx, const float *params, int32_t count);
scc_status scc_mlp_get_params(scc_ctx *ctx, float *params, int32_t count);
/* logits of N feature rows (host X[N][12] -> host logits[N]) */
scc_status scc_mlp_infer(scc_ctx *ctx, const float *X, int64_t N, float *logits);
/* gradient of the mean binary cross-entropy of sigmoid(logit) against y over
 * the N rows (host in, host grad[1505] out) */
scc_status scc_mlp_gradient(scc_ctx *ctx, const float *X, const float *y, int64_t N, float *grad,
                            float *loss);
/* Adam training on the GPU: `epochs` passes over the N rows in shuffled
 * mini-batches of `batch`; epoch_loss (may be NULL) receives the mean loss of
 * each epoch.  Starts from the current parameters. */
scc_status scc_mlp_train(scc_ctx *ctx, const float *X, const float *y, int64_t N, int32_t epochs,
                         int32_t batch, float lr, uint64_t seed, float *epoch_loss);

/* ---- measurement ----------------------------------------------------------
 * Work counters and kernel timing of the reachability levels, for the roofline
 * report (the reference times phases with gettimeofday + cudaDeviceSynchronize,
 * timing.h:14).  flags: 1 = count expanded vertices / scanned edges / new
 * visits with device counters; 2 = time every level kernel with HIP events on
 * the context's stream; 4 = also log one row per reach or trim level (implies
 * 1 and 2; levels are then launched one at a time, so wall times grow).  Flags
 * 1, 2 and 4 take the waiting path (no speculative steps, no one-workgroup
 * tail); 8 = time the phases of the one-workgroup tail's steps, on the fast
 * path. */
typedef struct {
    int64_t reach_launches;   /* level kernels launched, including ones that exited at once */
    int64_t reach_levels;     /* level kernels that did work */
    int64_t sweep_levels;     /* of those, levels that read every vertex word (q, t, h sweeps) */
    int64_t expanded;         /* vertices whose rows were scanned (flag 1) */
    int64_t edges;            /* arcs scanned (flag 1) */
    int64_t visited;          /* vertices newly reached (flag 1) */
    int64_t queue_levels;     /* levels that expanded a frontier queue (d, l, h) */
    int64_t trim_launches;
    double reach_kernel_ms;   /* sum of level-kernel durations (flag 2) */
    double trim_kernel_ms;    /* sum of trim-sweep durations (flag 2) */
    int64_t tile_runs;        /* tile reach: runs of a tile by a workgroup (flag 1) */
    int64_t tile_rounds;      /* tile reach: claim words of a cell taken and closed in those runs (flag 1) */
    double tile_phase_ms[4];  /* tile reach, summed over runs (flag 1): set-up and seeds, fixpoint,
                                 counting (instrumented runs only), claims across tile boundaries */
    double tile_busy_ms;      /* tile reach, summed over waves (flag 1): time spent on takes */
    int64_t tile_idle_polls;  /* tile reach (flag 1): polls of a wave that found nothing to take */
    double tail_phase_ms[4];  /* one-workgroup tail (flag 1): its steps' reach, trim, update, confirm */
    int64_t tail_local_live;  /* one-workgroup tail (flag 1): live vertices when its LDS copy was built
                                 (0: never), and the step of the tail that built it */
    int64_t tail_local_step;
} scc_profile_data;
scc_status scc_profile_set(scc_ctx *ctx, int flags);
scc_status scc_profile_get(scc_ctx *ctx, scc_profile_data *out, int reset);
/* Runs the calibration kernel k_stream_copy once: reads `bytes` bytes of
 * device memory and writes the same number (for scaling hardware traffic
 * counters against a kernel of known traffic). */
scc_status scc_profile_copy(scc_ctx *ctx, int64_t bytes);
/* The level log of flag 4, SCC_LEVEL_FIELDS doubles per row: kind (0 reach,
 * 1 trim), level, kernel ms, input queue F, input queue B, estimated arcs of
 * queue F, of queue B, pulled directions (bit 0 F, bit 1 B), expanded vertices,
 * scanned arcs, new visits (these three for reach levels), and the longest
 * early-exit row scan of one work-item in arcs (pull levels and trim levels).
 * Flag 4 implies flags 1 and 2.  Copies up to cap rows, sets *rows to the
 * rows logged; reset clears the log. */
#define SCC_LEVEL_FIELDS 12
scc_status scc_profile_levels(scc_ctx *ctx, double *out, int64_t cap, int64_t *rows, int reset);

/* Fault injection for tests: the nth device allocation (1-based) of the next
 * scc_upload reports out of memory, so that upload returns SCC_E_OOM and the
 * context stays usable for another upload; 0 disarms it.  No reference
 * counterpart: the reference exits on any failed allocation
 * (GPUerrchk.h:6-12). */
scc_status scc_test_fail_alloc(scc_ctx *ctx, int32_t nth);

/* ---- a decomposition that allocates nothing ---------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The caller passes a CSR and its transpose already in
 * device memory (0-based, int32, as scc_upload takes them) and a device buffer
 * of n labels; the call writes the canonical labels (minimum vertex id of each
 * SCC) there and the SCC count to *num_scc (host, may be NULL).  It runs the
 * plain FB rule (trim-1, lowest-id pivots, reach, trim, update) in one
 * workgroup on `stream` (a hipStream_t; NULL for the null stream) of the
 * current device and returns after that stream has finished the call.
 * Every buffer besides the caller's lives in one fixed device array of the
 * library (about 13.5 MB, held on each device the entry point has run on until
 * the process ends), so the call makes no allocation; it therefore accepts at
 * most SCC_NOALLOC_MAX_VERTICES vertices (SCC_E_OVERFLOW beyond) and checks
 * the offsets and ids of both arrays (SCC_E_INPUT).  Calls that overlap on one
 * device (other host threads, other streams) share that array and race. */
#define SCC_NOALLOC_MAX_VERTICES 262144
scc_status scc_run_noalloc(const int32_t *d_row, const int32_t *d_col, const int32_t *d_rowT,
                           const int32_t *d_colT, int32_t n, int64_t m, int32_t *d_labels, int64_t *num_scc,
                           void *stream);

/* ---- host input path (no GPU needed) ------------------------------------- */
/* load (load.cpp:18-207, load.h:5): edge list or MatrixMarket file; a line
 * "a b" is the arc b->a; self-loops kept, duplicates dropped.  Entry lines are
 * parsed in parallel.  A file written by scc_save_graph is read as binary CSR
 * (recognised by its magic bytes) and checked, with CSR^T rebuilt. */
scc_status scc_load_file(const char *path, scc_host_graph *g);
/* Writes g (CSR, statistics, batch offsets) as a binary CSR file that
 * scc_load_file reads without parsing text.  No reference counterpart: it
 * serves repeated runs on graphs whose text takes long to parse. */
scc_status scc_save_graph(const scc_host_graph *g, const char *path);
/* Builds CSR + CSR^T from arcs src[i]->dst[i] (0-based): rows sorted,
 * duplicates removed, self-loops dropped when drop_self_loops != 0. */
scc_status scc_csr_from_arcs(int32_t n, int64_t m, const int32_t *src, const int32_t *dst,
                             int drop_self_loops, scc_host_graph *g);
/* R-MAT (a, b, c, d = 1-a-b-c), 2^scale vertices, edge_factor * 2^scale arcs
 * drawn; self-loops and duplicates removed; permute != 0 relabels vertices by
 * a seeded bijection. */
scc_status scc_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                        int permute, scc_host_graph *g);
/* side x side grid, id = r*side + c; each lattice edge is bidirectional with
 * probability p_bidir, otherwise one uniformly random direction. */
scc_status scc_gen_mesh(int side, double p_bidir, uint64_t seed, scc_host_graph *g);
/* num_graphs ER digraphs packed into one CSR with graph_offsets: n_i
 * log-uniform in [n_min, n_max], mean out-degree uniform in [deg_lo, deg_hi],
 * plus planted directed cycles with Geometric(cycle_p) sizes over a
 * cycle_cover fraction of each graph's vertices. */
scc_status scc_gen_batch(int num_graphs, int n_min, int n_max, double deg_lo, double deg_hi,
                         double cycle_p, double cycle_cover, uint64_t seed, scc_host_graph *g);
void scc_graph_free(scc_host_graph *g);

#ifdef __cplusplus
}
#endif

#endif /* GPUSCC_H */
