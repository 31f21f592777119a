// Launchers of the HIP kernels (kernels_fb.hip, kernels_reach.hip), called by
// the host engine (engine.hip).  Host-compilable declarations only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpuscc {

struct Ctl;

// pivot scorer: 12 features -> 32 -> 32 -> 1
constexpr int MLP_IN = 12, MLP_HID = 32;
constexpr int MLP_PARAMS = MLP_HID * MLP_IN + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID + 1;

struct MlpWork {  // per-batch training buffers
    uint16_t *H1, *H2;          // BF16 activations [B][32]
    float *dH1, *dH2, *ds, *Xb;  // gradients of the activations, of the logit; gathered features [B][12]
    float *grad;                 // [MLP_PARAMS]
    float *loss;                 // [1] sum of per-sample losses
};

struct Launch {
    hipStream_t stream;
    uint32_t grid;   // workgroups (persistent grid-stride)
    uint32_t block;  // work-items per workgroup, multiple of 64
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;  // set: time the kernel itself
};

// Arguments of one reachability level (both directions): forward over
// (row, col) with bit F, backward over (rowT, colT) with bit B.  dirs bit 0 /
// bit 1 enables the forward / backward direction.  Queue schedules read
// inq[d] (length ctl->qsize[d][iter & 3]) and append to outq[d].
struct ReachArgs {
    const int32_t *row, *col, *rowT, *colT;
    uint32_t *word;
    int32_t n;
    const int32_t *inq[2];
    int32_t *outq[2];
    const int32_t *hin[2];  // huge rows deferred by the previous level (ctl->hsize[d][iter & 3])
    int32_t *hout[2];       // huge rows this level defers to the next
    Ctl *ctl;
    int iter;
    uint32_t dirs;
    uint32_t pull_thr;  // hybrid: a frontier of at least this many vertices pulls
    unsigned long long pull_arcs;  // hybrid: ... or one with at least this many (estimated) arcs
    int pull_cost;                 // hybrid: 1 = pull when frontier arcs >= 2 x (live - reached) instead;
                                   // 2 = direction-optimising (Beamer): a pushing direction pulls once its
                                   // frontier's arcs exceed (pull_arcs - arcs it pushed so far) / alpha, a
                                   // pulling one keeps pulling while its frontier grows or holds more than
                                   // pull_thr vertices
    uint32_t alpha;                // pull_cost 2: the divisor of the unexplored arcs
    const uint8_t *dcls[2];  // per vertex and direction: bit length of the degree | CLS_FITS | 0x80 (self-loop)
    const int32_t *head[2];  // per vertex and direction: first entry of the row (-1: empty row), read
                             // in place of col[row[v]] by the scans that stop at the first match
    const int4 *ell[2];      // per vertex and direction: the row's first four entries (-1 past its
                             // end); complete when the class byte has CLS_FITS (at most four)
    uint32_t nt;        // work-items per workgroup (set by the launcher; blockDim.x without its
                        // per-use load from the dispatch packet)
    uint32_t huge_thr;  // rows of at least this many arcs are split over all workgroups
    int local_rounds;   // d: rounds of workgroup-local expansion after the level's share
    int local_first;    // first level that may run local rounds
    int one_colour;     // every arc joins two vertices of one colour (the first FB step of a run): the
                        // local rounds claim without reading the target's word first
    int instr;          // count work (expanded vertices, scanned arcs, visits)
    uint32_t *vis[2];   // the first FB step of the hybrid schedule (else null): bit u of vis[d] is set
                        // once u is reached in direction d (seeds by the first election, visits by every
                        // claim), so the pull scans test a bit in an L2-resident bitmap instead of
                        // gathering whole vertex words
    // trim levels
    int32_t *root;      // trimmed vertices become their own SCC
    uint32_t *stamp;    // per-vertex tag of the last trim level that checked it
    uint32_t tag;       // this level's tag (unique per launch)
    int skip_fb;        // right after a complete reach from pivots: F and B without P means an SCC of
                        // two or more vertices, which no trim removes, so those vertices are not checked
    int init_list;      // the first level of trim1: check only the vertices k_init listed in inq[0]
    const int32_t *live;        // trim sweeps: only the vertices live[0 .. *live_cnt) (null: all n)
    const uint32_t *live_cnt;
    const int32_t *oid;         // the caller's ids (roots of trimmed vertices), null: the same
};

// Local rounds start at this level in the first FB step: its first levels are
// thin on every graph (one pivot), and a hub among them would be expanded by a
// single work-item.  Later steps start from many pivots on small colour sets,
// where local rounds from level 0 finish most reaches in one launch.
constexpr int LOCAL_FIRST_LEVEL = 4;

struct Graph {  // device CSR + CSR^T
    int32_t n;
    int64_t m;
    const int32_t *row, *col, *rowT, *colT;
};

// kernels_fb.hip
// colour 1 everywhere, or graph id + 1 when off (G + 1 offsets) is given; with cls
// (degree classes, out [0, n) and in [n, 2n)) the vertices whose out- or in-row
// holds no vertex but themselves start trimmed (the first level of trim1).
// Adds the live and trimmed counts to ctl->live and ctl->pcount (reset bit 1024 zeroes them).
// With hlive (k_head_live) the live vertices whose first out- or in-entry is not known to
// stay live are appended to unk, their count added to ctl->qsize[0][0] (zeroed before):
// the first level of trim1 checks only them.
// vis (may be null): the two visited bitmaps (vis_words words each) are cleared.
// oid (may be null): the caller's id of each vertex (see orig_id); roots are set to it.
void launch_init(const Launch &l, uint32_t *word, int32_t *root, int32_t n, const int32_t *off,
                 int32_t G, const uint8_t *cls, const uint8_t *hlive, int32_t *unk, Ctl *ctl,
                 uint32_t *vis = nullptr, int64_t vis_words = 0, int32_t *mn = nullptr,
                 const int32_t *oid = nullptr);
// ctl->count += arcs that cross graphs of the batch
void launch_batch_check(const Launch &l, const Graph &g, const int32_t *off, int32_t G, Ctl *ctl);
void launch_fill_u64(const Launch &l, unsigned long long *p, int64_t count, unsigned long long v);
void launch_stream_copy(const Launch &l, const void *src, void *dst, int64_t bytes);
void launch_fill_i32(const Launch &l, int32_t *p, int64_t count, int32_t v);
// out[v] = bit length of row[v + 1] - row[v], | 0x80 when the row holds v (a self-loop),
// | CLS_FITS when it has at most four entries; head[v] = col[row[v]] (-1 for an empty
// row); ell[v] = the row's first four entries, -1 past its end
constexpr uint32_t CLS_FITS = 0x40;
void launch_degree_class(const Launch &l, const int32_t *row, const int32_t *col, int32_t n, uint8_t *out,
                         int32_t *head, int4 *ell);
// hlive[v] (and hlive[n + v]) = 1 when the first entry of v's out-row (in-row) is a vertex
// other than v whose own rows hold another vertex each (k_init leaves it live); cls, head:
// out-rows [0, n), in-rows [n, 2n)
void launch_head_live(const Launch &l, const uint8_t *cls, const int32_t *head, int32_t n, uint8_t *hlive);
// control-block resets (bits: k_ctl_reset), after a chain guard when stage > 0 (a chain of
// last + 1 levels launched without a wait: its depth, and abort = stage unless it converged);
// with publish, the block is then copied to that (pinned host) mirror
void launch_ctl_reset(hipStream_t s, Ctl *ctl, int what, int last = 0, int stage = 0, Ctl *publish = nullptr);
// ctl (may be null): FB eliminations leave ctl->live.  lin (may be null): process only the
// vertices lin[0 .. *lin_cnt); lout (may be null): append the vertices live afterwards
void launch_update(const Launch &l, uint32_t *word, int32_t *root, int32_t n,
                   const unsigned long long *prev, unsigned long long *cur, int slot_bits,
                   unsigned long long tag, int recolour, int elect, const uint16_t *score,
                   Ctl *ctl, const int32_t *lin = nullptr, const uint32_t *lin_cnt = nullptr,
                   int32_t *lout = nullptr, uint32_t *lout_cnt = nullptr, int32_t *mn = nullptr,
                   const int32_t *oid = nullptr);
// vis (may be null): the winners are marked in both visited bitmaps (the first election
// of a run: the seeds of the first reach)
void launch_confirm(const Launch &l, uint32_t *word, int32_t n, const unsigned long long *cur,
                    int slot_bits, int32_t *piv, Ctl *ctl, const int32_t *lin = nullptr,
                    const uint32_t *lin_cnt = nullptr, uint32_t *vis = nullptr, int64_t vis_words = 0,
                    const int32_t *oid = nullptr);
void launch_count_state(const Launch &l, const uint32_t *word, int32_t n, int what, Ctl *ctl);
void launch_max_colour(const Launch &l, const uint32_t *word, int32_t n, Ctl *ctl);
// piv (may be null): visit only the pivot list of the last election (ctl->npiv entries)
void launch_unset_pivots(const Launch &l, uint32_t *word, int32_t n, Ctl *ctl, const int32_t *piv = nullptr);
void launch_trim2(const Launch &l, const Graph &g, uint32_t *word, int32_t *root, Ctl *ctl,
                  const int32_t *oid = nullptr);
// the WCC pass over the live list lin[0 .. *lin_cnt) (lin null: all n)
// (labels in the caller's ids: oid and its inverse nid, may be null)
void launch_wcc_init(const Launch &l, int32_t *lab, int32_t n, const int32_t *lin = nullptr,
                     const uint32_t *lin_cnt = nullptr, const int32_t *oid = nullptr);
void launch_wcc_iter(const Launch &l, const Graph &g, const uint32_t *word, int32_t *lab, Ctl *ctl,
                     int iter, const int32_t *lin = nullptr, const uint32_t *lin_cnt = nullptr,
                     const int32_t *nid = nullptr);
void launch_wcc_recolour(const Launch &l, uint32_t *word, const int32_t *lab, int32_t n,
                         const Ctl *ctl, const int32_t *lin = nullptr, const uint32_t *lin_cnt = nullptr);
// root[v] = mn[root[v]] in place (mn kept by the run's updates, see k_update)
void launch_root_to_min(const Launch &l, int32_t *root, const int32_t *mn, const uint32_t *word, int32_t n);
// out[i] = in[idx[i]], i < count
void launch_gather_i32(const Launch &l, const int32_t *in, const int32_t *idx, int32_t *out, int64_t count);
// renumbered CSR: row v becomes row map[v] at nrow[map[v]] in ncol, entries renamed by map
void launch_permute_rows(const Launch &l, const int32_t *row, const int32_t *col, const int32_t *map,
                         const int32_t *nrow, int32_t *ncol, int32_t n);
void launch_collect(const Launch &l, const uint32_t *word, int32_t n, uint32_t bit, int32_t *q,
                    Ctl *ctl);

// kernels_tail.hip -- the remaining FB steps of a run in one workgroup, over the live list
// list[lsel] (ctl->lcnt[lsel]) and the pivot list piv (ctl->npiv) of the last election,
// whose table is slots[prev_tab] with tag elect_tag(epoch).  Stops before max_steps steps or
// an epoch whose tag wraps (ctl->tail[1] = 1); ctl->tail and ctl->tail_pivots report the work.
struct TailArgs {
    const int32_t *row, *col, *rowT, *colT;
    uint32_t *word;
    int32_t *root;
    int32_t n;
    unsigned long long *slots[2];
    int slot_bits, prev_tab;
    uint32_t epoch;
    int32_t *piv;
    int32_t *list[2];
    int lsel;
    int32_t *q[4];  // queues, n entries each
    Ctl *ctl;
    uint32_t max_steps;  // live vertices + 2: bounds the steps, and each step's levels and trim rounds
                         // (0: the device's live count + 2)
    uint32_t gate_live;  // not 0: run only when at most this many vertices are live (else status 3)
    int32_t vbase;       // the queues are slices starting at this vertex (a batch graph's first), else 0
    int xbits;           // the run's schedule marks expanded vertices (q, t: FExt / BExt)
    int32_t *mn;         // may be null: FB-eliminated vertices lower mn[their pivot] (see k_update)
    int first_step;      // no pivots yet: the first step is trim-1 and the first election (lowest live id)
    int self_refill;     // before an election whose 16-bit tag would wrap, clear its slot table here
                         // (else hand back to the host, status 1)
    const int32_t *oid;  // the caller's ids (keys, roots), null: the same
    int instr;           // time the steps' phases (ctl->st_tailclk) and note when the LDS copy was built
};
void launch_fb_tail(hipStream_t s, const TailArgs &t);
// a batch in one persistent launch (k_fb_batch): graph g owns vertices [off[g], off[g + 1]);
// workgroups take graphs in the order `order` (largest first) from ctl->wnext; graph g's
// two election tables are tab[tab_off[g] ...] (2 * batch_table_size(n_g) words each)
struct BatchArgs {
    int32_t G;
    const int32_t *off;
    const int32_t *order;
    uint32_t *tab;
    const uint64_t *tab_off;
};
uint32_t batch_table_words(int32_t n_g);  // 4 * batch_table_size(n_g)
void launch_fb_batch(hipStream_t s, uint32_t grid, const TailArgs &t, const BatchArgs &b);

// kernels_tile.hip -- the reach of a first FB step over tiles (regular graphs renumbered tile by
// tile at upload).  A tile (one workgroup) is made of cells of 64 vertices (one wave lane each).
// Iteration 0 computes the cell closures and flags the tiles of the pivots (seed list,
// ctl->qsize[0][0]); iteration i >= 1 runs every flagged tile to its fixpoint and claims across
// tile boundaries, flagging the tiles it claims into.
constexpr uint32_t TILE_MAX = 4096;  // vertices per tile: 64 cells of 64
constexpr int32_t TILE_DEPTH = 256;  // BFS depth from vertex 0 above which the upload tiles by default
constexpr uint32_t CELL = 64;        // vertices per cell (the wavefront width)
constexpr int CELL_XARCS = 4;        // arcs to other cells of the tile kept per vertex and direction
struct TileArgs {
    const int32_t *row, *col, *rowT, *colT;
    uint32_t *word;
    int32_t n, ntiles;
    const int32_t *toff;      // tile t = internal vertices [toff[t], toff[t + 1]); cell k of it starts at toff[t] + 64k
    const unsigned long long *adj[2];  // per vertex and direction (0 out, 1 in): its neighbours in its cell, bit = lane
    unsigned long long *ct[2];         // written by iteration 0, per vertex: the vertices of its cell it is reached
                                       // from inside the cell ([0] forward: that reach it, [1] backward: that it
                                       // reaches), itself included
    const unsigned long long *xarc[2]; // up to 4 neighbours in other cells of its tile, 16 bits each (index from
                                       // the tile's first vertex; 0xFFFF none)
    const int2 *xout[2];      // per vertex and direction: up to two neighbours in other tiles (-1 none)
    const uint8_t *xflag;     // bit d: an arc of direction d leaves the tile; bit 2 + d: more than 4 arcs of
                              // direction d go to other cells of the tile; bit 4 + d: more than 2 leave the
                              // tile (for those, the row is scanned)
    int32_t tsize;            // vertices of every tile but the last: tile t starts at t * tsize
    uint8_t *flags;           // [3][ntiles]: iteration i runs the tiles flagged in [i % 3], flags tiles for
                              // iteration i + 1 in [(i + 1) % 3] and clears [(i + 2) % 3]
    unsigned long long *done; // [ntiles][128]: per cell and direction (claim-word order), the vertices the
                              // tile has expanded in this reach
    const int32_t *piv;       // iteration 0: the pivot list
    Ctl *ctl;
    int iter;
    int instr;                // count work: reached vertices and their arcs, tile runs and rounds
};
void launch_tile_reach(hipStream_t s, uint32_t grid, const TileArgs &a);

// Each source file's kernels are loaded onto a device (its code object) the first
// time one of them is launched there, which the first run of a process would pay
// (R-MAT scale 24: 2.54 ms for the first run after the upload, 1.91 ms with the
// loads done ahead); each of these loads one file's kernels on the current device
// (scc_create, once per device).
hipError_t preload_fb();
hipError_t preload_reach();
hipError_t preload_mlp();
hipError_t preload_tail();
hipError_t preload_tile();
hipError_t preload_noalloc();

// kernels_mlp.hip -- learned pivot scorer
void launch_static_features(const Launch &l, const Graph &g, float4 *out);
// live vertices per colour slot over the list lin[0 .. *lin_cnt) (lin null: all n)
void launch_colour_count(const Launch &l, const uint32_t *word, int32_t n, uint32_t *cnt, int slot_bits,
                         const int32_t *lin = nullptr, const uint32_t *lin_cnt = nullptr, const Ctl *ctl = nullptr);
// features (feat_out, n x 12), and with params the MFMA scorer: scores (score_out) and the
// learned election (keys offered to cur), over the list lin (null: all n); each output optional;
// static_deg: the first election of a run (within-colour degrees = degrees); one_set (may be
// null): every live vertex is in one colour set of *one_set vertices (cnt unused)
void launch_score(const Launch &l, const Graph &g, const uint32_t *word, const float4 *stat, const uint32_t *cnt,
                  const uint32_t *one_set, int slot_bits, float step_f, int static_deg, const int32_t *lin, const uint32_t *lin_cnt,
                  const uint16_t *params, uint16_t *score_out, float *feat_out, unsigned long long *cur,
                  unsigned long long tag, const Ctl *ctl, const int32_t *oid = nullptr);
void launch_mlp_infer(const Launch &l, const float *X, int64_t N, const uint16_t *params, float *logit);
void launch_mlp_fwdbwd(const Launch &l, const float *X, const float *y, const int32_t *idx, int64_t N,
                       const uint16_t *params, MlpWork &w);
void launch_mlp_grad(hipStream_t s, int64_t N, MlpWork &w);
void launch_adam(hipStream_t s, float *wf, const float *g, float *m, float *v, uint16_t *wb, float lr,
                 float beta1, float beta2, float eps, int t);
void launch_to_bf16(hipStream_t s, const float *w, uint16_t *wb);
void launch_fill_f32(const Launch &l, float *p, int64_t count, float v);
void launch_gather_rows(const Launch &l, const float *feat, const int32_t *idx, int64_t count, float *out);

// kernels_reach.hip -- data-driven trim: level 0 sweeps every live vertex and
// queues what it trims (outq[0]); level i >= 1 re-checks only the neighbours
// of the vertices level i-1 trimmed (inq[0])
void launch_trim_level(const Launch &l, const ReachArgs &a);

// kernels_reach.hip -- one level of a reachability chain; sched 0..4 = q t d l h
void launch_reach(const Launch &l, int sched, const ReachArgs &a);
// clears bits of every vertex word
void launch_clear_bits(const Launch &l, uint32_t *word, int32_t n, uint32_t bits);

}  // namespace gpuscc
