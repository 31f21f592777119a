// Host engine of the FB-SCC library: context, upload, the FB step loop and the
// phase-level entry points of include/gpuscc.h.  It restates computeSCC
// (scc.cu:150-1022) with one change of mechanism: the reference sets a flag,
// launches a level and copies the flag back for every BFS level, trim sweep
// and WCC sweep (reach_q.cu:98-109, scc_routines.cu:597-606, wcc.cu:166-171);
// here levels are chained on the device -- a level whose predecessor found
// nothing exits at once -- and the host launches them in growing batches,
// reading the control block back once per batch.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fb_device.hpp"
#include "gpuscc.h"
#include "host/graph.hpp"
#include "kernels.hpp"

using namespace gpuscc;

struct scc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    scc_opts opts{};
    std::string err;
    int cu_count = 256;
    char arch[64] = "unknown";
    Graph g{};
    int32_t *d_row = nullptr, *d_col = nullptr, *d_rowT = nullptr, *d_colT = nullptr;
    uint32_t *d_word = nullptr;
    int32_t *d_root = nullptr, *d_label = nullptr, *d_tmp = nullptr;
    int32_t *d_mn = nullptr;  // learned runs: minimum id of each pivot's SCC (k_update)
    // renumbering (opts.relabel): d_oid[v] = the caller's id of internal vertex v, d_nid the
    // inverse; both null when the upload kept the caller's numbering
    int32_t *d_oid = nullptr, *d_nid = nullptr;
    // tiles (opts.tile_reach; regular graphs): the upload numbered the vertices tile by tile;
    // the first FB step's reach runs over them (kernels_tile.hip)
    int32_t ntiles = 0;
    int32_t *d_toff = nullptr;
    uint8_t *d_tflags = nullptr;            // [3][ntiles]: tiles flagged for the coming iterations
    unsigned long long *d_tdone = nullptr;  // [ntiles][128]: expanded vertices per cell and direction
    unsigned long long *d_tadj = nullptr;   // [2n]: neighbours in the vertex's cell (out, then in)
    unsigned long long *d_txarc = nullptr;  // [2n]: up to four neighbours in other cells of its tile
    unsigned long long *d_tct = nullptr;    // [2n]: cell closures, written by each tile reach
    int2 *d_txout = nullptr;                // [2n]: up to two neighbours outside the vertex's tile
    int32_t tile_size = 0;                  // vertices of every tile but the last
    uint8_t *d_xflag = nullptr;

    bool track_mn = false;    // this run's updates keep d_mn
    int32_t *d_q[4] = {nullptr, nullptr, nullptr, nullptr};  // [direction * 2 + parity]
    int32_t *d_piv = nullptr;    // pivot list (also the forward seeds)
    int32_t *d_seedB = nullptr;  // backward seeds of a phase-level reach
    int32_t *d_huge[4] = {nullptr, nullptr, nullptr, nullptr};  // deferred huge rows [dir * 2 + parity]
    uint32_t *d_stamp = nullptr;  // trim levels: last level that checked each vertex
    uint8_t *d_dcls = nullptr;    // degree classes: out-degrees [0, n), in-degrees [n, 2n)
    int32_t *d_head = nullptr;    // first entry of each row: out-rows [0, n), in-rows [n, 2n)
    int4 *d_ell = nullptr;        // first four entries of each row, same halves
    uint8_t *d_hlive = nullptr;   // first entry of each row left live by k_init, same halves
    int32_t *d_live[2] = {nullptr, nullptr};  // live-vertex lists (written by elections, ctl->lcnt)
    bool use_lists = false;       // scc_run with the plain rule: elections keep a live-vertex list
    int lsel = -1, lsel_prev = -1;  // list holding the live vertices (-1: none yet), before the last election
    uint32_t trim_tag = 0;
    unsigned long long *d_slots[2] = {nullptr, nullptr};
    int slot_bits = 10;
    int prev_tab = 0;        // table holding the last election's winners
    uint32_t epoch = 0;
    Ctl *d_ctl = nullptr;
    Ctl *h_ctl = nullptr;    // pinned mirror
    uint16_t *d_score = nullptr;  // learned-rule scores (per vertex), may be null
    int32_t *d_goff = nullptr;    // batch: graph offsets [num_graphs + 1]
    int32_t num_graphs = 0;
    int32_t *d_border = nullptr;      // batch: graphs by decreasing size (the persistent solver's order)
    uint64_t *d_btab_off = nullptr;   // batch: each graph's election tables in d_btab (words)
    uint32_t *d_btab = nullptr;
    // learned pivot scorer
    float4 *d_stat = nullptr;     // static features per vertex (computed once per upload)
    uint32_t *d_cnt = nullptr;    // live vertices per colour slot
    float *d_feat = nullptr;      // features of every vertex (phase-level feature dumps)
    float *d_wf = nullptr, *d_m = nullptr, *d_v = nullptr;  // FP32 master weights, Adam moments
    uint16_t *d_wb = nullptr;     // BF16 weights used by every kernel
    bool have_params = false;
    bool stat_valid = false;
    int adam_t = 0;
    int cur_step = 0;
    bool one_colour = false;  // scc_run's first FB step: every arc joins two vertices of one colour
    bool best_step = false;   // BEST's first FB step: direction-optimising reach (alpha 15, beta 18)
    uint32_t *d_vis = nullptr;  // visited bitmaps of the first step's reach, F then B (vis_words each)
    int64_t vis_words = 0;
    bool vis_mark = false;   // the next confirm marks its winners in the bitmaps (a run's first election)
    bool vis_valid = false;  // this run cleared the bitmaps and marked its first pivots
    bool uploaded = false, labels_valid = false;
    bool label_in_root = false;  // the last run's labels are the root array itself (plain rule)
    int guess[4] = {4, 4, 2, 4};  // chain batch guesses: fwd, bwd, trim after a reach, first trim
    int guess_wcc = 4;            // the WCC pass's label chain (kept across runs)
    int64_t syncs = 0;
    int fail_alloc_at = 0;   // fault injection: the nth device allocation of the next upload fails
    int alloc_count = 0;
    // control-block work queued for the next k_ctl_reset launch (see ctl_op)
    int pend_bits = 0, pend_last = 0, pend_stage = 0;
    int instr = 0;                  // profile flag 1: device work counters
    int prof_flags = 0;
    scc_profile_data prof{};
    std::vector<hipEvent_t> evs;    // flag 2: one start/stop pair per level launch
    std::vector<double> level_log;  // flag 4: SCC_LEVEL_FIELDS per level
    // speculative steps: reach and trim chain lengths of step k in the last run
    std::vector<int> spec_reach, spec_trim;
    hipEvent_t spec_ev[6] = {};  // a speculative step's phase bounds; [4], [5]: trim-1's chain
    // reach chains captured as HIP graphs: key (schedule, levels, thresholds) -> executable
    struct GraphKey {
        int sched, levels;
        uint32_t pull_thr, huge_thr;
        unsigned long long pull_arcs;
        int local_rounds, local_first, one_colour;
        bool operator==(const GraphKey &o) const {
            return sched == o.sched && levels == o.levels && pull_thr == o.pull_thr && huge_thr == o.huge_thr &&
                   pull_arcs == o.pull_arcs && local_rounds == o.local_rounds && local_first == o.local_first &&
                   one_colour == o.one_colour;
        }
    };
    std::vector<std::pair<GraphKey, hipGraphExec_t>> reach_graphs;
};

namespace {

// live vertices at or below which the remaining FB steps run in one workgroup (k_fb_tail)
constexpr uint32_t kTailLive = 1u << 16;

scc_status fail(scc_ctx *c, scc_status s, const std::string &what) {
    if (c) c->err = what;
    return s;
}

scc_status hip_fail(scc_ctx *c, hipError_t e, const char *what) {
    scc_status s = (e == hipErrorOutOfMemory) ? SCC_E_OOM : SCC_E_HIP;
    return fail(c, s, std::string(what) + ": " + hipGetErrorString(e));
}

// A device allocation of scc_upload; with fault injection armed (scc_test_fail_alloc)
// the chosen one reports out of memory as hipMalloc would.
template <class T>
hipError_t dev_alloc(scc_ctx *c, T **p, size_t bytes) {
    if (c->fail_alloc_at > 0 && ++c->alloc_count == c->fail_alloc_at) {
        c->fail_alloc_at = 0;
        return hipErrorOutOfMemory;
    }
    return hipMalloc((void **)p, bytes);
}

#define HIPCHK(ctx, x)                                  \
    do {                                                \
        hipError_t e_ = (x);                            \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #x); \
    } while (0)

#define CHK(x)                          \
    do {                                \
        scc_status s_ = (x);            \
        if (s_ != SCC_OK) return s_;    \
    } while (0)

Launch launch_for(const scc_ctx *c) {
    uint32_t block = (uint32_t)(c->opts.block_size > 0 ? c->opts.block_size : 256);
    block = std::max<uint32_t>(64, std::min<uint32_t>(1024, block & ~63u));
    uint32_t grid = c->opts.grid_blocks > 0 ? (uint32_t)c->opts.grid_blocks
                                            : (uint32_t)c->cu_count * 4u;
    uint64_t need = ((uint64_t)std::max<int32_t>(c->g.n, 1) + block - 1) / block;
    if (grid > need) grid = (uint32_t)std::max<uint64_t>(need, 1);
    grid = (grid + 7u) & ~7u;
    return Launch{c->stream, grid, block};
}

Launch fixed_launch(const scc_ctx *c) {  // level kernels: LDS sized for up to 512 work-items
    Launch l = launch_for(c);
    l.block = c->opts.reach_block == 512 ? 512 : 256;
    uint64_t need = ((uint64_t)std::max<int32_t>(c->g.n, 1) + 255) / 256;
    uint32_t grid = c->opts.grid_blocks > 0 ? (uint32_t)c->opts.grid_blocks : (uint32_t)c->cu_count * 4u;
    if (grid > need) grid = (uint32_t)std::max<uint64_t>(need, 1);
    l.grid = (grid + 7u) & ~7u;
    return l;
}

// Control-block operations are batched: reset bits and a chain guard queued
// here go out with the next ctl_op, as one single-work-item kernel instead of
// several.  read_ctl sends a queued guard (and any reset but bit 64) first;
// bit 64 (seed lists from the pivot count) may stay queued across a read, as
// only the next chain start (bit 2, applied after it) reads the seeds.
void ctl_op(scc_ctx *c, int bits) {
    launch_ctl_reset(c->stream, c->d_ctl, c->pend_bits | bits, c->pend_last, c->pend_stage);
    c->pend_bits = 0;
    c->pend_stage = 0;
}
void ctl_queue(scc_ctx *c, int bits) { c->pend_bits |= bits; }
// the guard of a chain of last + 1 levels launched without a wait (stage 1, 2)
void ctl_guard(scc_ctx *c, int last, int stage) {
    if (c->pend_stage) ctl_op(c, 0);
    c->pend_last = last;
    c->pend_stage = stage;
}

// Before a call that writes the root array: labels held there move to the label buffer.
scc_status detach_labels(scc_ctx *c) {
    if (c->labels_valid && c->label_in_root && c->g.n > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_label, c->d_root, sizeof(int32_t) * (size_t)c->g.n, hipMemcpyDeviceToDevice,
                                 c->stream));
    c->label_in_root = false;
    return SCC_OK;
}

// The device copies the control block into the pinned mirror with the queued
// guard and resets (a queued bit 64 alone stays queued for the next reset), then
// the host waits.
scc_status read_ctl(scc_ctx *c) {
    if (c->pend_stage || (c->pend_bits & ~64)) {
        launch_ctl_reset(c->stream, c->d_ctl, c->pend_bits, c->pend_last, c->pend_stage, c->h_ctl);
        c->pend_bits = 0;
        c->pend_stage = 0;
    } else {
        launch_ctl_reset(c->stream, c->d_ctl, 0, 0, 0, c->h_ctl);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->syncs++;
    return SCC_OK;
}

scc_status ensure_events(scc_ctx *c, size_t count) {
    while (c->evs.size() < count) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->evs.push_back(e);
    }
    return SCC_OK;
}

// Runs a chain of level kernels to its fixpoint; *levels = kernels that did work.
// launch_level(iter, ev_start, ev_stop) launches level `iter`; with kernel
// timing on (profile flag 2) the events bracket the kernel itself.
// kind 0 = reachability, 1 = trim, 2 = other (for the profile counters).
// The chain starts with control-block reset reset_mask | 1, or as the caller
// left it when reset_mask < 0; resume > 0 continues a chain from that level.
template <class F>
scc_status run_chain(scc_ctx *c, F &&launch_level, int &guess, int reset_mask, int64_t *levels,
                     int kind = 2, int64_t resume = 0) {
    if (!resume && reset_mask >= 0) ctl_op(c, reset_mask | 1);
    int64_t launched = resume;
    const bool logged = (c->prof_flags & 4) && kind < 2;
    int L = logged ? 1 : std::max(1, std::min(guess, 64));
    const int64_t limit = (int64_t)c->g.n + 4;
    const bool timed = (c->prof_flags & 2) && kind < 2;
    unsigned long long last[3] = {0, 0, 0};
    if (logged) {  // the counters are cumulative over this chain
        CHK(read_ctl(c));
        last[0] = c->h_ctl->st_expanded;
        last[1] = c->h_ctl->st_edges;
        last[2] = c->h_ctl->st_visited;
    }
    for (;;) {
        if (timed) CHK(ensure_events(c, 2 * (size_t)L));
        for (int i = 0; i < L; i++) {
            if (timed) launch_level((int)(launched + i), c->evs[2 * i], c->evs[2 * i + 1]);
            else launch_level((int)(launched + i), (hipEvent_t) nullptr, (hipEvent_t) nullptr);
        }
        launched += L;
        CHK(read_ctl(c));
        if (timed) {
            double ms = 0;
            for (int i = 0; i < L; i++) {
                float t = 0;
                HIPCHK(c, hipEventElapsedTime(&t, c->evs[2 * i], c->evs[2 * i + 1]));
                ms += t;
            }
            if (kind == 0) c->prof.reach_kernel_ms += ms; else c->prof.trim_kernel_ms += ms;
        }
        if (kind == 0) c->prof.reach_launches += L;
        if (kind == 1) c->prof.trim_launches += L;
        uint32_t ran = c->h_ctl->ran;
        if (logged && (int64_t)ran >= launched) {
            const Ctl &h = *c->h_ctl;
            const int i = (int)((launched - 1) & 3);
            float t = 0;
            HIPCHK(c, hipEventElapsedTime(&t, c->evs[0], c->evs[1]));
            // a tile-reach iteration logs its tiles run and the most rounds in place of the
            // queues, and its longest tile run (microseconds) in place of the longest scan
            const bool tl = kind == 0 && h.tlog[i][0] > 0;
            double row[SCC_LEVEL_FIELDS] = {(double)kind, (double)(launched - 1), (double)t,
                                            (double)(tl ? h.tlog[i][0] : h.qsize[0][i]),
                                            (double)(tl ? h.tlog[i][1] : h.qsize[1][i]),
                                            (double)h.qarcs[0][i], (double)h.qarcs[1][i],
                                            (double)(kind == 0 ? h.pull_mask : 0),
                                            (double)(h.st_expanded - last[0]), (double)(h.st_edges - last[1]),
                                            (double)(h.st_visited - last[2]),
                                            tl ? h.tlog[i][2] * 0.01 : (double)h.st_maxscan[i]};
            last[0] = h.st_expanded;
            last[1] = h.st_edges;
            last[2] = h.st_visited;
            c->level_log.insert(c->level_log.end(), row, row + SCC_LEVEL_FIELDS);
        }
        if ((int64_t)ran < launched || c->h_ctl->changed[(launched - 1) & 3] == 0) {
            *levels = ran;
            break;
        }
        if (launched > limit) return fail(c, SCC_E_STATE, "level chain did not converge");
        L = logged ? 1 : std::min(L * 2, 1024);
    }
    guess = (int)std::max<int64_t>(1, *levels);
    return SCC_OK;
}

void clear_graphs(scc_ctx *c) {
    for (auto &kv : c->reach_graphs) (void)hipGraphExecDestroy(kv.second);
    c->reach_graphs.clear();
}

// Level launchers of the two chains (one object per chain run).
struct TrimChain {
    scc_ctx *c;
    Launch l;
    ReachArgs a{};
    bool after_init = false;  // the chain of trim1 right after k_init: level 0 checks k_init's list
    explicit TrimChain(scc_ctx *cc, bool skip_fb = false) : c(cc), l(fixed_launch(cc)) {
        a.row = c->g.row;
        a.col = c->g.col;
        a.rowT = c->g.rowT;
        a.colT = c->g.colT;
        a.word = c->d_word;
        a.n = c->g.n;
        a.ctl = c->d_ctl;
        a.root = c->d_root;
        a.oid = c->d_oid;
        a.stamp = c->d_stamp;
        a.skip_fb = skip_fb ? 1 : 0;
        a.instr = c->instr;
        a.head[0] = c->d_head;
        a.head[1] = c->d_head + c->g.n;
        a.ell[0] = c->d_ell;
        a.ell[1] = c->d_ell + c->g.n;
        a.dirs = 1;
    }
    void operator()(int it, hipEvent_t e0, hipEvent_t e1) {
        if (++c->trim_tag == 0) {  // tag wrap: forget every stamp
            launch_fill_i32(launch_for(c), (int32_t *)c->d_stamp, c->g.n, 0);
            c->trim_tag = 1;
        }
        a.iter = it;
        a.tag = c->trim_tag;
        a.init_list = after_init && it == 0 ? 1 : 0;
        // sweeps go over the live-vertex list of the last election, when there is one
        a.live = c->use_lists && c->lsel >= 0 ? c->d_live[c->lsel] : nullptr;
        a.live_cnt = a.live ? &c->d_ctl->lcnt[c->lsel] : nullptr;
        a.inq[0] = c->d_q[(it - 1) & 1];
        a.outq[0] = c->d_q[it & 1];
        a.inq[1] = a.inq[0];
        a.outq[1] = c->d_q[2 + (it & 1)];
        Launch lt = l;
        lt.ev_start = e0;
        lt.ev_stop = e1;
        launch_trim_level(lt, a);
    }
};

// One reachability chain for the directions in `dirs` (1 = F, 2 = B, 3 = both).
// Queue schedules start from the seed lists: d_piv (forward, ctl->seeds[0])
// and bseed (backward, ctl->seeds[1]).
struct ReachChain {
    scc_ctx *c;
    int sched;
    const int32_t *bseed;
    Launch l;
    ReachArgs a{};
    bool tiles = false;  // the first FB step of a tiled single graph: the tile reach (kernels_tile.hip)
    TileArgs ta{};
    ReachChain(scc_ctx *cc, uint32_t dirs, int sch, const int32_t *bs)
        : c(cc), sched(sch), bseed(bs), l(sch == SCC_SCHED_Q ? launch_for(cc) : fixed_launch(cc)) {
        a.row = c->g.row;
        a.col = c->g.col;
        a.rowT = c->g.rowT;
        a.colT = c->g.colT;
        a.word = c->d_word;
        a.n = c->g.n;
        a.ctl = c->d_ctl;
        a.dirs = dirs;
        a.instr = c->instr;
        a.pull_thr = c->opts.hybrid_threshold > 0 ? (uint32_t)c->opts.hybrid_threshold
                                                  : (uint32_t)std::max<int64_t>(1024, c->g.n / 20);
        a.huge_thr = c->opts.huge_row > 0 ? (uint32_t)c->opts.huge_row
                                          : (sched == SCC_SCHED_Q || sched == SCC_SCHED_D ? 1024u : 4096u);
        a.local_rounds = c->opts.local_rounds < 0 ? 0 : (c->opts.local_rounds > 0 ? c->opts.local_rounds : 64);
        a.local_first = c->cur_step > 0 ? 0 : LOCAL_FIRST_LEVEL;  // cur_step: elections before this reach
        a.one_colour = c->one_colour ? 1 : 0;
        const int div = c->opts.hybrid_arc_div > 0 ? c->opts.hybrid_arc_div : 14;
        a.pull_arcs = c->opts.hybrid_arc_div < 0 ? ~0ull
                                                 : std::max<unsigned long long>(1, (unsigned long long)c->g.m / div);
        a.dcls[0] = c->d_dcls;
        a.dcls[1] = c->d_dcls + c->g.n;
        a.head[0] = c->d_head;
        a.head[1] = c->d_head + c->g.n;
        a.ell[0] = c->d_ell;
        a.ell[1] = c->d_ell + c->g.n;
        a.pull_cost = c->opts.hybrid_policy == 2 ? 1 : 0;
        if (c->best_step) {  // reach_b.cu:540,645: alpha = 15 against the unexplored arcs, beta = 18
            a.pull_cost = 2;
            a.alpha = 15;
            a.pull_arcs = (unsigned long long)c->g.m;
            a.pull_thr = (uint32_t)std::max<int64_t>(1, c->g.n / 18);
        }
        a.vis[0] = a.vis[1] = nullptr;
        // the pulls test bitmap bits: batches only -- their pull levels scan many rows that
        // find nothing (4,096 graphs at different depths), and the bitmap gathers stay in L2
        // (batch 8.87 -> 8.62 ms); on single graphs the marks cost more than the pulls save
        // (R-MAT s22 0.486 -> 0.51 ms, mesh 6.4 -> 6.9 ms)
        if (sched == SCC_SCHED_H && c->one_colour && c->vis_valid && c->num_graphs > 0) {
            a.vis[0] = c->d_vis;
            a.vis[1] = c->d_vis + c->vis_words;
        }
        tiles = c->ntiles > 0 && sched == SCC_SCHED_H && dirs == 3 && c->one_colour && c->num_graphs == 0 &&
                !c->best_step && c->opts.hybrid_policy == 0;
        if (tiles) {
            const int32_t n = c->g.n;
            ta.row = c->g.row;
            ta.col = c->g.col;
            ta.rowT = c->g.rowT;
            ta.colT = c->g.colT;
            ta.word = c->d_word;
            ta.n = n;
            ta.ntiles = c->ntiles;
            ta.toff = c->d_toff;
            ta.adj[0] = c->d_tadj;
            ta.adj[1] = c->d_tadj + n;
            ta.xarc[0] = c->d_txarc;
            ta.xarc[1] = c->d_txarc + n;
            ta.ct[0] = c->d_tct;
            ta.ct[1] = c->d_tct + n;
            ta.xout[0] = c->d_txout;
            ta.xout[1] = c->d_txout + n;
            ta.tsize = c->tile_size;
            ta.xflag = c->d_xflag;
            ta.flags = c->d_tflags;
            ta.done = c->d_tdone;
            ta.piv = c->d_piv;
            ta.ctl = c->d_ctl;
            ta.instr = c->instr;
        }
    }
    int reset_mask() const { return sched >= SCC_SCHED_D ? 2 : 0; }
    void operator()(int it, hipEvent_t e0, hipEvent_t e1) {
        a.iter = it;
        for (int d = 0; d < 2; d++) {
            a.inq[d] = it == 0 ? (d == 0 ? c->d_piv : bseed) : c->d_q[d * 2 + ((it - 1) & 1)];
            a.outq[d] = c->d_q[d * 2 + (it & 1)];
            a.hin[d] = c->d_huge[d * 2 + ((it + 1) & 1)];
            a.hout[d] = c->d_huge[d * 2 + (it & 1)];
        }
        Launch lt = l;
        lt.ev_start = e0;
        lt.ev_stop = e1;
        if (tiles) {
            if (it == 0)  // no tile flagged (each iteration clears the flags it ran)
                launch_fill_i32(launch_for(c), (int32_t *)c->d_tflags, (3 * (int64_t)c->ntiles + 3) / 4, 0);
            ta.iter = it;
            if (e0) (void)hipEventRecord(e0, c->stream);
            launch_tile_reach(c->stream, (uint32_t)std::max(c->cu_count, 64), ta);
            if (e1) (void)hipEventRecord(e1, c->stream);
            return;
        }
        launch_reach(lt, sched, a);
    }
    void account(int64_t depth) {  // profile counters of a finished chain
        c->prof.reach_levels += depth;
        int64_t sw = 0;
        if (sched == SCC_SCHED_Q || sched == SCC_SCHED_T) sw = depth;
        else if (sched == SCC_SCHED_H) sw = c->h_ctl->pull_levels;
        c->prof.sweep_levels += sw;
        if (sched >= SCC_SCHED_D) c->prof.queue_levels += depth;
        if (c->instr) {
            c->prof.expanded += (int64_t)c->h_ctl->st_expanded;
            c->prof.edges += (int64_t)c->h_ctl->st_edges;
            c->prof.visited += (int64_t)c->h_ctl->st_visited;
            c->prof.tile_runs += (int64_t)c->h_ctl->st_truns;
            c->prof.tile_rounds += (int64_t)c->h_ctl->st_trounds;
            for (int q = 0; q < 4; q++) c->prof.tile_phase_ms[q] += (double)c->h_ctl->st_tclk[q] * 1e-5;
            c->prof.tile_busy_ms += (double)c->h_ctl->st_tbusy * 1e-5;
            c->prof.tile_idle_polls += (int64_t)c->h_ctl->st_tidle;
        }
    }
};

// Levels 0 .. L-1 of a fused reach chain from the pivot list.  Their
// arguments depend only on the schedule and the level (buffers are fixed per
// upload, options per set_opts), so the L launches are captured once as a HIP
// graph and replayed as one launch.  The chain holds no memset nodes.
scc_status launch_reach_levels(scc_ctx *c, ReachChain &rc, int L) {
    const scc_ctx::GraphKey key{rc.sched, L, rc.a.pull_thr, rc.a.huge_thr, rc.a.pull_arcs, rc.a.local_rounds,
                                rc.a.local_first, rc.a.one_colour};
    for (auto &kv : c->reach_graphs)
        if (kv.first == key) {
            HIPCHK(c, hipGraphLaunch(kv.second, c->stream));
            return SCC_OK;
        }
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < L; i++) rc(i, nullptr, nullptr);
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
    if (c->reach_graphs.size() >= 64) clear_graphs(c);  // bounded cache
    c->reach_graphs.emplace_back(key, exec);
    HIPCHK(c, hipGraphLaunch(exec, c->stream));
    return SCC_OK;
}

// resume > 0: continue a chain whose first `resume` levels already ran;
// after_reach: the trim of an FB step (the reach from the pivots is complete)
// first: trim1 after k_init (its guess slot; k_init's reset started the chain);
// after_reach: right after a complete reach.  Level 0 sweeps: it reads no
// input queue, so the seed lists stay as they are.
// first: trim1 right after k_init (states exactly as k_init left them)
scc_status do_trim(scc_ctx *c, int64_t *iters, bool first = false, int64_t resume = 0, bool after_reach = false) {
    TrimChain tc(c, after_reach);
    tc.after_init = first;
    return run_chain(c, tc, c->guess[first ? 3 : 2], first ? -1 : 2, iters, 1, resume);
}

scc_status do_reach(scc_ctx *c, uint32_t dirs, int sched, const int32_t *bseed, int64_t *depth,
                    int64_t resume = 0) {
    ReachChain rc(c, dirs, sched, bseed);
    if (c->instr && !resume) ctl_queue(c, 32);
    CHK(run_chain(c, rc, c->guess[dirs == 2 ? 1 : 0], rc.reset_mask(), depth, 0, resume));
    rc.account(*depth);
    return SCC_OK;
}

// per-vertex features of every vertex (and, with want_score, MLP scores) for the current words;
// the first election of a run (cur_step 0) takes the degrees as within-colour degrees
scc_status do_features(scc_ctx *c, bool want_score, float *feat_out) {
    Launch l = launch_for(c);
    launch_fill_i32(l, (int32_t *)c->d_cnt, (int64_t)1 << c->slot_bits, 0);
    launch_colour_count(l, c->d_word, c->g.n, c->d_cnt, c->slot_bits);
    float step_f = 0.25f * std::log1p((float)c->cur_step);
    launch_score(l, c->g, c->d_word, c->d_stat, c->d_cnt, nullptr, c->slot_bits, step_f, c->cur_step == 0 ? 1 : 0, nullptr,
                 nullptr, want_score ? c->d_wb : nullptr, want_score ? c->d_score : nullptr, feat_out, nullptr, 0,
                 nullptr);
    HIPCHK(c, hipGetLastError());
    return SCC_OK;
}

// one election into the table that is not `prev`; winners -> state 19 + pivot list.
// learned: the colour split (and the live list) first, then the colour-set
// sizes and the MFMA scorer over the live vertices, which offers the keys with
// the scores in them.  sync = false leaves *npiv alone: the count is then read
// from the control block with the next chain's results.
scc_status do_elect(scc_ctx *c, int recolour, int64_t *npiv, bool learned = false, bool sync = true) {
    Launch l = launch_for(c);
    c->epoch++;
    unsigned long long *prev = c->d_slots[c->prev_tab], *cur = c->d_slots[c->prev_tab ^ 1];
    int64_t nslots = (int64_t)1 << c->slot_bits;
    if ((c->epoch & 0xFFFFu) < 2u) launch_fill_u64(l, cur, nslots, ~0ull);  // tag wrap
    const bool lists = c->use_lists;
    const int lin_sel = c->lsel, lout_sel = lin_sel < 0 ? 0 : lin_sel ^ 1;
    ctl_op(c, 4 | (lists ? (lout_sel ? 4096 : 2048) : 0));  // with the output list's count zeroed
    if (learned) {
        const int32_t *lin = lists && lin_sel >= 0 ? c->d_live[lin_sel] : nullptr;
        const uint32_t *lin_cnt = lin ? &c->d_ctl->lcnt[lin_sel] : nullptr;
        int32_t *lout = lists ? c->d_live[lout_sel] : nullptr;
        uint32_t *lout_cnt = lists ? &c->d_ctl->lcnt[lout_sel] : nullptr;
        launch_update(l, c->d_word, c->d_root, c->g.n, prev, cur, c->slot_bits, elect_tag(c->epoch), recolour, 0,
                      nullptr, c->d_ctl, lin, lin_cnt, lout, lout_cnt, c->track_mn ? c->d_mn : nullptr, c->d_oid);
        if (lists) {
            c->lsel_prev = lin_sel;
            c->lsel = lout_sel;
        }
        // colour-set sizes: in the first election of a single graph every live vertex has
        // the one colour, and the device's live count is the size
        const uint32_t *one_set = !recolour && c->cur_step == 0 && c->num_graphs == 0 ? &c->d_ctl->live : nullptr;
        if (!one_set) {
            launch_fill_i32(l, (int32_t *)c->d_cnt, nslots, 0);
            launch_colour_count(l, c->d_word, c->g.n, c->d_cnt, c->slot_bits, lout, lout_cnt, c->d_ctl);
        }
        const float step_f = 0.25f * std::log1p((float)c->cur_step);
        launch_score(l, c->g, c->d_word, c->d_stat, c->d_cnt, one_set, c->slot_bits, step_f, c->cur_step == 0 ? 1 : 0,
                     lout, lout_cnt, c->d_wb, nullptr, nullptr, cur, elect_tag(c->epoch), c->d_ctl, c->d_oid);
    } else if (lists) {
        // the update reads the list of the election before (all n the first time)
        // and lists the vertices still live; the confirm pass reads that list
        const int32_t *lin = lin_sel >= 0 ? c->d_live[lin_sel] : nullptr;
        const uint32_t *lin_cnt = lin_sel >= 0 ? &c->d_ctl->lcnt[lin_sel] : nullptr;
        launch_update(l, c->d_word, c->d_root, c->g.n, prev, cur, c->slot_bits, elect_tag(c->epoch),
                      recolour, 1, nullptr, c->d_ctl, lin, lin_cnt, c->d_live[lout_sel],
                      &c->d_ctl->lcnt[lout_sel], c->track_mn ? c->d_mn : nullptr, c->d_oid);
        c->lsel_prev = lin_sel;
        c->lsel = lout_sel;
    } else {
        launch_update(l, c->d_word, c->d_root, c->g.n, prev, cur, c->slot_bits, elect_tag(c->epoch),
                      recolour, 1, nullptr, c->d_ctl, nullptr, nullptr, nullptr, nullptr,
                      c->track_mn ? c->d_mn : nullptr, c->d_oid);
    }
    uint32_t *vis = c->vis_mark ? c->d_vis : nullptr;
    if (lists)
        launch_confirm(l, c->d_word, c->g.n, cur, c->slot_bits, c->d_piv, c->d_ctl, c->d_live[lout_sel],
                       &c->d_ctl->lcnt[lout_sel], vis, c->vis_words, c->d_oid);
    else
        launch_confirm(l, c->d_word, c->g.n, cur, c->slot_bits, c->d_piv, c->d_ctl, nullptr, nullptr, vis,
                       c->vis_words, c->d_oid);
    c->vis_mark = false;
    ctl_queue(c, 64);  // the pivot list seeds both directions
    c->prev_tab ^= 1;
    if (!sync) return SCC_OK;
    CHK(read_ctl(c));
    *npiv = c->h_ctl->npiv;
    return SCC_OK;
}


scc_status do_count(scc_ctx *c, int what, int64_t *out) {
    ctl_op(c, 8);
    launch_count_state(launch_for(c), c->d_word, c->g.n, what, c->d_ctl);
    CHK(read_ctl(c));
    *out = (int64_t)c->h_ctl->count;
    return SCC_OK;
}

scc_status do_trim2(scc_ctx *c, int64_t *pairs) {
    ctl_op(c, 8);
    launch_trim2(launch_for(c), c->g, c->d_word, c->d_root, c->d_ctl, c->d_oid);
    CHK(read_ctl(c));
    *pairs = (int64_t)c->h_ctl->count;
    return SCC_OK;
}

// update_wcc (wcc.cu:146-226): labels to fixpoint, fresh colours, election
scc_status do_wcc(scc_ctx *c, int64_t *iters, int64_t *npiv, bool learned = false, bool sync = true) {
    Launch l = launch_for(c);
    ctl_op(c, 16);
    launch_max_colour(l, c->d_word, c->g.n, c->d_ctl);  // over every vertex, as thrust::reduce (scc.cu:802)
    // the live vertices: the list of the last election when the run keeps one
    const int32_t *lin = c->use_lists && c->lsel >= 0 ? c->d_live[c->lsel] : nullptr;
    const uint32_t *lin_cnt = lin ? &c->d_ctl->lcnt[c->lsel] : nullptr;
    launch_wcc_init(l, c->d_tmp, c->g.n, lin, lin_cnt, c->d_oid);
    CHK(run_chain(c, [&](int it, hipEvent_t, hipEvent_t) {
            launch_wcc_iter(l, c->g, c->d_word, c->d_tmp, c->d_ctl, it, lin, lin_cnt, c->d_nid);
        }, c->guess_wcc, 0, iters));
    launch_wcc_recolour(l, c->d_word, c->d_tmp, c->g.n, c->d_ctl, lin, lin_cnt);
    return do_elect(c, 0, npiv, learned, sync);
}

// Hub-first rows: every row of the CSR and of the CSR^T lists its neighbours by
// decreasing total degree (ties by id).  The order inside a row changes no
// result; the scans that stop at the first live or reached neighbour (trim
// sweeps, pull levels) meet the hubs first, and the hubs' words stay in cache.
void hub_first_rows(int32_t n, int64_t m, const int32_t *row, const int32_t *col, const int32_t *rowT,
                    const int32_t *colT, std::vector<int32_t> &oc, std::vector<int32_t> &ocT) {
    std::vector<int32_t> deg((size_t)n);
    for (int32_t v = 0; v < n; v++) deg[v] = (row[v + 1] - row[v]) + (rowT[v + 1] - rowT[v]);
    oc.assign(col, col + m);
    ocT.assign(colT, colT + m);
    auto before = [&](int32_t a, int32_t b) { return deg[a] != deg[b] ? deg[a] > deg[b] : a < b; };
    const int32_t T = (int32_t)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int32_t t = 0; t < T; t++) {
        pool.emplace_back([&, t] {
            int32_t lo = (int32_t)((int64_t)n * t / T), hi = (int32_t)((int64_t)n * (t + 1) / T);
            for (int32_t v = lo; v < hi; v++) {
                std::sort(oc.begin() + row[v], oc.begin() + row[v + 1], before);
                std::sort(ocT.begin() + rowT[v], ocT.begin() + rowT[v + 1], before);
            }
        });
    }
    for (auto &th : pool) th.join();
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void free_batch(scc_ctx *c) {
    for (void *p : {(void *)c->d_border, (void *)c->d_btab_off, (void *)c->d_btab})
        if (p) (void)hipFree(p);
    c->d_border = nullptr;
    c->d_btab_off = nullptr;
    c->d_btab = nullptr;
}

void free_tiles(scc_ctx *c) {
    for (void *p : {(void *)c->d_toff, (void *)c->d_tflags, (void *)c->d_tdone, (void *)c->d_tadj,
                    (void *)c->d_txarc, (void *)c->d_tct, (void *)c->d_txout, (void *)c->d_xflag})
        if (p) (void)hipFree(p);
    c->d_toff = nullptr;
    c->d_tflags = nullptr;
    c->d_tdone = nullptr;
    c->d_tadj = c->d_txarc = c->d_tct = nullptr;
    c->d_txout = nullptr;
    c->d_xflag = nullptr;
    c->ntiles = 0;
}

void free_graph(scc_ctx *c) {
    free_tiles(c);
    clear_graphs(c);  // captured launches point at the buffers freed here
    int32_t **bufs[] = {&c->d_row, &c->d_col, &c->d_rowT, &c->d_colT, &c->d_root, &c->d_label,
                        &c->d_tmp, &c->d_mn, &c->d_q[0], &c->d_q[1], &c->d_q[2], &c->d_q[3],
                        &c->d_piv, &c->d_seedB, &c->d_huge[0], &c->d_huge[1], &c->d_huge[2],
                        &c->d_huge[3], (int32_t **)&c->d_stamp, (int32_t **)&c->d_dcls, &c->d_head};
    for (auto b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    if (c->d_word) (void)hipFree(c->d_word);
    c->d_word = nullptr;
    if (c->d_ell) (void)hipFree(c->d_ell);
    c->d_ell = nullptr;
    if (c->d_hlive) (void)hipFree(c->d_hlive);
    c->d_hlive = nullptr;
    for (auto &p : c->d_live) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    if (c->d_vis) (void)hipFree(c->d_vis);
    c->d_vis = nullptr;
    c->vis_words = 0;
    c->lsel = c->lsel_prev = -1;
    for (auto &t : c->d_slots) {
        if (t) (void)hipFree(t);
        t = nullptr;
    }
    if (c->d_score) (void)hipFree(c->d_score);
    c->d_score = nullptr;
    if (c->d_stat) (void)hipFree(c->d_stat);
    c->d_stat = nullptr;
    if (c->d_cnt) (void)hipFree(c->d_cnt);
    c->d_cnt = nullptr;
    if (c->d_feat) (void)hipFree(c->d_feat);
    c->d_feat = nullptr;
    if (c->d_goff) (void)hipFree(c->d_goff);
    c->d_goff = nullptr;
    for (int32_t **p : {&c->d_oid, &c->d_nid}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    free_batch(c);
    c->num_graphs = 0;
    c->uploaded = false;
    c->labels_valid = false;
}

// ---- renumbering (opts.relabel) ----------------------------------------------
// The caller's ids in the new internal order: by graph (offsets off[0..G], may
// be null), then by decreasing total degree, then by id -- a stable LSD radix
// sort on 16-bit digits of the key (graph, ~degree), the ids entering in order.
std::vector<int32_t> degree_order(int32_t n, const std::vector<int32_t> &deg, const int32_t *off, int32_t G) {
    std::vector<uint64_t> key((size_t)n);
    int32_t g = 0;
    for (int32_t v = 0; v < n; v++) {
        while (off && g < G - 1 && off[g + 1] <= v) g++;
        key[v] = ((uint64_t)(uint32_t)g << 32) | (uint32_t) ~(uint32_t)deg[v];
    }
    int passes = 2;
    for (uint32_t x = off && G > 1 ? (uint32_t)(G - 1) : 0u; x; x >>= 16) passes++;
    std::vector<int32_t> a((size_t)n), b((size_t)n);
    for (int32_t v = 0; v < n; v++) a[v] = v;
    std::vector<int64_t> cnt(65537);
    for (int p = 0; p < passes; p++) {
        const int sh = 16 * p;
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int32_t i = 0; i < n; i++) cnt[((key[a[i]] >> sh) & 0xFFFFu) + 1]++;
        for (int d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
        for (int32_t i = 0; i < n; i++) b[cnt[(key[a[i]] >> sh) & 0xFFFFu]++] = a[i];
        a.swap(b);
    }
    return a;
}

// The row indexes derived from the device CSR as it stands: degree classes,
// heads, ELL entries, head-live bits and the scorer's static features.
void row_indexes(scc_ctx *c) {
    const int32_t n = c->g.n;
    Launch l = launch_for(c);
    launch_degree_class(l, c->d_row, c->d_col, n, c->d_dcls, c->d_head, c->d_ell);
    launch_degree_class(l, c->d_rowT, c->d_colT, n, c->d_dcls + n, c->d_head + n, c->d_ell + n);
    launch_head_live(l, c->d_dcls, c->d_head, n, c->d_hlive);
    // the scorer's static features (degrees, 2-hop degree sums): a property of the graph
    launch_static_features(l, c->g, c->d_stat);
    c->stat_valid = true;
}

// Whether a BFS from vertex 0 over the arcs taken both ways reaches depth `depth`
// (the graph's diameter is at least that); stops there.
bool bfs_depth_at_least(int32_t n, const int32_t *row, const int32_t *col, const int32_t *rowT,
                        const int32_t *colT, int32_t depth) {
    std::vector<int32_t> lev((size_t)n, -1), cur{0}, nxt;
    lev[0] = 0;
    for (int32_t l = 0; !cur.empty(); l++) {
        if (l >= depth) return true;
        nxt.clear();
        for (int32_t v : cur)
            for (int d = 0; d < 2; d++) {
                const int32_t *r = d ? rowT : row, *cl = d ? colT : col;
                for (int32_t e = r[v]; e < r[v + 1]; e++)
                    if (lev[cl[e]] < 0) {
                        lev[cl[e]] = l + 1;
                        nxt.push_back(cl[e]);
                    }
            }
        cur.swap(nxt);
    }
    return false;
}

// Tiles for the tile reach: regions of up to T vertices grown by BFS over the
// arcs taken both ways, each from the lowest vertex not yet taken and, when the
// region runs out of unvisited neighbours before T, from the next lowest, so
// every tile but the last holds T vertices.  Returns the caller's ids in tile
// order (a region's vertices in BFS order) and the tile offsets.
std::vector<int32_t> tile_partition(int32_t n, const int32_t *row, const int32_t *col, const int32_t *rowT,
                                    const int32_t *colT, int32_t T, std::vector<int32_t> &toff) {
    std::vector<int32_t> order;
    order.reserve((size_t)n);
    std::vector<uint8_t> seen((size_t)n, 0);
    toff.assign(1, 0);
    int32_t seed = 0;
    while ((int32_t)order.size() < n) {
        const size_t start = order.size();
        size_t head = start;
        while ((int64_t)(order.size() - start) < T && (int32_t)order.size() < n) {
            if (head == order.size()) {  // the region is closed: go on from the next lowest vertex
                while (seen[seed]) seed++;
                seen[seed] = 1;
                order.push_back(seed);
            }
            const int32_t v = order[head++];
            for (int d = 0; d < 2; d++) {
                const int32_t *r = d ? rowT : row, *cl = d ? colT : col;
                for (int32_t e = r[v]; e < r[v + 1] && (int64_t)(order.size() - start) < T; e++) {
                    const int32_t u = cl[e];
                    if (seen[u]) continue;
                    seen[u] = 1;
                    order.push_back(u);
                }
            }
        }
        toff.push_back((int32_t)order.size());
    }
    return order;
}

// Cells: every tile's vertices regrouped into cells of CELL vertices grown by
// BFS over the tile's own vertices (arcs taken both ways), each from the
// tile's first vertex not yet taken (in the tile's order) and, when the cell's
// BFS runs out of neighbours in the tile, from the next one; so every cell of
// a tile but its last is full.  order/toff: tile_partition's result, changed
// in place (the tiles keep their vertex sets and offsets).
void cell_order(int32_t n, const int32_t *row, const int32_t *col, const int32_t *rowT, const int32_t *colT,
                std::vector<int32_t> &order, const std::vector<int32_t> &toff) {
    const int32_t nt = (int32_t)toff.size() - 1;
    std::vector<int32_t> tile((size_t)n), out;
    out.reserve((size_t)n);
    for (int32_t t = 0; t < nt; t++)
        for (int32_t i = toff[t]; i < toff[t + 1]; i++) tile[order[i]] = t;
    std::vector<uint8_t> taken((size_t)n, 0);
    for (int32_t t = 0; t < nt; t++) {
        int32_t next = toff[t];  // the tile's next seed candidate, in the tile's order
        while ((int32_t)out.size() < toff[t + 1]) {
            const size_t start = out.size();
            size_t head = start;
            auto full = [&] { return out.size() - start >= CELL || (int32_t)out.size() >= toff[t + 1]; };
            while (!full()) {
                if (head == out.size()) {
                    while (taken[order[next]]) next++;
                    taken[order[next]] = 1;
                    out.push_back(order[next]);
                    continue;
                }
                const int32_t v = out[head++];
                for (int d = 0; d < 2 && !full(); d++) {
                    const int32_t *r = d ? rowT : row, *cl = d ? colT : col;
                    for (int32_t e = r[v]; e < r[v + 1] && !full(); e++) {
                        const int32_t u = cl[e];
                        if (taken[u] || tile[u] != t) continue;
                        taken[u] = 1;
                        out.push_back(u);
                    }
                }
            }
        }
    }
    order.swap(out);
}

// The tile structures of a graph renumbered by `order` (tiles toff, cells of CELL
// vertices from each tile's first vertex): per vertex and direction its neighbours
// in its cell as a mask (bit = lane), up to CELL_XARCS neighbours in other cells of
// its tile as 16-bit indices from the tile's first vertex, up to two neighbours in
// other tiles, and flags (bit d: an arc of direction d leaves the tile; bit 2 + d:
// more than CELL_XARCS go to other cells; bit 4 + d: more than two leave the tile).
// row .. colT: the CSR pair in the caller's ids.
scc_status build_tiles(scc_ctx *c, const int32_t *row, const int32_t *col, const int32_t *rowT,
                       const int32_t *colT, const std::vector<int32_t> &order, const std::vector<int32_t> &toff) {
    const int32_t n = c->g.n, nt = (int32_t)toff.size() - 1;
    for (int32_t t = 1; t < nt; t++)  // the kernel finds a vertex's tile by division
        if (toff[t] != t * toff[1]) return fail(c, SCC_E_STATE, "tiles of unequal size");
    std::vector<int32_t> nid((size_t)n), tof((size_t)n);
    for (int32_t i = 0; i < n; i++) nid[order[i]] = i;
    for (int32_t t = 0; t < nt; t++)
        for (int32_t i = toff[t]; i < toff[t + 1]; i++) tof[i] = t;
    std::vector<unsigned long long> adj(2 * (size_t)n, 0ull), xarc(2 * (size_t)n, ~0ull);
    std::vector<int2> xout(2 * (size_t)n, int2{-1, -1});
    std::vector<uint8_t> xf((size_t)n, 0);
    for (int d = 0; d < 2; d++) {
        const int32_t *r = d ? rowT : row, *cl = d ? colT : col;
        for (int32_t i = 0; i < n; i++) {
            const int32_t o = order[i], t = tof[i], lo = toff[t];
            const int32_t cell = (i - lo) / (int32_t)CELL;
            unsigned long long m = 0ull, x = ~0ull;
            int nx = 0, no = 0;
            int2 &xo = xout[(size_t)d * n + i];
            for (int32_t e = r[o]; e < r[o + 1]; e++) {
                const int32_t u = nid[cl[e]];
                if (u == i) continue;  // a self-loop reaches nothing new
                if (tof[u] != t) {
                    xf[i] |= (uint8_t)(1u << d);
                    if (no == 0) xo.x = u;
                    else if (no == 1) xo.y = u;
                    else xf[i] |= (uint8_t)(16u << d);
                    no++;
                } else if ((u - lo) / (int32_t)CELL == cell) {
                    m |= 1ull << ((u - lo) % (int32_t)CELL);
                } else if (nx < CELL_XARCS) {
                    x = (x & ~(0xFFFFull << (16 * nx))) | ((unsigned long long)(u - lo) << (16 * nx));
                    nx++;
                } else {
                    xf[i] |= (uint8_t)(4u << d);
                }
            }
            adj[(size_t)d * n + i] = m;
            xarc[(size_t)d * n + i] = x;
        }
    }
    free_tiles(c);
    HIPCHK(c, dev_alloc(c, &c->d_toff, sizeof(int32_t) * toff.size()));
    HIPCHK(c, dev_alloc(c, &c->d_tadj, sizeof(unsigned long long) * adj.size()));
    HIPCHK(c, dev_alloc(c, &c->d_txarc, sizeof(unsigned long long) * xarc.size()));
    HIPCHK(c, dev_alloc(c, &c->d_tct, sizeof(unsigned long long) * adj.size()));
    HIPCHK(c, dev_alloc(c, &c->d_txout, sizeof(int2) * xout.size()));
    HIPCHK(c, dev_alloc(c, &c->d_xflag, (size_t)n));
    HIPCHK(c, dev_alloc(c, &c->d_tflags, 3 * (size_t)nt + 4));
    HIPCHK(c, dev_alloc(c, &c->d_tdone, sizeof(unsigned long long) * 2 * TILE_MAX / CELL * (size_t)nt));
    HIPCHK(c, hipMemcpy(c->d_toff, toff.data(), sizeof(int32_t) * toff.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_tadj, adj.data(), sizeof(unsigned long long) * adj.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_txarc, xarc.data(), sizeof(unsigned long long) * xarc.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_xflag, xf.data(), (size_t)n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_txout, xout.data(), sizeof(int2) * xout.size(), hipMemcpyHostToDevice));
    const Launch l = launch_for(c);
    launch_fill_i32(l, (int32_t *)c->d_tflags, (3 * (int64_t)nt + 3) / 4, 0);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ntiles = nt;
    c->tile_size = toff.size() > 1 ? toff[1] : 0;
    return SCC_OK;
}

// Renumbers the device CSR so that internal vertex i is the caller's vertex
// order[i].  row, rowT: the current internal row offsets (host); cur_oid: the
// caller's id of each current internal vertex (empty: the identity).  A row
// keeps the order of its entries, so a hub-first row (decreasing degree, then
// id) comes out ascending in a numbering that follows the same order.
scc_status permute_graph(scc_ctx *c, const int32_t *row, const int32_t *rowT, const std::vector<int32_t> &order,
                         const std::vector<int32_t> &cur_oid) {
    const int32_t n = c->g.n;
    const int64_t m = c->g.m;
    std::vector<int32_t> nid((size_t)n), map((size_t)n), cur_of((size_t)n);
    for (int32_t i = 0; i < n; i++) nid[order[i]] = i;
    for (int32_t v = 0; v < n; v++) {
        const int32_t o = cur_oid.empty() ? v : cur_oid[v];
        map[v] = nid[o];
        cur_of[o] = v;
    }
    std::vector<int32_t> nrow((size_t)n + 1), nrowT((size_t)n + 1);
    nrow[0] = nrowT[0] = 0;
    for (int32_t i = 0; i < n; i++) {
        const int32_t u = cur_of[order[i]];
        nrow[i + 1] = nrow[i] + (row[u + 1] - row[u]);
        nrowT[i + 1] = nrowT[i] + (rowT[u + 1] - rowT[u]);
    }
    const size_t nb = sizeof(int32_t) * ((size_t)n + 1), mb = sizeof(int32_t) * std::max<size_t>((size_t)m, 1);
    const size_t vb = sizeof(int32_t) * std::max<size_t>((size_t)n, 1);
    int32_t *nr = nullptr, *nc = nullptr, *nrT = nullptr, *ncT = nullptr;
    auto drop = [&]() {
        for (int32_t *p : {nr, nc, nrT, ncT})
            if (p) (void)hipFree(p);
    };
    hipError_t e = dev_alloc(c, &nr, nb);
    if (e == hipSuccess) e = dev_alloc(c, &nrT, nb);
    if (e == hipSuccess) e = dev_alloc(c, &nc, mb);
    if (e == hipSuccess) e = dev_alloc(c, &ncT, mb);
    if (e == hipSuccess && !c->d_oid) e = dev_alloc(c, &c->d_oid, vb);
    if (e == hipSuccess && !c->d_nid) e = dev_alloc(c, &c->d_nid, vb);
    const size_t ib = sizeof(int32_t) * (size_t)n;  // n > 0 (the caller's check)
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_tmp, map.data(), ib, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(nr, nrow.data(), nb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(nrT, nrowT.data(), nb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_oid, order.data(), ib, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_nid, nid.data(), ib, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const Launch l = launch_for(c);
        launch_permute_rows(l, c->d_row, c->d_col, c->d_tmp, nr, nc, n);
        launch_permute_rows(l, c->d_rowT, c->d_colT, c->d_tmp, nrT, ncT, n);
        e = hipGetLastError();
    }
    const hipError_t es = hipStreamSynchronize(c->stream);  // the host vectors go out of scope
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) {
        drop();
        return hip_fail(c, e, "renumbering");
    }
    for (int32_t *p : {c->d_row, c->d_col, c->d_rowT, c->d_colT}) (void)hipFree(p);
    c->d_row = nr;
    c->d_col = nc;
    c->d_rowT = nrT;
    c->d_colT = ncT;
    c->g = Graph{n, m, c->d_row, c->d_col, c->d_rowT, c->d_colT};
    clear_graphs(c);  // captured launches point at the freed rows
    return SCC_OK;
}

}  // namespace

namespace {

// ---- learned pivot scorer ---------------------------------------------------
scc_status mlp_alloc(scc_ctx *c) {
    if (c->d_wf) return SCC_OK;
    HIPCHK(c, hipMalloc(&c->d_wf, sizeof(float) * MLP_PARAMS));
    HIPCHK(c, hipMalloc(&c->d_m, sizeof(float) * MLP_PARAMS));
    HIPCHK(c, hipMalloc(&c->d_v, sizeof(float) * MLP_PARAMS));
    HIPCHK(c, hipMalloc(&c->d_wb, sizeof(uint16_t) * MLP_PARAMS));
    return SCC_OK;
}

struct DevBuf {  // scoped device allocation for the training entry points
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

scc_status mlp_work_alloc(scc_ctx *c, int64_t B, MlpWork &w, DevBuf *bufs) {
    size_t b = (size_t)std::max<int64_t>(B, 1);
    size_t sizes[8] = {b * MLP_HID * 2, b * MLP_HID * 2, b * MLP_HID * 4, b * MLP_HID * 4, b * 4,
                       b * MLP_IN * 4, MLP_PARAMS * 4, 4};
    for (int i = 0; i < 8; i++) HIPCHK(c, hipMalloc(&bufs[i].p, sizes[i]));
    w.H1 = (uint16_t *)bufs[0].p;
    w.H2 = (uint16_t *)bufs[1].p;
    w.dH1 = (float *)bufs[2].p;
    w.dH2 = (float *)bufs[3].p;
    w.ds = (float *)bufs[4].p;
    w.Xb = (float *)bufs[5].p;
    w.grad = (float *)bufs[6].p;
    w.loss = (float *)bufs[7].p;
    return SCC_OK;
}

Launch mlp_launch(const scc_ctx *c, int64_t N) {
    uint64_t need = ((uint64_t)std::max<int64_t>(N, 1) + 255) / 256;
    uint32_t grid = (uint32_t)std::min<uint64_t>(need, (uint64_t)c->cu_count * 4);
    return Launch{c->stream, std::max<uint32_t>(grid, 1), 256};
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *scc_status_string(scc_status s) {
    switch (s) {
    case SCC_OK: return "ok";
    case SCC_E_HIP: return "HIP runtime error";
    case SCC_E_OOM: return "out of memory";
    case SCC_E_INPUT: return "invalid input";
    case SCC_E_OVERFLOW: return "size limit exceeded";
    case SCC_E_STATE: return "invalid call order";
    }
    return "unknown status";
}

const char *scc_last_error(const scc_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

scc_status scc_create(int device, const scc_opts *opts, scc_ctx **out) {
    if (!out) return SCC_E_INPUT;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return SCC_E_HIP;
    if (device < 0 || device >= ndev) return SCC_E_INPUT;
    scc_ctx *c = new (std::nothrow) scc_ctx();
    if (!c) return SCC_E_OOM;
    c->device = device;
    if (opts) c->opts = *opts;
    if ((c->opts.block_size != 0 && (c->opts.block_size < 64 || c->opts.block_size > 1024 ||
                                     (c->opts.block_size & 63) != 0)) ||
        (c->opts.reach_block != 0 && c->opts.reach_block != 256 && c->opts.reach_block != 512) ||
        c->opts.tail_live < -1 || c->opts.relabel < -1 || c->opts.relabel > 1 || c->opts.tile_reach < -1) {
        delete c;
        return SCC_E_INPUT;
    }
    if (hipSetDevice(device) != hipSuccess) { delete c; return SCC_E_HIP; }
    {  // the kernels loaded onto this device once per process, outside every run
        static std::mutex m;
        static std::vector<bool> loaded;
        std::lock_guard<std::mutex> lock(m);
        if ((size_t)device >= loaded.size()) loaded.resize((size_t)device + 1, false);
        if (!loaded[(size_t)device]) {
            for (hipError_t (*f)() : {preload_fb, preload_reach, preload_mlp, preload_tail, preload_tile,
                                      preload_noalloc})
                if (f() != hipSuccess) { delete c; return SCC_E_HIP; }
            loaded[(size_t)device] = true;
        }
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        if (prop.multiProcessorCount > 0) c->cu_count = prop.multiProcessorCount;
        size_t k = 0;  // "gfx950:sramecc+:xnack-" -> "gfx950"
        while (k + 1 < sizeof(c->arch) && prop.gcnArchName[k] && prop.gcnArchName[k] != ':') {
            c->arch[k] = prop.gcnArchName[k];
            k++;
        }
        c->arch[k] = 0;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SCC_E_HIP;
    }
    if (hipMalloc(&c->d_ctl, sizeof(Ctl)) != hipSuccess ||
        hipHostMalloc(&c->h_ctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess) {
        scc_destroy(c);
        return SCC_E_OOM;
    }
    std::memset(c->h_ctl, 0, sizeof(Ctl));
    if (hipMemcpy(c->d_ctl, c->h_ctl, sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess) {
        scc_destroy(c);
        return SCC_E_HIP;
    }
    *out = c;
    return SCC_OK;
}

scc_status scc_set_opts(scc_ctx *c, const scc_opts *opts) {
    if (!c || !opts) return SCC_E_INPUT;
    if (opts->block_size != 0 &&
        (opts->block_size < 64 || opts->block_size > 1024 || (opts->block_size & 63) != 0))
        return fail(c, SCC_E_INPUT, "block_size must be a multiple of 64 in [64, 1024]");
    if (opts->reach_block != 0 && opts->reach_block != 256 && opts->reach_block != 512)
        return fail(c, SCC_E_INPUT, "reach_block must be 0, 256 or 512");
    if (opts->tail_live < -1) return fail(c, SCC_E_INPUT, "tail_live must be -1 (off), 0 (default) or a count");
    if (opts->grid_blocks < 0) return fail(c, SCC_E_INPUT, "grid_blocks < 0");
    if (opts->relabel < -1 || opts->relabel > 1) return fail(c, SCC_E_INPUT, "relabel must be -1, 0 or 1");
    if (opts->tile_reach < -1) return fail(c, SCC_E_INPUT, "tile_reach must be -1, 0, 1 or a tile size");
    if (c->uploaded && opts->slot_bits != 0 && opts->slot_bits != c->slot_bits)
        return fail(c, SCC_E_STATE, "slot_bits is fixed once a graph is uploaded");
    c->opts = *opts;
    clear_graphs(c);        // captured launches hold the old options
    c->spec_reach.clear();  // chain depths depend on the options
    c->spec_trim.clear();
    return SCC_OK;
}

scc_status scc_device_info(const scc_ctx *c, char *arch, int32_t len, int32_t *cu_count) {
    if (!c || (len > 0 && !arch)) return SCC_E_INPUT;
    if (len > 0) {
        std::strncpy(arch, c->arch, (size_t)len - 1);
        arch[len - 1] = 0;
    }
    if (cu_count) *cu_count = c->cu_count;
    return SCC_OK;
}

void scc_destroy(scc_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_graph(c);
    for (auto e : c->evs) (void)hipEventDestroy(e);
    for (auto e : c->spec_ev)
        if (e) (void)hipEventDestroy(e);
    for (void *p : {(void *)c->d_wf, (void *)c->d_m, (void *)c->d_v, (void *)c->d_wb})
        if (p) (void)hipFree(p);
    if (c->d_ctl) (void)hipFree(c->d_ctl);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

scc_status scc_upload(scc_ctx *c, int32_t n, int64_t m, const int32_t *row, const int32_t *col,
                      const int32_t *rowT, const int32_t *colT) {
    if (!c) return SCC_E_INPUT;
    if (n < 0 || m < 0) return fail(c, SCC_E_INPUT, "negative size");
    if (m >= (int64_t)INT32_MAX || n >= INT32_MAX - 1) return fail(c, SCC_E_OVERFLOW, "graph too large for int32 CSR");
    if (!row || (m > 0 && !col)) return fail(c, SCC_E_INPUT, "null CSR");
    if (row[0] != 0 || row[n] != m) return fail(c, SCC_E_INPUT, "row[0] != 0 or row[n] != m");
    if ((rowT == nullptr) != (colT == nullptr)) return fail(c, SCC_E_INPUT, "rowT/colT: both or neither");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_graph(c);
    c->alloc_count = 0;
    // validate rows and build the transpose on the host when it is not given
    std::vector<int32_t> hrT, hcT;
    for (int32_t v = 0; v < n; v++)
        if (row[v + 1] < row[v]) return fail(c, SCC_E_INPUT, "row offsets decrease");
    for (int64_t i = 0; i < m; i++)
        if (col[i] < 0 || col[i] >= n) return fail(c, SCC_E_INPUT, "column id out of range");
    if (!rowT) {
        scc_host_graph tg;
        std::vector<int32_t> src((size_t)m);
        for (int32_t v = 0; v < n; v++)
            for (int32_t i = row[v]; i < row[v + 1]; i++) src[i] = v;
        scc_status s = csr_from_arcs(n, m, col, src.data(), 0, &tg);
        if (s != SCC_OK) return fail(c, s, "transpose failed");
        if (tg.m != m) {
            graph_free(&tg);
            return fail(c, SCC_E_INPUT, "CSR rows are not duplicate-free");
        }
        hrT.assign(tg.row, tg.row + n + 1);
        hcT.assign(tg.col, tg.col + m);
        graph_free(&tg);
        rowT = hrT.data();
        colT = hcT.data();
    } else {
        if (rowT[0] != 0 || rowT[n] != m) return fail(c, SCC_E_INPUT, "rowT[0] != 0 or rowT[n] != m");
        for (int32_t v = 0; v < n; v++)
            if (rowT[v + 1] < rowT[v]) return fail(c, SCC_E_INPUT, "rowT offsets decrease");
        for (int64_t i = 0; i < m; i++)
            if (colT[i] < 0 || colT[i] >= n) return fail(c, SCC_E_INPUT, "colT id out of range");
    }
    std::vector<int32_t> hub_col, hub_colT;
    hub_first_rows(n, m, row, col, rowT, colT, hub_col, hub_colT);
    col = hub_col.data();
    colT = hub_colT.data();
    size_t nb = sizeof(int32_t) * ((size_t)n + 1), mb = sizeof(int32_t) * std::max<size_t>((size_t)m, 1);
    size_t vb = sizeof(int32_t) * std::max<size_t>((size_t)n, 1);
    HIPCHK(c, dev_alloc(c, &c->d_row, nb));
    HIPCHK(c, dev_alloc(c, &c->d_col, mb));
    HIPCHK(c, dev_alloc(c, &c->d_rowT, nb));
    HIPCHK(c, dev_alloc(c, &c->d_colT, mb));
    HIPCHK(c, dev_alloc(c, &c->d_word, vb));
    HIPCHK(c, dev_alloc(c, &c->d_root, vb));
    HIPCHK(c, dev_alloc(c, &c->d_label, vb));
    HIPCHK(c, dev_alloc(c, &c->d_tmp, vb));
    HIPCHK(c, dev_alloc(c, &c->d_mn, vb));
    for (auto &q : c->d_q) HIPCHK(c, dev_alloc(c, &q, vb));
    for (auto &q : c->d_huge) HIPCHK(c, dev_alloc(c, &q, vb));
    HIPCHK(c, dev_alloc(c, &c->d_stamp, vb));
    HIPCHK(c, dev_alloc(c, &c->d_dcls, 2 * std::max<size_t>((size_t)n, 1)));
    HIPCHK(c, dev_alloc(c, &c->d_head, 2 * vb));
    HIPCHK(c, dev_alloc(c, &c->d_ell, 2 * sizeof(int4) * std::max<size_t>((size_t)n, 1)));
    HIPCHK(c, dev_alloc(c, &c->d_hlive, 2 * std::max<size_t>((size_t)n, 1)));
    for (auto &p : c->d_live) HIPCHK(c, dev_alloc(c, &p, vb));
    c->vis_words = ((int64_t)n + 63) / 64 * 2 + 2;
    HIPCHK(c, dev_alloc(c, &c->d_vis, sizeof(uint32_t) * 2 * (size_t)c->vis_words));
    HIPCHK(c, hipMemsetAsync(c->d_stamp, 0, vb, c->stream));
    c->trim_tag = 0;
    HIPCHK(c, dev_alloc(c, &c->d_seedB, vb));
    HIPCHK(c, dev_alloc(c, &c->d_piv, vb));
    // election table: about n slots (at least 2^10, at most 2^22), or the option
    int sb = 10;
    while (sb < 22 && ((int64_t)1 << sb) < (int64_t)n) sb++;
    if (c->opts.slot_bits > 0) sb = std::min(24, std::max(1, (int)c->opts.slot_bits));
    c->slot_bits = sb;
    size_t sbytes = sizeof(unsigned long long) << sb;
    HIPCHK(c, dev_alloc(c, &c->d_slots[0], sbytes));
    HIPCHK(c, dev_alloc(c, &c->d_slots[1], sbytes));
    // scorer buffers (used by the learned pivot rule; allocated here so runs never allocate)
    HIPCHK(c, dev_alloc(c, &c->d_stat, sizeof(float4) * std::max<size_t>((size_t)n, 1)));
    HIPCHK(c, dev_alloc(c, &c->d_cnt, sizeof(uint32_t) << sb));
    HIPCHK(c, dev_alloc(c, &c->d_score, sizeof(uint16_t) * std::max<size_t>((size_t)n, 1)));
    HIPCHK(c, hipMemcpyAsync(c->d_row, row, nb, hipMemcpyHostToDevice, c->stream));
    if (m > 0) HIPCHK(c, hipMemcpyAsync(c->d_col, col, sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_rowT, rowT, nb, hipMemcpyHostToDevice, c->stream));
    if (m > 0) HIPCHK(c, hipMemcpyAsync(c->d_colT, colT, sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice, c->stream));
    c->g = Graph{n, m, c->d_row, c->d_col, c->d_rowT, c->d_colT};
    Launch l = launch_for(c);
    launch_fill_u64(l, c->d_slots[0], (int64_t)1 << sb, ~0ull);
    launch_fill_u64(l, c->d_slots[1], (int64_t)1 << sb, ~0ull);
    // renumbering by decreasing degree: always, never, or when the degrees are skewed
    std::vector<int32_t> deg((size_t)n);
    int64_t dmax = 0;
    for (int32_t v = 0; v < n; v++) {
        deg[v] = (row[v + 1] - row[v]) + (rowT[v + 1] - rowT[v]);
        dmax = std::max<int64_t>(dmax, deg[v]);
    }
    const bool skewed = n >= 65536 && dmax * (int64_t)n > 64 * 2 * m;  // max above 64 x the mean (2m / n)
    // tiles for the tile reach: when asked for, and by default for graphs of high diameter
    // (a long BFS from vertex 0), where the level kernels pay a launch or a barrier per level
    const bool tiles = n > 0 && (c->opts.tile_reach > 0 ||
                                 (c->opts.tile_reach == 0 && c->opts.relabel == 0 && !skewed && n >= 65536 &&
                                  bfs_depth_at_least(n, row, col, rowT, colT, TILE_DEPTH)));
    if (tiles) {
        const int32_t T = c->opts.tile_reach > 1 ? std::min<int32_t>(c->opts.tile_reach, (int32_t)TILE_MAX)
                                                 : (int32_t)TILE_MAX;
        std::vector<int32_t> toff;
        std::vector<int32_t> order = tile_partition(n, row, col, rowT, colT, T, toff);
        cell_order(n, row, col, rowT, colT, order, toff);
        CHK(permute_graph(c, row, rowT, order, {}));
        CHK(build_tiles(c, row, col, rowT, colT, order, toff));
    } else if (n > 0 && (c->opts.relabel > 0 || (c->opts.relabel == 0 && skewed))) {
        CHK(permute_graph(c, row, rowT, degree_order(n, deg, nullptr, 0), {}));
    }
    row_indexes(c);
    c->epoch = 0;
    c->prev_tab = 0;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->uploaded = true;
    c->spec_reach.clear();
    c->spec_trim.clear();
    return SCC_OK;
}

scc_status scc_run(scc_ctx *c, int schedule, int variant, int pivot_rule, int32_t *labels_out,
                   scc_stats *stats) {
    if (!c) return SCC_E_INPUT;
    if (!c->uploaded) return fail(c, SCC_E_STATE, "scc_run before scc_upload");
    if (schedule < SCC_SCHED_Q || schedule > SCC_SCHED_B) return fail(c, SCC_E_INPUT, "bad schedule");
    if (variant < SCC_VARIANT_PLAIN || variant > SCC_VARIANT_2W) return fail(c, SCC_E_INPUT, "bad variant");
    if (pivot_rule != SCC_PIVOT_PLAIN && pivot_rule != SCC_PIVOT_LEARNED)
        return fail(c, SCC_E_INPUT, "bad pivot rule");
    if (pivot_rule == SCC_PIVOT_LEARNED && !c->have_params)
        return fail(c, SCC_E_STATE, "learned pivot rule needs parameters (scc_mlp_set_params or scc_mlp_train)");
    const bool learned = pivot_rule == SCC_PIVOT_LEARNED;
    HIPCHK(c, hipSetDevice(c->device));
    scc_stats st;
    std::memset(&st, 0, sizeof(st));
    st.first_pivot = -1;
    st.renumbered = c->ntiles > 0 ? 2 : c->d_oid ? 1 : 0;
    const int32_t n = c->g.n;
    c->syncs = 0;
    c->cur_step = 0;
    c->use_lists = true;  // elections of this run keep the live-vertex list (none yet)
    c->lsel = c->lsel_prev = -1;
    if (c->label_in_root) c->labels_valid = false;  // the run rewrites the root array
    HIPCHK(c, hipStreamSynchronize(c->stream));
    auto t_start = std::chrono::steady_clock::now();
    scc_status s = SCC_OK;
    // a batch with the plain variant and rule: every graph solved on its own in one
    // persistent launch (k_fb_batch); the graphs' schedule is the tail's (queue reach)
    const bool batch_solve = c->num_graphs > 0 && variant == SCC_VARIANT_PLAIN && !learned &&
                             c->opts.batch_solver == 1 && schedule != SCC_SCHED_C && schedule != SCC_SCHED_B &&
                             !(schedule == SCC_SCHED_H && c->opts.hybrid_policy == 1);
    auto solve_batch = [&]() -> scc_status {
        c->pend_bits = c->pend_stage = 0;
        ctl_op(c, 512 | 1024 | 8192);
        TailArgs ta{};
        ta.row = c->g.row;
        ta.col = c->g.col;
        ta.rowT = c->g.rowT;
        ta.colT = c->g.colT;
        ta.word = c->d_word;
        ta.root = c->d_root;
        ta.oid = c->d_oid;
        ta.n = n;
        ta.slot_bits = c->slot_bits;
        ta.piv = c->d_piv;
        ta.list[0] = c->d_live[0];
        ta.list[1] = c->d_live[1];
        for (int i = 0; i < 4; i++) ta.q[i] = c->d_q[i];
        ta.ctl = c->d_ctl;
        ta.xbits = schedule == SCC_SCHED_Q || schedule == SCC_SCHED_T ? 1 : 0;
        const BatchArgs ba{c->num_graphs, c->d_goff, c->d_border, c->d_btab, c->d_btab_off};
        // one workgroup per CU: each holds the 150 KB LDS copy of a small graph
        launch_fb_batch(c->stream, (uint32_t)c->cu_count, ta, ba);
        CHK(read_ctl(c));
        const Ctl &h = *c->h_ctl;
        if (h.tail[1] == 2) return fail(c, SCC_E_STATE, "batch steps did not converge");
        st.steps = h.tail[0];
        st.tail_steps = h.tail[0];
        st.pivots_total = (int64_t)h.tail_pivots;
        st.fwd_depths = st.bwd_depths = h.tail[2];
        st.trim_iters = h.tail[3];
        st.first_pivot = h.first_piv == 0xFFFFFFFFu ? -1 : (int64_t)h.first_piv;
        st.num_scc = (int64_t)h.pcount;
        c->label_in_root = true;  // lowest-id pivots: the roots are the labels
        return SCC_OK;
    };
    auto body = [&]() -> scc_status {
        if (batch_solve && n > 0) return solve_batch();
        Launch l = launch_for(c);
        // k_init runs the first level of trim1 (a degree test) on the words it
        // writes and lists, as the input of trim1's first level, the live
        // vertices whose first out- or in-entry may not stay live; its
        // control-block reset also starts trim1's chain (qsize[0][0] = 0, where
        // k_init counts its list)
        c->pend_bits = c->pend_stage = 0;
        ctl_op(c, 512 | 1024 | 128 | 2 | 1);
        // learned pivots are not the lowest ids of their SCCs: the updates keep each SCC's
        // minimum in d_mn (k_init sets mn[v] = v), and the labels are mn[root]
        c->track_mn = learned;
        launch_init(l, c->d_word, c->d_root, n, c->num_graphs ? c->d_goff : nullptr, c->num_graphs, c->d_dcls,
                    c->d_hlive, c->d_q[1], c->d_ctl, c->d_vis, c->vis_words, learned ? c->d_mn : nullptr,
                    c->d_oid);
        if (n == 0) return SCC_OK;
        int64_t it = 0, npiv = 0, ne = 0;
        bool wcc_done = false;
        bool unsynced = false;  // the last election's pivot count is not read yet
        const int64_t one_pct = n / 100;
        // schedules that switch once n/100 vertices are FB-eliminated: the
        // reference's HYBRID (t, then q; scc.cu:493-563), COMBINE (l, then q;
        // scc.cu:565-638) and BEST (the direction-optimising reach in step 1, t in
        // later steps, then q; scc.cu:640-700, reach_b.cu:374-725)
        const bool by_elim = (schedule == SCC_SCHED_H && c->opts.hybrid_policy == 1) || schedule == SCC_SCHED_C ||
                             schedule == SCC_SCHED_B;
        // Speculative steps: reach chain, trim chain and election go out without
        // host waits between them, each chain as long as the same step's chain
        // was in the last run.  A guard after each chain sets ctl->abort when the
        // chain may need more levels; every later kernel of the step then exits
        // at once, and the host resumes that chain and finishes the step with
        // waits.  The results are those of the waiting path.
        const bool spec_ok = !c->opts.sequential_reach && !by_elim && (c->prof_flags & 7) == 0;
        // scc_w / scc_2_w (scc.cu:787-821): once n/100 vertices are FB-eliminated, the
        // pivots just elected are unset, scc_2_w peels size-2 SCCs, and one WCC pass gives
        // fresh colours and a new election.  The eliminated count is the device's (updates
        // add to it), read with the step's wait; the pair count arrives with the WCC
        // chain's first wait; with sync false the new election's pivot count is read by
        // the next step (after_wcc: no pivot there means the run is over).
        bool after_wcc = false;
        auto wcc_pass = [&](bool sync) -> scc_status {
            wcc_done = true;
            st.wcc_step = st.steps;
            launch_unset_pivots(launch_for(c), c->d_word, n, c->d_ctl, c->d_piv);  // the pivot list
            if (variant == SCC_VARIANT_2W) {
                ctl_op(c, 8);
                launch_trim2(launch_for(c), c->g, c->d_word, c->d_root, c->d_ctl, c->d_oid);
            }
            int64_t wi = 0;
            CHK(do_wcc(c, &wi, &npiv, learned, sync));
            if (variant == SCC_VARIANT_2W) st.trim2_pairs += (int64_t)c->h_ctl->count;
            if (!sync) {
                unsynced = true;
                after_wcc = true;
            }
            return SCC_OK;
        };
        auto tp = std::chrono::steady_clock::now();
        // The tail: the remaining FB steps in one workgroup and one launch (k_fb_tail).
        // gate 0: the host has read that few vertices are live; otherwise the kernel
        // checks the device's live count against gate and does nothing above it (status
        // 3).  Either way *npiv is the last election's pivot count afterwards.
        const uint32_t tail_max = c->opts.tail_live ? (uint32_t)c->opts.tail_live : kTailLive;
        auto run_tail = [&](uint32_t gate) -> scc_status {
            auto t0 = std::chrono::steady_clock::now();
            TailArgs ta{};
            ta.row = c->g.row;
            ta.col = c->g.col;
            ta.rowT = c->g.rowT;
            ta.colT = c->g.colT;
            ta.word = c->d_word;
            ta.root = c->d_root;
            ta.oid = c->d_oid;
            ta.n = n;
            ta.slots[0] = c->d_slots[0];
            ta.slots[1] = c->d_slots[1];
            ta.slot_bits = c->slot_bits;
            ta.prev_tab = c->prev_tab;
            ta.epoch = c->epoch;
            ta.piv = c->d_piv;
            ta.list[0] = c->d_live[0];
            ta.list[1] = c->d_live[1];
            ta.lsel = c->lsel;
            for (int i = 0; i < 4; i++) ta.q[i] = c->d_q[i];
            ta.ctl = c->d_ctl;
            ta.max_steps = gate ? 0u : c->h_ctl->live + 2;  // each step eliminates its pivots
            ta.gate_live = gate;
            ta.xbits = schedule == SCC_SCHED_Q || schedule == SCC_SCHED_T ? 1 : 0;
            ta.mn = c->track_mn ? c->d_mn : nullptr;
            ta.instr = (c->prof_flags & 8) ? 1 : 0;
            if (c->pend_stage || c->pend_bits) ctl_op(c, 0);  // queued resets go out first
            launch_fb_tail(c->stream, ta);
            CHK(read_ctl(c));
            const Ctl &h = *c->h_ctl;
            if (c->prof_flags & 8) {
                for (int q = 0; q < 4; q++) c->prof.tail_phase_ms[q] += (double)h.st_tailclk[q] * 1e-5;
                c->prof.tail_local_live = h.tail_L;
                c->prof.tail_local_step = h.tail_lstep;
            }
            npiv = h.npiv;
            if (h.tail[1] == 3) return SCC_OK;  // not run: too many live vertices
            if (h.tail[1] == 2) return fail(c, SCC_E_STATE, "tail steps did not converge");
            const uint32_t ts = h.tail[0];
            st.tail_live = h.tail_live0;
            st.steps += ts;
            st.tail_steps += ts;
            st.pivots_total += (int64_t)h.tail_pivots;
            st.fwd_depths += h.tail[2];
            st.bwd_depths += h.tail[2];
            st.trim_iters += h.tail[3];
            st.ms_fwd += ms_since(t0);
            c->epoch += ts;
            if (ts & 1u) {
                c->prev_tab ^= 1;
                c->lsel ^= 1;
            }
            return SCC_OK;
        };
        // With speculative steps the rest of trim-1 goes out without a wait as well:
        // its chain (as long as trim-1's in the last run), a guard (stage 3), the first
        // election and the first step, whose wait tells whether trim-1 converged.
        // The first step must then take the speculative path: no tag wrap in the
        // first two elections.
        bool spec_t1 = spec_ok && ((c->epoch + 2) & 0xFFFFu) >= 2u;
        const int Lt1 = std::max(1, std::min(c->guess[3], 1024));
        if (spec_t1) {
            if (!c->spec_ev[0])
                for (auto &e : c->spec_ev) HIPCHK(c, hipEventCreate(&e));
            HIPCHK(c, hipEventRecord(c->spec_ev[4], c->stream));
            TrimChain t1(c, false);
            t1.after_init = true;  // k_init's reset started the chain
            for (int i = 0; i < Lt1; i++) t1(i, nullptr, nullptr);
            HIPCHK(c, hipEventRecord(c->spec_ev[5], c->stream));
            ctl_guard(c, Lt1 - 1, 3);  // goes out with the election's first reset
            c->vis_mark = true;        // the first pivots seed the visited bitmaps k_init cleared
            CHK(do_elect(c, 0, &npiv, learned, false));
            c->vis_valid = true;
            unsynced = true;  // live vertices are not known yet: the first step's wait tells
        } else {
            CHK(do_trim(c, &it, true));  // the rest of trim1 (scc.cu:232)
            st.trim_iters += it + 1;
            st.ms_trim += ms_since(tp);
            // The device counts live vertices: after trim-1 every live vertex still
            // has its initial colour, so live vertices mean at least one pivot and
            // the first election needs no wait of its own (its count arrives with
            // the first reach's).
            unsynced = c->h_ctl->live > 0;
            if (unsynced) {
                tp = std::chrono::steady_clock::now();
                c->vis_mark = true;  // the first pivots seed the visited bitmaps k_init cleared
                CHK(do_elect(c, 0, &npiv, learned, false));  // first pivot: lowest live id (scc.cu:235-245)
                c->vis_valid = true;
                st.ms_update += ms_since(tp);
            }
        }
        while (unsynced || npiv > 0) {
            // a WCC election left unread, and the next step waits anyway (an election tag
            // wrap): read its pivot count here
            if (unsynced && after_wcc && !(spec_ok && ((c->epoch + 1) & 0xFFFFu) >= 2u)) {
                CHK(read_ctl(c));
                npiv = c->h_ctl->npiv;
                unsynced = after_wcc = false;
                if (npiv == 0) break;
            }
            // The tail: once few vertices are live, the remaining steps run in one
            // workgroup and one launch (k_fb_tail); it stops early only before an
            // election tag would wrap, and the loop then goes on as before.
            if (spec_ok && !unsynced && (variant == SCC_VARIANT_PLAIN || wcc_done) && c->lsel >= 0 &&
                c->opts.tail_live >= 0 && c->h_ctl->live <= tail_max) {
                CHK(run_tail(0));
                if (npiv == 0) break;
            }
            st.steps++;
            c->one_colour = st.steps == 1;  // no recolouring yet: every arc stays inside a colour
            int sched = schedule;
            if (by_elim) sched = ne < one_pct ? (schedule == SCC_SCHED_C ? SCC_SCHED_L : SCC_SCHED_T) : SCC_SCHED_Q;
            c->best_step = schedule == SCC_SCHED_B && st.steps == 1;
            if (c->best_step) sched = SCC_SCHED_H;
            int64_t d = 0;
            tp = std::chrono::steady_clock::now();
            const size_t k = (size_t)st.steps - 1;
            if (spec_ok && ((c->epoch + 1) & 0xFFFFu) >= 2u) {  // no table refill inside the step
                if (!c->spec_ev[0])
                    for (auto &e : c->spec_ev) HIPCHK(c, hipEventCreate(&e));
                int Lr = k < c->spec_reach.size() ? c->spec_reach[k] : c->guess[0];
                int Lt = k < c->spec_trim.size() ? c->spec_trim[k] : c->guess[2];
                Lr = std::max(1, std::min(Lr, 1024));
                Lt = std::max(1, std::min(Lt, 1024));
                HIPCHK(c, hipEventRecord(c->spec_ev[0], c->stream));
                ReachChain rc(c, 3, sched, c->d_piv);
                ctl_op(c, rc.reset_mask() | 1);
                if (c->opts.reach_graphs > 0) {
                    CHK(launch_reach_levels(c, rc, Lr));
                } else {
                    for (int i = 0; i < Lr; i++) rc(i, nullptr, nullptr);
                }
                ctl_guard(c, Lr - 1, 1);
                HIPCHK(c, hipEventRecord(c->spec_ev[1], c->stream));
                TrimChain tc(c, true);
                ctl_op(c, 2 | 1);  // with the reach chain's guard
                for (int i = 0; i < Lt; i++) tc(i, nullptr, nullptr);
                ctl_guard(c, Lt - 1, 2);  // goes out with the election's first reset
                HIPCHK(c, hipEventRecord(c->spec_ev[2], c->stream));
                c->cur_step = (int)st.steps;
                int64_t unused = 0;
                CHK(do_elect(c, 1, &unused, learned, false));
                HIPCHK(c, hipEventRecord(c->spec_ev[3], c->stream));
                CHK(read_ctl(c));
                const Ctl h = *c->h_ctl;
                if (spec_t1) {  // the first step also tells how trim-1 ended
                    spec_t1 = false;
                    if (h.abort == 3) {
                        // trim-1 needs more levels, and nothing after its chain ran: undo the
                        // two elections' bookkeeping, finish trim-1 with waits, start over
                        c->epoch -= 2;  // prev_tab flipped twice: as before
                        c->lsel = c->lsel_prev = -1;
                        c->cur_step = 0;  // the first step's reach starts its local rounds late again
                        st.steps--;
                        ctl_op(c, 512);
                        CHK(do_trim(c, &it, true, Lt1));  // sets the next guess
                        st.trim_iters += it + 1;
                        unsynced = c->h_ctl->live > 0;
                        if (unsynced) {
                            c->vis_mark = true;
                            CHK(do_elect(c, 0, &npiv, learned, false));
                            c->vis_valid = true;
                        }
                        continue;
                    }
                    st.trim_iters += (int64_t)h.depth[2] + 1;
                    c->guess[3] = std::max(1, (int)h.depth[2]);
                    float t1ms = 0;
                    HIPCHK(c, hipEventElapsedTime(&t1ms, c->spec_ev[4], c->spec_ev[5]));
                    st.ms_trim += t1ms;
                    if ((h.abort == 0 ? h.npiv_prev : h.npiv) == 0) {
                        // trim-1 left no live vertex: no pivot, and the step found nothing
                        ctl_op(c, 512);
                        st.steps--;
                        unsynced = false;
                        npiv = 0;
                        continue;
                    }
                }
                const bool elected = h.abort == 0;
                // the election before this step's reach (the step's own election,
                // when it ran, moved those counts to the _prev fields)
                const int64_t np = elected ? h.npiv_prev : h.npiv;
                if (unsynced) {
                    if (!after_wcc) st.first_pivot = (int64_t)(elected ? h.first_piv_prev : h.first_piv);
                    unsynced = false;
                }
                if (np == 0) {
                    ctl_op(c, 512);
                    if (!after_wcc) return fail(c, SCC_E_STATE, "live vertices left but no pivot elected");
                    // the WCC pass left no live vertex: the step found nothing
                    st.steps--;
                    after_wcc = false;
                    npiv = 0;
                    continue;
                }
                after_wcc = false;
                st.pivots_total += np;
                int64_t dr = h.depth[0], dt = h.depth[1];
                if (elected) {
                    float t[3];
                    for (int j = 0; j < 3; j++) HIPCHK(c, hipEventElapsedTime(&t[j], c->spec_ev[j], c->spec_ev[j + 1]));
                    st.ms_fwd += t[0];
                    st.ms_trim += t[1];
                    st.ms_update += t[2];
                    npiv = h.npiv;
                } else {  // undo the election's bookkeeping and finish the step with waits
                    c->epoch--;
                    c->prev_tab ^= 1;
                    c->lsel = c->lsel_prev;  // its update did not run: the list before it holds
                    ctl_op(c, 512);
                    if (h.abort == 1) CHK(do_reach(c, 3, sched, c->d_piv, &dr, Lr));
                    st.ms_fwd += ms_since(tp);
                    tp = std::chrono::steady_clock::now();
                    CHK(do_trim(c, &dt, false, h.abort == 2 ? Lt : 0, true));
                    st.ms_trim += ms_since(tp);
                    tp = std::chrono::steady_clock::now();
                    CHK(do_elect(c, 1, &npiv, learned));
                    st.ms_update += ms_since(tp);
                }
                st.fwd_depths += dr;
                st.bwd_depths += dr;
                st.trim_iters += dt;
                if (c->spec_reach.size() <= k) {
                    c->spec_reach.resize(k + 1, 1);
                    c->spec_trim.resize(k + 1, 1);
                }
                // the longest chains seen: a pull level may reach further in one run than in
                // another (claims made earlier in the same level are seen by later rows), so
                // the depth of a step varies by a level or so; a chain one level too long
                // costs an empty launch, one too short a wait and a resumed chain
                c->spec_reach[k] = std::max(c->spec_reach[k], (int)dr);
                c->spec_trim[k] = std::max(c->spec_trim[k], (int)dt);
                if (variant != SCC_VARIANT_PLAIN && !wcc_done) {
                    ne = (int64_t)(elected ? h.fbelim : c->h_ctl->fbelim);
                    if (ne >= one_pct) {
                        if (ne == n) break;
                        CHK(wcc_pass(false));
                        // the tail right behind the WCC election, gated on the device's live
                        // count; one wait reads the election's pivot count either way
                        if (c->opts.tail_live >= 0 && ((c->epoch + 1) & 0xFFFFu) >= 2u) {
                            CHK(run_tail(tail_max));
                            unsynced = after_wcc = false;
                        }
                    }
                }
                continue;
            }
            if (c->opts.sequential_reach) {  // scc.cu:285-704 order: forward, then backward
                CHK(do_reach(c, 1, sched, c->d_piv, &d));
                st.fwd_depths += d;
                st.ms_fwd += ms_since(tp);
                tp = std::chrono::steady_clock::now();
                CHK(do_reach(c, 2, sched, c->d_piv, &d));
                st.bwd_depths += d;
                st.ms_bwd += ms_since(tp);
            } else {  // both directions in the same level kernels
                CHK(do_reach(c, 3, sched, c->d_piv, &d));
                st.fwd_depths += d;
                st.bwd_depths += d;
                st.ms_fwd += ms_since(tp);
            }
            if (unsynced) {  // the first election's results, read with the reach's
                npiv = c->h_ctl->npiv;
                st.first_pivot = (int64_t)c->h_ctl->first_piv;
                if (npiv == 0) return fail(c, SCC_E_STATE, "live vertices left but no pivot elected");
                unsynced = false;
            }
            st.pivots_total += npiv;
            tp = std::chrono::steady_clock::now();
            CHK(do_trim(c, &it, false, 0, true));
            st.trim_iters += it;
            st.ms_trim += ms_since(tp);
            tp = std::chrono::steady_clock::now();
            c->cur_step = (int)st.steps;
            CHK(do_elect(c, 1, &npiv, learned));
            if (variant != SCC_VARIANT_PLAIN && !wcc_done) {
                CHK(do_count(c, 0, &ne));  // countElimmed (scc.cu:42-51, 787-790)
                if (ne >= one_pct) {       // scc.cu:787-821
                    if (ne == n) { st.pivots_total += npiv; st.ms_update += ms_since(tp); break; }
                    CHK(wcc_pass(true));
                }
            } else if (by_elim && ne < one_pct) {
                CHK(do_count(c, 0, &ne));  // scc.cu:879-890
            }
            st.ms_update += ms_since(tp);
        }
        // canonical labels + #SCC (scc.cu:912-919): the device counts P bits as
        // they are set, and the last wait above read the count.  With the plain
        // rule every pivot is the lowest live id of its colour set, so root is
        // already the minimum of each SCC (trimmed vertices and pairs point to
        // themselves and to the lower id); with the learned rule each root takes its
        // SCC's minimum, kept by the updates in d_mn.
        if (learned) launch_root_to_min(l, c->d_root, c->d_mn, c->d_word, n);
        c->label_in_root = true;
        st.num_scc = (int64_t)c->h_ctl->pcount;
        return SCC_OK;
    };
    s = body();
    if (s == SCC_OK && c->ntiles > 0 && c->h_ctl->terr != 0)  // never: the upload cuts tiles of <= TILE_MAX
        s = fail(c, SCC_E_STATE, "a tile larger than TILE_MAX vertices");
    if (s == SCC_OK && c->d_oid && n > 0) {  // renumbered: the labels into the caller's order
        launch_gather_i32(launch_for(c), c->d_root, c->d_nid, c->d_label, n);
        c->label_in_root = false;
    }
    c->track_mn = false;
    c->one_colour = false;
    c->best_step = false;
    c->vis_mark = c->vis_valid = false;
    c->use_lists = false;  // phase entry points work on states of their own
    c->lsel = c->lsel_prev = -1;
    if (s != SCC_OK) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    st.ms_total = ms_since(t_start);
    st.host_syncs = c->syncs;
    c->labels_valid = true;
    if (labels_out && n > 0) {
        HIPCHK(c, hipMemcpy(labels_out, c->label_in_root ? c->d_root : c->d_label, sizeof(int32_t) * (size_t)n,
                            hipMemcpyDeviceToHost));
        std::vector<int32_t> cnt((size_t)n, 0);
        for (int32_t v = 0; v < n; v++) cnt[labels_out[v]]++;
        for (int32_t v = 0; v < n; v++) {
            if (labels_out[v] != v) continue;
            if (cnt[v] == 1) st.trivial++;
            st.biggest = std::max<int64_t>(st.biggest, cnt[v]);
        }
    }
    if (stats) *stats = st;
    return SCC_OK;
}

scc_status scc_profile_set(scc_ctx *c, int flags) {
    if (!c || flags < 0 || flags > 15) return SCC_E_INPUT;
    if (flags & 4) flags |= 3;
    c->prof_flags = flags;
    c->instr = flags & 1;
    return SCC_OK;
}

scc_status scc_profile_get(scc_ctx *c, scc_profile_data *out, int reset) {
    if (!c || !out) return SCC_E_INPUT;
    *out = c->prof;
    if (reset) std::memset(&c->prof, 0, sizeof(c->prof));
    return SCC_OK;
}

scc_status scc_profile_levels(scc_ctx *c, double *out, int64_t cap, int64_t *rows, int reset) {
    if (!c || !rows || cap < 0 || (cap > 0 && !out)) return SCC_E_INPUT;
    const int64_t have = (int64_t)c->level_log.size() / SCC_LEVEL_FIELDS;
    *rows = have;
    if (cap > 0 && have > 0)
        std::memcpy(out, c->level_log.data(), sizeof(double) * SCC_LEVEL_FIELDS * (size_t)std::min(have, cap));
    if (reset) c->level_log.clear();
    return SCC_OK;
}

scc_status scc_set_batch(scc_ctx *c, const int32_t *off, int32_t G) {
    if (!c) return SCC_E_INPUT;
    if (!c->uploaded) return fail(c, SCC_E_STATE, "scc_set_batch before scc_upload");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->spec_reach.clear();
    c->spec_trim.clear();
    if (c->d_goff) HIPCHK(c, hipFree(c->d_goff));
    c->d_goff = nullptr;
    free_batch(c);
    c->num_graphs = 0;
    if (G == 0) return SCC_OK;
    if (G < 0 || !off) return fail(c, SCC_E_INPUT, "bad batch");
    if (off[0] != 0 || off[G] != c->g.n) return fail(c, SCC_E_INPUT, "offsets must run from 0 to n");
    for (int32_t i = 0; i < G; i++)
        if (off[i + 1] < off[i]) return fail(c, SCC_E_INPUT, "offsets decrease");
    free_tiles(c);  // the tile reach serves single graphs
    if (c->d_oid && c->g.n > 0) {  // renumbered: again, by degree within each graph, so graphs keep their ranges
        const int32_t n = c->g.n;
        std::vector<int32_t> row((size_t)n + 1), rowT((size_t)n + 1), oid((size_t)n), deg((size_t)n);
        HIPCHK(c, hipMemcpy(row.data(), c->d_row, sizeof(int32_t) * row.size(), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(rowT.data(), c->d_rowT, sizeof(int32_t) * rowT.size(), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(oid.data(), c->d_oid, sizeof(int32_t) * oid.size(), hipMemcpyDeviceToHost));
        for (int32_t i = 0; i < n; i++) deg[oid[i]] = (row[i + 1] - row[i]) + (rowT[i + 1] - rowT[i]);
        CHK(permute_graph(c, row.data(), rowT.data(), degree_order(n, deg, off, G), oid));
        row_indexes(c);
        c->labels_valid = false;
    }
    HIPCHK(c, hipMalloc(&c->d_goff, sizeof(int32_t) * ((size_t)G + 1)));
    HIPCHK(c, hipMemcpy(c->d_goff, off, sizeof(int32_t) * ((size_t)G + 1), hipMemcpyHostToDevice));
    if (c->g.n > 0) {
        ctl_op(c, 8);
        launch_batch_check(launch_for(c), c->g, c->d_goff, G, c->d_ctl);
        CHK(read_ctl(c));
        if (c->h_ctl->count != 0) {
            (void)hipFree(c->d_goff);
            c->d_goff = nullptr;
            return fail(c, SCC_E_INPUT, "an arc crosses graphs of the batch");
        }
    }
    // the persistent solver: graphs by decreasing vertex count (the largest start first,
    // so that no workgroup is left with a big graph at the end) and their election tables
    std::vector<int32_t> order((size_t)G);
    std::vector<uint64_t> toff((size_t)G);
    for (int32_t i = 0; i < G; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
    uint64_t words = 0;
    for (int32_t i = 0; i < G; i++) {
        toff[i] = words;
        words += batch_table_words(off[i + 1] - off[i]);
    }
    HIPCHK(c, hipMalloc(&c->d_border, sizeof(int32_t) * (size_t)G));
    HIPCHK(c, hipMalloc(&c->d_btab_off, sizeof(uint64_t) * (size_t)G));
    HIPCHK(c, hipMalloc(&c->d_btab, sizeof(uint32_t) * (size_t)std::max<uint64_t>(words, 1)));
    HIPCHK(c, hipMemcpy(c->d_border, order.data(), sizeof(int32_t) * (size_t)G, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_btab_off, toff.data(), sizeof(uint64_t) * (size_t)G, hipMemcpyHostToDevice));
    c->num_graphs = G;
    return SCC_OK;
}

scc_status scc_mlp_set_params(scc_ctx *c, const float *params, int32_t count) {
    if (!c || !params) return SCC_E_INPUT;
    if (count != MLP_PARAMS) return fail(c, SCC_E_INPUT, "the scorer has 1505 parameters");
    HIPCHK(c, hipSetDevice(c->device));
    CHK(mlp_alloc(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_wf, params, sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->d_m, 0, sizeof(float) * MLP_PARAMS));
    HIPCHK(c, hipMemset(c->d_v, 0, sizeof(float) * MLP_PARAMS));
    launch_to_bf16(c->stream, c->d_wf, c->d_wb);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->adam_t = 0;
    c->have_params = true;
    return SCC_OK;
}

scc_status scc_mlp_get_params(scc_ctx *c, float *params, int32_t count) {
    if (!c || !params) return SCC_E_INPUT;
    if (count != MLP_PARAMS) return fail(c, SCC_E_INPUT, "the scorer has 1505 parameters");
    if (!c->have_params) return fail(c, SCC_E_STATE, "no parameters");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(params, c->d_wf, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost));
    return SCC_OK;
}

scc_status scc_mlp_infer(scc_ctx *c, const float *X, int64_t N, float *logits) {
    if (!c || N < 0 || (N > 0 && (!X || !logits))) return SCC_E_INPUT;
    if (!c->have_params) return fail(c, SCC_E_STATE, "no parameters");
    if (N == 0) return SCC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf dx, dl;
    HIPCHK(c, hipMalloc(&dx.p, sizeof(float) * MLP_IN * (size_t)N));
    HIPCHK(c, hipMalloc(&dl.p, sizeof(float) * (size_t)N));
    HIPCHK(c, hipMemcpy(dx.p, X, sizeof(float) * MLP_IN * (size_t)N, hipMemcpyHostToDevice));
    launch_mlp_infer(mlp_launch(c, N), (const float *)dx.p, N, c->d_wb, (float *)dl.p);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(logits, dl.p, sizeof(float) * (size_t)N, hipMemcpyDeviceToHost));
    return SCC_OK;
}

scc_status scc_mlp_gradient(scc_ctx *c, const float *X, const float *y, int64_t N, float *grad,
                            float *loss) {
    if (!c || N <= 0 || !X || !y || !grad) return SCC_E_INPUT;
    if (!c->have_params) return fail(c, SCC_E_STATE, "no parameters");
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf dx, dy, wb[8];
    MlpWork w;
    CHK(mlp_work_alloc(c, N, w, wb));
    HIPCHK(c, hipMalloc(&dx.p, sizeof(float) * MLP_IN * (size_t)N));
    HIPCHK(c, hipMalloc(&dy.p, sizeof(float) * (size_t)N));
    HIPCHK(c, hipMemcpy(dx.p, X, sizeof(float) * MLP_IN * (size_t)N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dy.p, y, sizeof(float) * (size_t)N, hipMemcpyHostToDevice));
    launch_fill_f32(mlp_launch(c, 1), w.loss, 1, 0.f);
    launch_mlp_fwdbwd(mlp_launch(c, N), (const float *)dx.p, (const float *)dy.p, nullptr, N, c->d_wb, w);
    launch_mlp_grad(c->stream, N, w);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(grad, w.grad, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost));
    if (loss) {
        float ls = 0;
        HIPCHK(c, hipMemcpy(&ls, w.loss, sizeof(float), hipMemcpyDeviceToHost));
        *loss = ls / (float)N;
    }
    return SCC_OK;
}

scc_status scc_mlp_train(scc_ctx *c, const float *X, const float *y, int64_t N, int32_t epochs,
                         int32_t batch, float lr, uint64_t seed, float *epoch_loss) {
    if (!c || N <= 0 || !X || !y || epochs < 0 || batch <= 0 || !(lr > 0.f)) return SCC_E_INPUT;
    if (!c->have_params) return fail(c, SCC_E_STATE, "initialise the parameters first (scc_mlp_set_params)");
    HIPCHK(c, hipSetDevice(c->device));
    int64_t B = std::min<int64_t>(batch, N);
    DevBuf dx, dy, di, wb[8];
    MlpWork w;
    CHK(mlp_work_alloc(c, B, w, wb));
    HIPCHK(c, hipMalloc(&dx.p, sizeof(float) * MLP_IN * (size_t)N));
    HIPCHK(c, hipMalloc(&dy.p, sizeof(float) * (size_t)N));
    HIPCHK(c, hipMalloc(&di.p, sizeof(int32_t) * (size_t)N));
    HIPCHK(c, hipMemcpy(dx.p, X, sizeof(float) * MLP_IN * (size_t)N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dy.p, y, sizeof(float) * (size_t)N, hipMemcpyHostToDevice));
    std::vector<int32_t> perm((size_t)N);
    for (int64_t i = 0; i < N; i++) perm[i] = (int32_t)i;
    uint64_t rs = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {  // xorshift64*
        rs ^= rs >> 12; rs ^= rs << 25; rs ^= rs >> 27;
        return rs * 0x2545F4914F6CDD1Dull;
    };
    for (int32_t ep = 0; ep < epochs; ep++) {
        for (int64_t i = N - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
        HIPCHK(c, hipMemcpy(di.p, perm.data(), sizeof(int32_t) * (size_t)N, hipMemcpyHostToDevice));
        launch_fill_f32(mlp_launch(c, 1), w.loss, 1, 0.f);
        int64_t nb = 0;
        for (int64_t b0 = 0; b0 + B <= N; b0 += B, nb++) {
            launch_mlp_fwdbwd(mlp_launch(c, B), (const float *)dx.p, (const float *)dy.p,
                              (const int32_t *)di.p + b0, B, c->d_wb, w);
            launch_mlp_grad(c->stream, B, w);
            launch_adam(c->stream, c->d_wf, w.grad, c->d_m, c->d_v, c->d_wb, lr, 0.9f, 0.999f, 1e-8f,
                        ++c->adam_t);
        }
        HIPCHK(c, hipGetLastError());
        float ls = 0;
        HIPCHK(c, hipMemcpyAsync(&ls, w.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (epoch_loss) epoch_loss[ep] = nb > 0 ? ls / (float)(nb * B) : 0.f;
    }
    return SCC_OK;
}

scc_status scc_phase_recolour(scc_ctx *c) {
    if (!c) return SCC_E_INPUT;
    CHK(detach_labels(c));
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    if (c->g.n == 0) return SCC_OK;
    Launch l = launch_for(c);
    unsigned long long *prev = c->d_slots[c->prev_tab], *cur = c->d_slots[c->prev_tab ^ 1];
    launch_update(l, c->d_word, c->d_root, c->g.n, prev, cur, c->slot_bits, elect_tag(c->epoch), 1, 0,
                  nullptr, c->d_ctl, nullptr, nullptr, nullptr, nullptr, nullptr, c->d_oid);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SCC_OK;
}

scc_status scc_phase_features(scc_ctx *c, int32_t step, const int32_t *idx, int64_t count, float *X_out) {
    if (!c || count < 0 || (count > 0 && (!idx || !X_out))) return SCC_E_INPUT;
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    if (c->g.n == 0 || count == 0) return SCC_OK;
    for (int64_t i = 0; i < count; i++)
        if (idx[i] < 0 || idx[i] >= c->g.n) return fail(c, SCC_E_INPUT, "vertex index out of range");
    if (!c->d_feat) HIPCHK(c, hipMalloc(&c->d_feat, sizeof(float) * MLP_IN * (size_t)c->g.n));
    c->cur_step = step;
    CHK(do_features(c, false, c->d_feat));
    DevBuf di, dx;
    HIPCHK(c, hipMalloc(&di.p, sizeof(int32_t) * (size_t)count));
    HIPCHK(c, hipMalloc(&dx.p, sizeof(float) * MLP_IN * (size_t)count));
    HIPCHK(c, hipMemcpyAsync(di.p, idx, sizeof(int32_t) * (size_t)count, hipMemcpyHostToDevice, c->stream));
    if (c->d_oid) {  // renumbered: the caller's ids to internal ones, in place through dx's first count words
        launch_gather_i32(launch_for(c), c->d_nid, (const int32_t *)di.p, (int32_t *)dx.p, count);
        HIPCHK(c, hipMemcpyAsync(di.p, dx.p, sizeof(int32_t) * (size_t)count, hipMemcpyDeviceToDevice, c->stream));
    }
    launch_gather_rows(mlp_launch(c, count * MLP_IN), c->d_feat, (const int32_t *)di.p, count, (float *)dx.p);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(X_out, dx.p, sizeof(float) * MLP_IN * (size_t)count, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SCC_OK;
}

scc_status scc_profile_copy(scc_ctx *c, int64_t bytes) {
    if (!c || bytes <= 0 || (bytes % 16) != 0) return SCC_E_INPUT;
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf src, dst;
    HIPCHK(c, hipMalloc(&src.p, (size_t)bytes));
    HIPCHK(c, hipMalloc(&dst.p, (size_t)bytes));
    HIPCHK(c, hipMemsetAsync(src.p, 1, (size_t)bytes, c->stream));
    Launch l{c->stream, (uint32_t)c->cu_count * 8u, 256};
    launch_stream_copy(l, src.p, dst.p, bytes);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SCC_OK;
}

scc_status scc_labels_device(scc_ctx *c, const int32_t **dptr) {
    if (!c || !dptr) return SCC_E_INPUT;
    if (!c->labels_valid) return fail(c, SCC_E_STATE, "no labels yet");
    *dptr = c->label_in_root ? c->d_root : c->d_label;
    return SCC_OK;
}

// ---- phase-level entry points ---------------------------------------------
scc_status scc_state_set(scc_ctx *c, const uint32_t *words, const int32_t *roots) {
    if (!c || !words) return SCC_E_INPUT;
    CHK(detach_labels(c));
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    size_t vb = sizeof(uint32_t) * (size_t)c->g.n;
    ctl_op(c, 512);  // phase-level use starts from a clean control block
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (vb == 0) return SCC_OK;
    const int32_t n = c->g.n;
    std::vector<int32_t> oid;  // renumbered: internal vertex i is the caller's oid[i]
    if (c->d_oid) {
        oid.resize((size_t)n);
        HIPCHK(c, hipMemcpy(oid.data(), c->d_oid, vb, hipMemcpyDeviceToHost));
    }
    auto at = [&](int32_t i) { return oid.empty() ? i : oid[i]; };
    std::vector<uint32_t> w((size_t)n);
    std::vector<int32_t> r((size_t)n);
    for (int32_t i = 0; i < n; i++) {
        w[i] = words[at(i)];
        r[i] = roots ? roots[at(i)] : at(i);
    }
    HIPCHK(c, hipMemcpy(c->d_word, w.data(), vb, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_root, r.data(), vb, hipMemcpyHostToDevice));
    return SCC_OK;
}

scc_status scc_state_get(scc_ctx *c, uint32_t *words, int32_t *roots) {
    if (!c) return SCC_E_INPUT;
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    size_t vb = sizeof(uint32_t) * (size_t)c->g.n;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (vb == 0) return SCC_OK;
    if (words) HIPCHK(c, hipMemcpy(words, c->d_word, vb, hipMemcpyDeviceToHost));
    if (roots) HIPCHK(c, hipMemcpy(roots, c->d_root, vb, hipMemcpyDeviceToHost));
    if (c->d_oid) {  // renumbered: into the caller's order
        const int32_t n = c->g.n;
        std::vector<int32_t> oid((size_t)n);
        HIPCHK(c, hipMemcpy(oid.data(), c->d_oid, vb, hipMemcpyDeviceToHost));
        if (words) {
            std::vector<uint32_t> t(words, words + n);
            for (int32_t i = 0; i < n; i++) words[oid[i]] = t[i];
        }
        if (roots) {
            std::vector<int32_t> t(roots, roots + n);
            for (int32_t i = 0; i < n; i++) roots[oid[i]] = t[i];
        }
    }
    return SCC_OK;
}

scc_status scc_phase_trim(scc_ctx *c, int64_t *iterations) {
    if (!c) return SCC_E_INPUT;
    CHK(detach_labels(c));
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    int64_t it = 0;
    if (c->g.n > 0) CHK(do_trim(c, &it));
    if (iterations) *iterations = it;
    return SCC_OK;
}

scc_status scc_phase_reach(scc_ctx *c, int dir, int schedule, int64_t *depth) {
    if (!c) return SCC_E_INPUT;
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    if (dir < 0 || dir > 2) return fail(c, SCC_E_INPUT, "dir must be 0, 1 or 2");
    if (schedule < SCC_SCHED_Q || schedule > SCC_SCHED_H) return fail(c, SCC_E_INPUT, "bad schedule");
    int64_t d = 0;
    uint32_t dirs = dir == 0 ? 1u : dir == 1 ? 2u : 3u;
    if (c->g.n > 0) {
        if (schedule >= SCC_SCHED_D) {  // seeds: live vertices holding the bit
            ctl_op(c, 128);
            if (dirs & 1u) launch_collect(launch_for(c), c->d_word, c->g.n, ST_F, c->d_piv, c->d_ctl);
            if (dirs & 2u) launch_collect(launch_for(c), c->d_word, c->g.n, ST_B, c->d_seedB, c->d_ctl);
        }
        CHK(do_reach(c, dirs, schedule, c->d_seedB, &d));
    }
    if (depth) *depth = d;
    return SCC_OK;
}

scc_status scc_phase_update(scc_ctx *c, int recolour, int64_t *pivots) {
    if (!c) return SCC_E_INPUT;
    CHK(detach_labels(c));
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    int64_t np = 0;
    if (c->g.n > 0) CHK(do_elect(c, recolour ? 1 : 0, &np));
    if (pivots) *pivots = np;
    return SCC_OK;
}

scc_status scc_phase_trim2(scc_ctx *c, int64_t *pairs) {
    if (!c) return SCC_E_INPUT;
    CHK(detach_labels(c));
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    int64_t p = 0;
    if (c->g.n > 0) CHK(do_trim2(c, &p));
    if (pairs) *pairs = p;
    return SCC_OK;
}

scc_status scc_phase_wcc(scc_ctx *c, int64_t *iterations, int64_t *pivots) {
    if (!c) return SCC_E_INPUT;
    CHK(detach_labels(c));
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    int64_t it = 0, np = 0;
    if (c->g.n > 0) CHK(do_wcc(c, &it, &np));
    if (iterations) *iterations = it;
    if (pivots) *pivots = np;
    return SCC_OK;
}

scc_status scc_phase_unset_pivots(scc_ctx *c) {
    if (!c) return SCC_E_INPUT;
    CHK(detach_labels(c));
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    if (c->g.n == 0) return SCC_OK;
    launch_unset_pivots(launch_for(c), c->d_word, c->g.n, c->d_ctl);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SCC_OK;
}

scc_status scc_test_fail_alloc(scc_ctx *c, int32_t nth) {
    if (!c || nth < 0) return SCC_E_INPUT;
    c->fail_alloc_at = nth;
    c->alloc_count = 0;
    return SCC_OK;
}

scc_status scc_phase_count_elimmed(scc_ctx *c, int64_t *count) {
    if (!c || !count) return SCC_E_INPUT;
    if (!c->uploaded) return fail(c, SCC_E_STATE, "no graph");
    *count = 0;
    if (c->g.n > 0) CHK(do_count(c, 0, count));
    return SCC_OK;
}

}  // extern "C"
