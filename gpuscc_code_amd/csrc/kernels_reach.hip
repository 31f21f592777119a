// Forward / backward reachability inside colours for gfx950.
//
// One kernel per BFS level, chained on the device (a level whose predecessor
// found nothing exits at once).  The reference runs the forward reach to its
// fixpoint and then the backward one (scc.cu:285-704); here one level kernel
// advances both: the forward frontier over the CSR (bit F) and the backward
// frontier over CSR^T (bit B) share each launch, so a step costs
// max(depth_F, depth_B) launches instead of their sum.  Schedules:
//   q  sweep, one lane per vertex                 (reach_q.cu:12-79)
//   t  sweep, workgroup-balanced edge expansion   (reach_t.cu:17-270)
//   d  frontier queues, one lane per item         (reach_d.cu:11-64)
//   l  frontier queues, workgroup-balanced edges  (reach_l.cu:53-253)
//   h  per level and direction: balanced queue expansion (push) while the
//      frontier is small, bottom-up (pull) once it is large: every live
//      unreached vertex looks for one reached in-neighbour and stops at the
//      first (scc.cu:493-563 switches by eliminated count instead)
//
// Every visit is claimed with one atomicOr on the target's word, whose old
// value says who set the bit first, so a vertex enters a queue exactly once:
// queues of n entries cannot overflow (the reference's can, worklistc.h:80).
// Pushes go through per-direction LDS staging buffers, one global atomicAdd
// per flush.
#include "fb_device.hpp"
#include "kernels.hpp"

namespace gpuscc {

constexpr uint32_t SQ_CAP = 1024;  // LDS staging entries per direction per workgroup

__device__ inline uint32_t dbit(int d) { return d == 0 ? ST_F : ST_B; }
__device__ inline uint32_t dxbit(int d) { return d == 0 ? ST_FX : ST_BX; }
__device__ inline const int32_t *drow(const ReachArgs &a, int d) { return d == 0 ? a.row : a.rowT; }
__device__ inline const int32_t *dcol(const ReachArgs &a, int d) { return d == 0 ? a.col : a.colT; }

__device__ inline bool try_claim(uint32_t *word, int32_t u, uint32_t c, uint32_t bit) {
    uint32_t wu = word[u];
    if ((wu & (bit | ST_E)) != 0 || colour_of(wu) != c) return false;
    uint32_t old = atomicOr(&word[u], bit);
    return !(old & bit);
}

// ---------------------------------------------------------------------------
// LDS staging of the two output queues
// ---------------------------------------------------------------------------
struct Stage {
    int32_t buf[2][SQ_CAP];
    uint32_t cnt[2];
    uint32_t base[2];
};

__device__ inline void st_init(Stage &s) {
    if (threadIdx.x < 2) s.cnt[threadIdx.x] = 0;
}

__device__ inline void st_push(Stage &s, int d, int32_t v, const ReachArgs &a) {
    uint32_t i = atomicAdd(&s.cnt[d], 1u);
    if (i < SQ_CAP) {
        s.buf[d][i] = v;
    } else {
        uint32_t g = atomicAdd(&a.ctl->qsize[d][(a.iter + 1) & 3], 1u);
        a.outq[d][g] = v;
    }
}

// all work-items of the workgroup call this
__device__ inline void st_flush(Stage &s, const ReachArgs &a) {
    __syncthreads();
    uint32_t c0 = s.cnt[0] < SQ_CAP ? s.cnt[0] : SQ_CAP;
    uint32_t c1 = s.cnt[1] < SQ_CAP ? s.cnt[1] : SQ_CAP;
    if (threadIdx.x == 0 && c0) s.base[0] = atomicAdd(&a.ctl->qsize[0][(a.iter + 1) & 3], c0);
    if (threadIdx.x == 1 && c1) s.base[1] = atomicAdd(&a.ctl->qsize[1][(a.iter + 1) & 3], c1);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < c0; i += blockDim.x) a.outq[0][s.base[0] + i] = s.buf[0][i];
    for (uint32_t i = threadIdx.x; i < c1; i += blockDim.x) a.outq[1][s.base[1] + i] = s.buf[1][i];
    __syncthreads();
    if (threadIdx.x < 2) s.cnt[threadIdx.x] = 0;
    __syncthreads();
}

// all work-items call this; flushes when another pass could overflow a buffer
__device__ inline void st_maybe_flush(Stage &s, const ReachArgs &a) {
    __syncthreads();
    bool full = s.cnt[0] > SQ_CAP - blockDim.x || s.cnt[1] > SQ_CAP - blockDim.x;
    __syncthreads();  // every work-item has read the counts before anyone pushes again
    if (full) st_flush(s, a);
}

__device__ inline void instr_flush(Ctl *ctl, unsigned long long x, unsigned long long e,
                                   unsigned long long v, unsigned long long *sh) {
    x = block_sum(x, sh);
    if (threadIdx.x == 0 && x) atomicAdd(&ctl->st_expanded, x);
    e = block_sum(e, sh);
    if (threadIdx.x == 0 && e) atomicAdd(&ctl->st_edges, e);
    v = block_sum(v, sh);
    if (threadIdx.x == 0 && v) atomicAdd(&ctl->st_visited, v);
}

// exclusive prefix sum over the workgroup (blockDim.x <= 1024, wave64)
__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t *shw, uint32_t &total) {
    int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    uint32_t inc = x;
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, WAVE);
        if (lane >= o) inc += y;
    }
    __syncthreads();
    if (lane == WAVE - 1) shw[wid] = inc;
    __syncthreads();
    int nw = (blockDim.x + WAVE - 1) / WAVE;
    uint32_t pre = 0, tot = 0;
    for (int i = 0; i < nw; i++) {
        uint32_t s = shw[i];
        if (i < wid) pre += s;
        tot += s;
    }
    total = tot;
    return pre + inc - x;
}

// largest k < cnt with off[k] <= e (off[0] == 0 <= e)
__device__ inline uint32_t find_item(const uint32_t *off, uint32_t cnt, uint32_t e) {
    uint32_t lo = 0, hi = cnt;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// balanced expansion of the items held in LDS: item k owns arcs
// [start[k], start[k] + deg[k]) of direction cd[k] >> 31; work-item t takes
// arcs t, t + blockDim, ... of their concatenation (the CTA gather of
// reach_t.cu:195-265 and reach_l.cu:190-245, without their 256-arc windows)
// ---------------------------------------------------------------------------
struct Items {
    uint32_t off[1024];
    int32_t start[1024];
    uint32_t cd[1024];  // colour | direction << 31
    int32_t lim[1024];  // end of the row (huge rows)
    uint32_t shw[16];
};

template <bool PUSH>
__device__ inline void expand_items(const ReachArgs &a, Items &it, uint32_t nitems, uint32_t total,
                                    Stage *st, bool &found, unsigned long long &cv) {
    for (uint32_t base = 0; base < total; base += blockDim.x) {  // block-uniform trip count
        uint32_t e = base + threadIdx.x;
        if (e < total) {
            uint32_t k = find_item(it.off, nitems, e);
            uint32_t cd = it.cd[k];
            int d = (int)(cd >> 31);
            int32_t u = dcol(a, d)[it.start[k] + (int32_t)(e - it.off[k])];
            if (try_claim(a.word, u, cd & 0x7FFFFFFFu, dbit(d))) {
                found = true;
                cv++;
                if (PUSH) st_push(*st, d, u, a);
            }
        }
        if (PUSH) st_maybe_flush(*st, a);
    }
}

// Gate of a reach level: exits when the previous level made no progress;
// block 0 clears the output counters of the level after next.
__device__ inline bool level_begin(const ReachArgs &a, bool queues) {
    if (chain_exit(a.ctl, a.iter)) return false;
    if (blockIdx.x == 0 && threadIdx.x < 2) {
        a.ctl->hsize[threadIdx.x][(a.iter + 2) & 3] = 0;
        if (queues) a.ctl->qsize[threadIdx.x][(a.iter + 2) & 3] = 0;
    }
    return true;
}

// a row of at least huge_thr arcs is not expanded by its own lane or
// workgroup: it is deferred to the next level, where all workgroups split it
__device__ inline void push_huge(const ReachArgs &a, int d, int32_t v) {
    uint32_t g = atomicAdd(&a.ctl->hsize[d][(a.iter + 1) & 3], 1u);
    a.hout[d][g] = v;
}

constexpr uint32_t HCHUNK = 1024;  // arcs per chunk of a huge row

// Expands the huge rows deferred by the previous level: their arcs are cut
// into chunks of HCHUNK, numbered across the list, and workgroup b takes the
// chunks k with k mod gridDim == b.  All work-items call this.
template <bool PUSH>
__device__ inline void huge_phase(const ReachArgs &a, Items &it, Stage *st, bool &found,
                                  unsigned long long &cx, unsigned long long &ce,
                                  unsigned long long &cv) {
    const uint32_t G = gridDim.x;
    for (int d = 0; d < 2; d++) {
        if (!((a.dirs >> d) & 1u)) continue;
        uint32_t hs = a.ctl->hsize[d][a.iter & 3];
        if (hs == 0) continue;
        const int32_t *row = drow(a, d), *col = dcol(a, d);
        uint32_t bit = dbit(d);
        uint32_t acc = 0;
        for (uint32_t base = 0; base < hs; base += blockDim.x) {
            uint32_t i = base + threadIdx.x;
            uint32_t nch = 0, c = 0;
            int32_t e0 = 0, e1 = 0;
            if (i < hs) {
                int32_t v = a.hin[d][i];
                e0 = row[v];
                e1 = row[v + 1];
                nch = ((uint32_t)(e1 - e0) + HCHUNK - 1) / HCHUNK;
                c = colour_of(a.word[v]);
                if (blockIdx.x == 0) {
                    cx++;
                    ce += (unsigned long long)(e1 - e0);
                }
            }
            uint32_t total;
            uint32_t off = block_excl_scan(nch, it.shw, total);
            it.off[threadIdx.x] = off;
            it.start[threadIdx.x] = e0;
            it.lim[threadIdx.x] = e1;
            it.cd[threadIdx.x] = c;
            __syncthreads();
            for (uint32_t k = acc + (blockIdx.x + G - acc % G) % G; k < acc + total; k += G) {
                uint32_t local = k - acc;
                uint32_t j = find_item(it.off, blockDim.x, local);
                int32_t cs = it.start[j] + (int32_t)((local - it.off[j]) * HCHUNK);
                int32_t cend = cs + (int32_t)HCHUNK < it.lim[j] ? cs + (int32_t)HCHUNK : it.lim[j];
                uint32_t cc = it.cd[j];
                for (int32_t pb = cs; pb < cend; pb += (int32_t)blockDim.x) {  // block-uniform
                    int32_t e = pb + (int32_t)threadIdx.x;
                    if (e < cend && try_claim(a.word, col[e], cc, bit)) {
                        found = true;
                        cv++;
                        if (PUSH) st_push(*st, d, col[e], a);
                    }
                    if (PUSH) st_maybe_flush(*st, a);
                }
            }
            acc += total;
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// q: one lane per vertex; a live vertex with the bit and without the expanded
// bit sets the expanded bit and scans its row serially (reach_q.cu:44-79)
// ---------------------------------------------------------------------------
__global__ void k_reach_sweep(ReachArgs a) {
    __shared__ Items it;
    __shared__ unsigned long long sh[16];
    if (!level_begin(a, false)) return;
    int64_t beg, end;
    my_range(a.n, beg, end);
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    huge_phase<false>(a, it, nullptr, found, cx, ce, cv);
    for (int64_t vv = beg + threadIdx.x; vv < end; vv += blockDim.x) {
        int32_t v = (int32_t)vv;
        uint32_t w = a.word[v];
        if (w & ST_E) continue;
        uint32_t c = colour_of(w);
        for (int d = 0; d < 2; d++) {
            uint32_t bit = dbit(d), xbit = dxbit(d);
            if (!((a.dirs >> d) & 1u) || (w & (bit | xbit)) != bit) continue;
            atomicOr(&a.word[v], xbit);
            const int32_t *row = drow(a, d), *col = dcol(a, d);
            int32_t e0 = row[v], e1 = row[v + 1];
            if ((uint32_t)(e1 - e0) >= a.huge_thr) {
                push_huge(a, d, v);
                found = true;
                continue;
            }
            cx++;
            ce += (unsigned long long)(e1 - e0);
            for (int32_t e = e0; e < e1; e++) {
                if (try_claim(a.word, col[e], c, bit)) {
                    found = true;
                    cv++;
                }
            }
        }
    }
    chain_progress(a.ctl, a.iter, found);
    if (a.instr) instr_flush(a.ctl, cx, ce, cv, sh);
}

// ---------------------------------------------------------------------------
// t: sweep with balanced expansion: the workgroup's 256 vertices contribute up
// to 512 items (forward and backward rows), expanded arc-parallel
// ---------------------------------------------------------------------------
__global__ void k_reach_topo(ReachArgs a) {
    __shared__ Items it;
    __shared__ unsigned long long sh[16];
    if (!level_begin(a, false)) return;
    int64_t beg, end;
    my_range(a.n, beg, end);
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    huge_phase<false>(a, it, nullptr, found, cx, ce, cv);
    for (int64_t base = beg; base < end; base += blockDim.x) {
        int64_t vv = base + threadIdx.x;
        uint32_t deg[2] = {0, 0}, c = 0;
        int32_t start[2] = {0, 0};
        if (vv < end) {
            int32_t v = (int32_t)vv;
            uint32_t w = a.word[v];
            if (!(w & ST_E)) {
                c = colour_of(w);
                uint32_t xb = 0;
                for (int d = 0; d < 2; d++) {
                    if (!((a.dirs >> d) & 1u) || (w & (dbit(d) | dxbit(d))) != dbit(d)) continue;
                    xb |= dxbit(d);
                    const int32_t *row = drow(a, d);
                    start[d] = row[v];
                    deg[d] = (uint32_t)(row[v + 1] - start[d]);
                    if (deg[d] >= a.huge_thr) {
                        push_huge(a, d, v);
                        found = true;
                        deg[d] = 0;
                        continue;
                    }
                    cx++;
                    ce += deg[d];
                }
                if (xb) atomicOr(&a.word[v], xb);
            }
        }
        uint32_t total;
        uint32_t off = block_excl_scan(deg[0] + deg[1], it.shw, total);
        it.off[2 * threadIdx.x] = off;
        it.off[2 * threadIdx.x + 1] = off + deg[0];
        it.start[2 * threadIdx.x] = start[0];
        it.start[2 * threadIdx.x + 1] = start[1];
        it.cd[2 * threadIdx.x] = c;
        it.cd[2 * threadIdx.x + 1] = c | 0x80000000u;
        __syncthreads();
        if (total > 0) expand_items<false>(a, it, 2 * blockDim.x, total, nullptr, found, cv);
        __syncthreads();
    }
    chain_progress(a.ctl, a.iter, found);
    if (a.instr) instr_flush(a.ctl, cx, ce, cv, sh);
}

// ---------------------------------------------------------------------------
// d: frontier queues, one lane per item, serial row scan (reach_d.cu:36-64);
// the item space is the forward queue followed by the backward queue, dealt to
// workgroups in contiguous chunks.  After its chunk, a workgroup keeps
// expanding the vertices it discovered itself, in LDS, for up to
// local_rounds rounds: reachability is monotone, so reaching a vertex before
// its BFS level changes only the number of levels.  What is left (and any
// staging overflow) goes to the next level's queue.
// ---------------------------------------------------------------------------
struct LocalQ {
    int32_t v[2 * SQ_CAP];
};

// expands item v of direction d serially; new visits are staged
__device__ inline void expand_serial(const ReachArgs &a, Stage &st, int d, int32_t v, bool &found,
                                     unsigned long long &cx, unsigned long long &ce,
                                     unsigned long long &cv) {
    uint32_t c = colour_of(a.word[v]), bit = dbit(d);
    const int32_t *row = drow(a, d), *col = dcol(a, d);
    int32_t e0 = row[v], e1 = row[v + 1];
    if ((uint32_t)(e1 - e0) >= a.huge_thr) {
        push_huge(a, d, v);
        found = true;
        return;
    }
    cx++;
    ce += (unsigned long long)(e1 - e0);
    for (int32_t e = e0; e < e1; e++) {
        int32_t u = col[e];
        if (try_claim(a.word, u, c, bit)) {
            found = true;
            cv++;
            st_push(st, d, u, a);
        }
    }
}

constexpr int32_t LOCAL_MAX_DEG = 64;       // longer rows leave the local rounds for the next level
constexpr uint32_t LOCAL_ITEMS_PER_WG = 8;  // local rounds only when a level's frontier is this thin
constexpr int LOCAL_FIRST_LEVEL = 4;        // ... and the reach is this many levels deep

// Workgroup-local rounds on the visits staged by this workgroup: each round
// moves the staged vertices into lq and expands them one per lane; rows longer
// than LOCAL_MAX_DEG go to the next level's queue instead.  All work-items call this.
__device__ inline void local_rounds(const ReachArgs &a, uint32_t level_items, Stage &st, LocalQ &lq,
                                    bool &found, unsigned long long &cx, unsigned long long &ce,
                                    unsigned long long &cv) {
    // a thick frontier keeps every workgroup busy without them; a thin one
    // (high-diameter graphs) would leave most workgroups idle per level.  The
    // first levels of a reach are thin on every graph, so they never qualify.
    if (a.iter < LOCAL_FIRST_LEVEL || level_items > LOCAL_ITEMS_PER_WG * gridDim.x) return;
    for (int r = 0; r < a.local_rounds; r++) {
        __syncthreads();
        uint32_t c0 = st.cnt[0] < SQ_CAP ? st.cnt[0] : SQ_CAP;
        uint32_t c1 = st.cnt[1] < SQ_CAP ? st.cnt[1] : SQ_CAP;
        if (c0 + c1 == 0) break;  // block-uniform: read after the barrier
        for (uint32_t k = threadIdx.x; k < c0; k += blockDim.x) lq.v[k] = st.buf[0][k];
        for (uint32_t k = threadIdx.x; k < c1; k += blockDim.x) lq.v[SQ_CAP + k] = st.buf[1][k];
        __syncthreads();
        if (threadIdx.x < 2) st.cnt[threadIdx.x] = 0;
        __syncthreads();
        for (uint32_t k0 = 0; k0 < c0 + c1; k0 += blockDim.x) {
            uint32_t k = k0 + threadIdx.x;
            if (k < c0 + c1) {
                int d = k < c0 ? 0 : 1;
                int32_t v = d == 0 ? lq.v[k] : lq.v[SQ_CAP + (k - c0)];
                const int32_t *row = drow(a, d);
                if (row[v + 1] - row[v] > LOCAL_MAX_DEG) {
                    uint32_t g = atomicAdd(&a.ctl->qsize[d][(a.iter + 1) & 3], 1u);
                    a.outq[d][g] = v;
                } else {
                    expand_serial(a, st, d, v, found, cx, ce, cv);
                }
            }
            st_maybe_flush(st, a);
        }
    }
}

__global__ void k_reach_queue(ReachArgs a) {
    __shared__ Items it;
    __shared__ Stage st;
    __shared__ LocalQ lq;
    __shared__ unsigned long long sh[16];
    if (!level_begin(a, true)) return;
    st_init(st);
    __syncthreads();
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    huge_phase<true>(a, it, &st, found, cx, ce, cv);
    st_flush(st, a);  // the queue loop below may not run
    uint32_t qs0 = (a.dirs & 1u) ? a.ctl->qsize[0][a.iter & 3] : 0u;
    uint32_t qs1 = (a.dirs & 2u) ? a.ctl->qsize[1][a.iter & 3] : 0u;
    uint32_t tot = qs0 + qs1;
    uint32_t chunk = (tot + gridDim.x - 1) / gridDim.x;
    uint32_t b0 = blockIdx.x * chunk, b1 = b0 + chunk < tot ? b0 + chunk : tot;
    for (uint32_t base = b0; base < b1; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        if (i < b1) {
            int d = i < qs0 ? 0 : 1;
            int32_t v = d == 0 ? a.inq[0][i] : a.inq[1][i - qs0];
            expand_serial(a, st, d, v, found, cx, ce, cv);
        }
        st_maybe_flush(st, a);
    }
    local_rounds(a, tot, st, lq, found, cx, ce, cv);
    st_flush(st, a);
    chain_progress(a.ctl, a.iter, found);
    if (a.instr) instr_flush(a.ctl, cx, ce, cv, sh);
}

// balanced expansion of queue items [0, tot): forward items first (qs0 of them)
__device__ inline void queue_balanced(const ReachArgs &a, uint32_t qs0, uint32_t tot, Items &it,
                                      Stage &st, bool &found, unsigned long long &cx,
                                      unsigned long long &ce, unsigned long long &cv) {
    uint32_t chunk = (tot + gridDim.x - 1) / gridDim.x;
    uint32_t b0 = blockIdx.x * chunk, b1 = b0 + chunk < tot ? b0 + chunk : tot;
    for (uint32_t base = b0; base < b1; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        uint32_t deg = 0, cd = 0;
        int32_t start = 0;
        if (i < b1) {
            int d = i < qs0 ? 0 : 1;
            int32_t v = d == 0 ? a.inq[0][i] : a.inq[1][i - qs0];
            const int32_t *row = drow(a, d);
            start = row[v];
            deg = (uint32_t)(row[v + 1] - start);
            cd = colour_of(a.word[v]) | ((uint32_t)d << 31);
            if (deg >= a.huge_thr) {
                push_huge(a, d, v);
                found = true;
                deg = 0;
            } else {
                cx++;
                ce += deg;
            }
        }
        uint32_t total;
        uint32_t off = block_excl_scan(deg, it.shw, total);
        it.off[threadIdx.x] = off;
        it.start[threadIdx.x] = start;
        it.cd[threadIdx.x] = cd;
        __syncthreads();
        if (total > 0) expand_items<true>(a, it, blockDim.x, total, &st, found, cv);
        st_maybe_flush(st, a);
    }
}

// ---------------------------------------------------------------------------
// l: frontier queues with balanced expansion (reach_l.cu:151-253)
// ---------------------------------------------------------------------------
__global__ void k_reach_lb(ReachArgs a) {
    __shared__ Items it;
    __shared__ Stage st;
    __shared__ LocalQ lq;
    __shared__ unsigned long long sh[16];
    if (!level_begin(a, true)) return;
    st_init(st);
    __syncthreads();
    uint32_t qs0 = (a.dirs & 1u) ? a.ctl->qsize[0][a.iter & 3] : 0u;
    uint32_t qs1 = (a.dirs & 2u) ? a.ctl->qsize[1][a.iter & 3] : 0u;
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    huge_phase<true>(a, it, &st, found, cx, ce, cv);
    st_flush(st, a);
    queue_balanced(a, qs0, qs0 + qs1, it, st, found, cx, ce, cv);
    local_rounds(a, qs0 + qs1, st, lq, found, cx, ce, cv);
    st_flush(st, a);
    chain_progress(a.ctl, a.iter, found);
    if (a.instr) instr_flush(a.ctl, cx, ce, cv, sh);
}

// ---------------------------------------------------------------------------
// h: a direction whose frontier holds fewer than pull_thr vertices expands its
// queue (balanced push); a larger one runs bottom-up: every live vertex of
// the direction that is not reached scans its reverse row for a reached
// same-colour neighbour and stops at the first.  Pulled visits are queued too,
// so the next level may push again.
// ---------------------------------------------------------------------------
__global__ void k_reach_hybrid(ReachArgs a) {
    __shared__ Items it;
    __shared__ Stage st;
    __shared__ LocalQ lq;
    __shared__ unsigned long long sh[16];
    if (!level_begin(a, true)) return;
    st_init(st);
    __syncthreads();
    uint32_t qs0 = (a.dirs & 1u) ? a.ctl->qsize[0][a.iter & 3] : 0u;
    uint32_t qs1 = (a.dirs & 2u) ? a.ctl->qsize[1][a.iter & 3] : 0u;
    bool pull0 = qs0 > 0 && qs0 >= a.pull_thr, pull1 = qs1 > 0 && qs1 >= a.pull_thr;
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    huge_phase<true>(a, it, &st, found, cx, ce, cv);
    st_flush(st, a);
    // push part: queue items of the directions that do not pull
    uint32_t p0 = pull0 ? 0u : qs0, p1 = pull1 ? 0u : qs1;
    if (p0 + p1 > 0) queue_balanced(a, p0, p0 + p1, it, st, found, cx, ce, cv);
    if (pull0 || pull1) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            a.ctl->pull_levels += (pull0 ? 1u : 0u) + (pull1 ? 1u : 0u);
        int64_t beg, end;
        my_range(a.n, beg, end);
        for (int64_t base = beg; base < end; base += blockDim.x) {
            int64_t vv = base + threadIdx.x;
            if (vv < end) {
                int32_t v = (int32_t)vv;
                uint32_t w = a.word[v];
                if (!(w & ST_E)) {
                    uint32_t c = colour_of(w);
                    for (int d = 0; d < 2; d++) {
                        if (!(d == 0 ? pull0 : pull1)) continue;
                        uint32_t bit = dbit(d);
                        if (w & bit) continue;
                        // reverse direction: a forward visit needs a reached in-neighbour
                        const int32_t *row = drow(a, 1 - d), *col = dcol(a, 1 - d);
                        int32_t e0 = row[v], e1 = row[v + 1];
                        cx++;
                        for (int32_t e = e0; e < e1; e++) {
                            ce++;
                            uint32_t wu = a.word[col[e]];
                            if ((wu & (bit | ST_E)) == bit && colour_of(wu) == c) {
                                uint32_t old = atomicOr(&a.word[v], bit);
                                if (!(old & bit)) {
                                    found = true;
                                    cv++;
                                    st_push(st, d, v, a);
                                }
                                break;
                            }
                        }
                    }
                }
            }
            st_maybe_flush(st, a);
        }
    }
    local_rounds(a, (pull0 || pull1) ? 0xFFFFFFFFu : qs0 + qs1, st, lq, found, cx, ce, cv);
    st_flush(st, a);
    chain_progress(a.ctl, a.iter, found);
    if (a.instr) instr_flush(a.ctl, cx, ce, cv, sh);
}

// ---------------------------------------------------------------------------
// data-driven trim within colours (trim_kernel, scc_routines.cu:336-394, run to
// its fixpoint by trim(), :616-643): the fixpoint is unique, so instead of
// re-sweeping every live vertex per iteration, level 0 sweeps once and every
// later level re-checks only the live same-colour neighbours of the vertices
// trimmed in the level before (each at most once per level, by stamp).
// ---------------------------------------------------------------------------
__device__ inline bool live_nb(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                               const uint32_t *word, int32_t v, uint32_t c) {
    for (int32_t e = row[v]; e < row[v + 1]; e++) {
        int32_t u = col[e];
        if (u == v) continue;
        uint32_t wu = word[u];
        if (!(wu & ST_E) && colour_of(wu) == c) return true;
    }
    return false;
}

// trims v (colour c) when it has no live same-colour in- or out-neighbour
__device__ inline bool trim_check(const ReachArgs &a, int32_t v, uint32_t c) {
    if (live_nb(a.row, a.col, a.word, v, c) && live_nb(a.rowT, a.colT, a.word, v, c)) return false;
    a.word[v] = make_word(c, 28);  // P|T|E (scc_routines.cu:381)
    a.root[v] = v;
    return true;
}

__global__ void k_trim_level(ReachArgs a) {
    __shared__ Items it;
    __shared__ Stage st;
    if (!level_begin(a, true)) return;
    st_init(st);
    __syncthreads();
    bool found = false;
    if (a.iter == 0) {
        int64_t beg, end;
        my_range(a.n, beg, end);
        for (int64_t base = beg; base < end; base += blockDim.x) {
            int64_t vv = base + threadIdx.x;
            if (vv < end) {
                int32_t v = (int32_t)vv;
                uint32_t w = a.word[v];
                if (!(w & ST_E) && trim_check(a, v, colour_of(w))) {
                    found = true;
                    st_push(st, 0, v, a);
                }
            }
            st_maybe_flush(st, a);
        }
    } else {
        uint32_t qs = a.ctl->qsize[0][a.iter & 3];
        uint32_t chunk = (qs + gridDim.x - 1) / gridDim.x;
        uint32_t b0 = blockIdx.x * chunk, b1 = b0 + chunk < qs ? b0 + chunk : qs;
        for (uint32_t base = b0; base < b1; base += blockDim.x) {
            // each trimmed vertex contributes its out-row and its in-row
            uint32_t i = base + threadIdx.x;
            uint32_t deg[2] = {0, 0}, c = 0;
            int32_t start[2] = {0, 0};
            if (i < b1) {
                int32_t x = a.inq[0][i];
                c = colour_of(a.word[x]);
                start[0] = a.row[x];
                deg[0] = (uint32_t)(a.row[x + 1] - start[0]);
                start[1] = a.rowT[x];
                deg[1] = (uint32_t)(a.rowT[x + 1] - start[1]);
            }
            uint32_t total;
            uint32_t off = block_excl_scan(deg[0] + deg[1], it.shw, total);
            it.off[2 * threadIdx.x] = off;
            it.off[2 * threadIdx.x + 1] = off + deg[0];
            it.start[2 * threadIdx.x] = start[0];
            it.start[2 * threadIdx.x + 1] = start[1];
            it.cd[2 * threadIdx.x] = c;
            it.cd[2 * threadIdx.x + 1] = c | 0x80000000u;
            __syncthreads();
            for (uint32_t pb = 0; pb < total; pb += blockDim.x) {  // block-uniform
                uint32_t e = pb + threadIdx.x;
                if (e < total) {
                    uint32_t k = find_item(it.off, 2 * blockDim.x, e);
                    uint32_t cd = it.cd[k];
                    const int32_t *col = (cd >> 31) ? a.colT : a.col;
                    int32_t u = col[it.start[k] + (int32_t)(e - it.off[k])];
                    uint32_t cc = cd & 0x7FFFFFFFu;
                    uint32_t wu = a.word[u];
                    if (!(wu & ST_E) && colour_of(wu) == cc &&
                        atomicExch(&a.stamp[u], a.tag) != a.tag && trim_check(a, u, cc)) {
                        found = true;
                        st_push(st, 0, u, a);
                    }
                }
                st_maybe_flush(st, a);
            }
        }
    }
    st_flush(st, a);
    chain_progress(a.ctl, a.iter, found);
}

__global__ void k_clear_bits(uint32_t *word, int32_t n, uint32_t bits) {
    int64_t beg, end;
    my_range(n, beg, end);
    for (int64_t v = beg + threadIdx.x; v < end; v += blockDim.x) {
        uint32_t w = word[v];
        if (w & bits) word[v] = w & ~bits;
    }
}

// ---------------------------------------------------------------------------
// launchers: with events set, the launch records them around the kernel
// itself (hipExtLaunchKernelGGL), so their interval is the kernel's duration
// ---------------------------------------------------------------------------
template <typename K>
static void launch_reach_kernel(K kernel, const Launch &l, const ReachArgs &a) {
    if (l.ev_start)
        hipExtLaunchKernelGGL(kernel, dim3(l.grid), dim3(l.block), 0, l.stream, l.ev_start, l.ev_stop,
                              0, a);
    else
        hipLaunchKernelGGL(kernel, dim3(l.grid), dim3(l.block), 0, l.stream, a);
}

void launch_reach(const Launch &l, int sched, const ReachArgs &a) {
    switch (sched) {
    case 0: launch_reach_kernel(k_reach_sweep, l, a); break;
    case 1: launch_reach_kernel(k_reach_topo, l, a); break;
    case 2: launch_reach_kernel(k_reach_queue, l, a); break;
    case 3: launch_reach_kernel(k_reach_lb, l, a); break;
    default: launch_reach_kernel(k_reach_hybrid, l, a); break;
    }
}

void launch_trim_level(const Launch &l, const ReachArgs &a) { launch_reach_kernel(k_trim_level, l, a); }

void launch_clear_bits(const Launch &l, uint32_t *word, int32_t n, uint32_t bits) {
    hipLaunchKernelGGL(k_clear_bits, dim3(l.grid), dim3(l.block), 0, l.stream, word, n, bits);
}

}  // namespace gpuscc
