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

constexpr uint32_t SQ_CAP = 2048;  // LDS staging entries per direction per workgroup
constexpr int EXP_ILP = 4;          // arcs per work-item per pass of a balanced expansion

__device__ inline uint32_t dbit(int d) { return d == 0 ? ST_F : ST_B; }
__device__ inline uint32_t dxbit(int d) { return d == 0 ? ST_FX : ST_BX; }
// Per-direction arrays of the arguments: both members are read first and the
// value is selected.  Written as `d == 0 ? a.x[0] : a.x[1]` (or a.x[d]), the
// compiler merges the two reads into one read at a per-lane address -- a
// vector load of the pointer, one more dependent round trip before every use;
// reading both keeps them scalar loads hoisted to the kernel's start.
template <typename T>
__device__ inline T dpick(int d, T x0, T x1) {
    return d == 0 ? x0 : x1;
}
__device__ inline const int32_t *drow(const ReachArgs &a, int d) { return dpick(d, a.row, a.rowT); }
__device__ inline const int32_t *dcol(const ReachArgs &a, int d) { return dpick(d, a.col, a.colT); }
__device__ inline const uint8_t *ddcls(const ReachArgs &a, int d) { return dpick(d, a.dcls[0], a.dcls[1]); }
__device__ inline const int4 *dell(const ReachArgs &a, int d) { return dpick(d, a.ell[0], a.ell[1]); }
__device__ inline int32_t *doutq(const ReachArgs &a, int d) { return dpick(d, a.outq[0], a.outq[1]); }
__device__ inline int32_t *dhout(const ReachArgs &a, int d) { return dpick(d, a.hout[0], a.hout[1]); }
__device__ inline const int32_t *dhin(const ReachArgs &a, int d) { return dpick(d, a.hin[0], a.hin[1]); }

__device__ inline uint32_t *dvis(const ReachArgs &a, int d) { return dpick(d, a.vis[0], a.vis[1]); }

// u is being claimed in direction d from a reached vertex: its bit in the
// visited bitmap (first FB step of the hybrid schedule only), set with the claim
// itself (no wait for the claim's result).  A claim without a look at the
// target's word may mark an eliminated vertex; in the first step those are
// trim-1's, and a trimmed vertex with a live in-neighbour has no live
// out-neighbour (else trim-1 would have kept it), so no pull ever meets such a
// mark: the marks are the reached live vertices as far as the pulls can tell.
__device__ inline void mark_vis(const ReachArgs &a, int d, int32_t u) {
    uint32_t *bm = dvis(a, d);
    if (bm) atomicOr(&bm[u >> 5], 1u << (u & 31));
}

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
    unsigned long long arcs[2];  // estimated arcs of vertices queued past the buffers since the last flush
};

__device__ inline void st_init(Stage &s) {
    if (threadIdx.x < 2) {
        s.cnt[threadIdx.x] = 0;
        s.arcs[threadIdx.x] = 0;
    }
}

// arcs of a vertex of degree class k (bit length k, bit 7 the self-loop flag): about 3/4 of 2^k
__device__ inline unsigned long long class_arcs(uint32_t k) { return (3ull << (k & 0x3Fu)) >> 2; }

// A push past the buffers goes straight to the queue (and its arc estimate
// to the control block).
__device__ inline void st_overflow(int d, int32_t v, const ReachArgs &a) {
    uint32_t g = atomicAdd(&a.ctl->qsize[d][(a.iter + 1) & 3], 1u);
    doutq(a, d)[g] = v;
    if (ddcls(a, d)) atomicAdd(&a.ctl->qarcs[d][(a.iter + 1) & 3], class_arcs(ddcls(a, d)[v]));
}

// One lane's push: one LDS atomic.  The arc estimate of the next frontier
// (degree classes) is taken when the staged vertices reach the queue
// (st_flush), not per push.  Used where few lanes push at a time (the
// latency-bound local rounds).
__device__ inline void st_push(Stage &s, int d, int32_t v, const ReachArgs &a) {
    uint32_t i = atomicAdd(&s.cnt[d], 1u);
    if (i < SQ_CAP) s.buf[d][i] = v;
    else st_overflow(d, v, a);
}

// the same with the two counts at cnt (the local rounds rotate their counts)
__device__ inline void st_push_at(Stage &s, uint32_t *cnt, int d, int32_t v, const ReachArgs &a) {
    uint32_t i = atomicAdd(&cnt[d], 1u);
    if (i < SQ_CAP) s.buf[d][i] = v;
    else st_overflow(d, v, a);
}

// Wave-aggregated push, for passes where many lanes push together: the active
// lanes pushing direction d ballot, the lowest of them reserves their slots
// with one LDS atomic, and each lane writes at that base plus the number of
// pushing lanes below it.
__device__ inline void st_push_wave(Stage &s, int d, int32_t v, const ReachArgs &a) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int dd = 0; dd < 2; dd++) {
        const unsigned long long m = __ballot(d == dd);  // among the lanes calling now
        if (d != dd) continue;
        const int leader = __ffsll((long long)m) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&s.cnt[dd], (uint32_t)__popcll(m));
        base = __shfl(base, leader, WAVE);
        const uint32_t i = base + (uint32_t)__popcll(m & below);
        if (i < SQ_CAP) s.buf[dd][i] = v;
        else st_overflow(dd, v, a);
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
    unsigned long long ar0 = 0, ar1 = 0;
    for (uint32_t i = threadIdx.x; i < c0; i += a.nt) {
        int32_t v = s.buf[0][i];
        a.outq[0][s.base[0] + i] = v;
        if (a.dcls[0]) ar0 += class_arcs(a.dcls[0][v]);
    }
    for (uint32_t i = threadIdx.x; i < c1; i += a.nt) {
        int32_t v = s.buf[1][i];
        a.outq[1][s.base[1] + i] = v;
        if (a.dcls[1]) ar1 += class_arcs(a.dcls[1][v]);
    }
    if (ar0) atomicAdd(&s.arcs[0], ar0);
    if (ar1) atomicAdd(&s.arcs[1], ar1);
    __syncthreads();
    if (threadIdx.x < 2) {
        if (s.arcs[threadIdx.x]) atomicAdd(&a.ctl->qarcs[threadIdx.x][(a.iter + 1) & 3], s.arcs[threadIdx.x]);
        s.cnt[threadIdx.x] = 0;
        s.arcs[threadIdx.x] = 0;
    }
    __syncthreads();
}

// all work-items call this; flushes when another pass (at most `per_item`
// pushes per work-item) could overflow a buffer
__device__ inline void st_maybe_flush(Stage &s, const ReachArgs &a, uint32_t per_item = 1) {
    __syncthreads();
    const uint32_t room = per_item * a.nt;
    bool full = s.cnt[0] > SQ_CAP - room || s.cnt[1] > SQ_CAP - room;
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
constexpr uint32_t ITEM_CAP = 1024;  // two items per work-item of a workgroup of up to 512

struct Items {
    uint32_t off[ITEM_CAP];
    int32_t start[ITEM_CAP];
    uint32_t cd[ITEM_CAP];  // colour | direction << 31
    int32_t lim[ITEM_CAP];  // end of the row (huge rows)
    uint32_t shw[16];
};

template <bool PUSH>
__device__ inline void expand_items(const ReachArgs &a, Items &it, uint32_t nitems, uint32_t total,
                                    Stage *st, bool &found, unsigned long long &cv, bool keep_last = false) {
    // EXP_ILP arcs per work-item and pass: their column ids, then their words,
    // then the claims of the candidates, each group of loads in flight together
    for (uint32_t base = 0; base < total; base += a.nt * EXP_ILP) {  // block-uniform trip count
        int32_t u[EXP_ILP];
        uint32_t cd[EXP_ILP];
#pragma unroll
        for (int j = 0; j < EXP_ILP; j++) {
            uint32_t e = base + (uint32_t)j * a.nt + threadIdx.x;
            u[j] = -1;
            cd[j] = 0;
            if (e < total) {
                uint32_t k = find_item(it.off, nitems, e);
                cd[j] = it.cd[k];
                u[j] = dcol(a, (int)(cd[j] >> 31))[it.start[k] + (int32_t)(e - it.off[k])];
            }
        }
        uint32_t wu[EXP_ILP], old[EXP_ILP];
#pragma unroll
        for (int j = 0; j < EXP_ILP; j++) wu[j] = u[j] >= 0 ? a.word[u[j]] : ST_E;
#pragma unroll
        for (int j = 0; j < EXP_ILP; j++) {
            uint32_t bit = dbit((int)(cd[j] >> 31));
            old[j] = bit;
            if (!(wu[j] & (bit | ST_E)) && colour_of(wu[j]) == (cd[j] & 0x7FFFFFFFu)) {
                old[j] = atomicOr(&a.word[u[j]], bit);
                mark_vis(a, (int)(cd[j] >> 31), u[j]);  // with the claim: u is reached by it or another
            }
        }
#pragma unroll
        for (int j = 0; j < EXP_ILP; j++) {
            if (old[j] & dbit((int)(cd[j] >> 31))) continue;
            found = true;
            cv++;
            if (PUSH) st_push_wave(*st, (int)(cd[j] >> 31), u[j], a);
        }
        // room for the next pass (with 512 work-items one pass can fill the whole
        // buffer); keep_last leaves the last pass's finds staged for the local rounds
        if (PUSH && !(keep_last && base + a.nt * EXP_ILP >= total)) st_maybe_flush(*st, a, EXP_ILP);
    }
}

// Gate of a reach level: exits when the previous level made no progress;
// block 0 clears the output counters of the level after next.
__device__ inline bool level_begin(const ReachArgs &a, bool queues) {
    if (chain_exit(a.ctl, a.iter)) return false;
    if (blockIdx.x == 0 && threadIdx.x < 2) {
        a.ctl->hsize[threadIdx.x][(a.iter + 2) & 3] = 0;
        if (queues) {
            a.ctl->qsize[threadIdx.x][(a.iter + 2) & 3] = 0;
            a.ctl->qarcs[threadIdx.x][(a.iter + 2) & 3] = 0;
        }
        if (threadIdx.x == 0) a.ctl->st_maxscan[(a.iter + 2) & 3] = 0;
    }
    return true;
}

// instrumented: the level's longest serial row scan (wave maximum, one atomic per wave)
__device__ inline void instr_maxscan(const ReachArgs &a, uint32_t steps) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t x = (uint32_t)__shfl_xor((int)steps, o, 64);
        steps = x > steps ? x : steps;
    }
    if ((threadIdx.x & 63) == 0 && steps) atomicMax(&a.ctl->st_maxscan[a.iter & 3], steps);
}

// a row of at least huge_thr arcs is not expanded by its own lane or
// workgroup: it is deferred to the next level, where all workgroups split it
__device__ inline void push_huge(const ReachArgs &a, int d, int32_t v) {
    uint32_t g = atomicAdd(&a.ctl->hsize[d][(a.iter + 1) & 3], 1u);
    dhout(a, d)[g] = v;
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
        const uint32_t step = a.nt < ITEM_CAP ? a.nt : ITEM_CAP;  // rows per pass
        for (uint32_t base = 0; base < hs; base += step) {
            uint32_t i = base + threadIdx.x;
            uint32_t nch = 0, c = 0;
            int32_t e0 = 0, e1 = 0;
            if (threadIdx.x < step && i < hs) {
                int32_t v = dhin(a, d)[i];
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
            uint32_t off = block_excl_scan(nch, it.shw, total, a.nt);
            if (threadIdx.x < step) {
                it.off[threadIdx.x] = off;
                it.start[threadIdx.x] = e0;
                it.lim[threadIdx.x] = e1;
                it.cd[threadIdx.x] = c;
            }
            __syncthreads();
            for (uint32_t k = acc + (blockIdx.x + G - acc % G) % G; k < acc + total; k += G) {
                uint32_t local = k - acc;
                uint32_t j = find_item(it.off, step, local);
                int32_t cs = it.start[j] + (int32_t)((local - it.off[j]) * HCHUNK);
                int32_t cend = cs + (int32_t)HCHUNK < it.lim[j] ? cs + (int32_t)HCHUNK : it.lim[j];
                uint32_t cc = it.cd[j];
                for (int32_t pb = cs; pb < cend; pb += (int32_t)a.nt) {  // block-uniform
                    int32_t e = pb + (int32_t)threadIdx.x;
                    if (e < cend && try_claim(a.word, col[e], cc, bit)) {
                        found = true;
                        cv++;
                        mark_vis(a, d, col[e]);
                        if (PUSH) st_push_wave(*st, d, col[e], a);
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
    for (int64_t vv = beg + threadIdx.x; vv < end; vv += a.nt) {
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
    for (int64_t base = beg; base < end; base += a.nt) {
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
        uint32_t off = block_excl_scan(deg[0] + deg[1], it.shw, total, a.nt);
        it.off[2 * threadIdx.x] = off;
        it.off[2 * threadIdx.x + 1] = off + deg[0];
        it.start[2 * threadIdx.x] = start[0];
        it.start[2 * threadIdx.x + 1] = start[1];
        it.cd[2 * threadIdx.x] = c;
        it.cd[2 * threadIdx.x + 1] = c | 0x80000000u;
        __syncthreads();
        if (total > 0) expand_items<false>(a, it, 2 * a.nt, total, nullptr, found, cv);
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
constexpr uint32_t LOCAL_CAP = 512;   // items of one local round (the widest workgroup: one per work-item)
constexpr int32_t HEAD_UNKNOWN = -2;  // hd[.].x of an item whose row head is not at hand (ELL entries are >= -1)

struct LocalQ {
    int32_t v[2][LOCAL_CAP];  // the items of a round and of the next: vertex | direction << 31
    int4 hd[2][LOCAL_CAP];    // their row heads, loaded with the claim that found them (complete rows only)
    uint32_t rc[3];           // item counts, rotating (see local_rounds)
};

// The local rounds start after the last use of the items of a level, so the
// two share LDS (fewer bytes per workgroup, more workgroups per CU).
union ItemsLocal {
    Items it;
    LocalQ lq;
};

// expands item v (colour c) of direction d serially over its CSR row; new
// visits are staged (wave-aggregated, or one by one with the counts at cnt)
template <bool WAVE_PUSH>
__device__ inline void expand_serial(const ReachArgs &a, Stage &st, uint32_t *cnt, int d, int32_t v,
                                     uint32_t c, bool &found, unsigned long long &cx,
                                     unsigned long long &ce, unsigned long long &cv) {
    const uint32_t bit = dbit(d);
    const int32_t *row = drow(a, d), *col = dcol(a, d);
    int32_t e0 = row[v], e1 = row[v + 1];
    if ((uint32_t)(e1 - e0) >= a.huge_thr) {
        push_huge(a, d, v);
        found = true;
        return;
    }
    cx++;
    ce += (unsigned long long)(e1 - e0);
    // four arcs per round: their column ids, then their words, then the claims
    // of the candidates, each group of loads in flight together
    constexpr int A = 4;
    for (int32_t e = e0; e < e1; e += A) {
        int32_t u[A];
        uint32_t wu[A];
#pragma unroll
        for (int j = 0; j < A; j++) u[j] = e + j < e1 ? col[e + j] : -1;
#pragma unroll
        for (int j = 0; j < A; j++) wu[j] = u[j] >= 0 ? a.word[u[j]] : ST_E;
        uint32_t old[A];
#pragma unroll
        for (int j = 0; j < A; j++) {
            old[j] = bit;
            if (!(wu[j] & (bit | ST_E)) && colour_of(wu[j]) == c) {
                old[j] = atomicOr(&a.word[u[j]], bit);
                mark_vis(a, d, u[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < A; j++) {
            if (old[j] & bit) continue;
            found = true;
            cv++;
            if (WAVE_PUSH) st_push_wave(st, d, u[j], a);
            else st_push_at(st, cnt, d, u[j], a);
        }
    }
}

// Expands item v of direction d whose row fits its four-entry head (class byte
// k has CLS_FITS): the column ids come with v's own word in the first round
// trip, so a short row costs three dependent accesses (head, target words,
// claims) instead of five (row bounds, own word, column ids, words, claims).
//
// Without PRECHECK (only when every arc joins two vertices of one colour) the
// claims go out without reading the target words first: a target that is
// already reached keeps its word, and one that is eliminated (E) gets a bit no
// later kernel reads (every one skips E) and is not pushed, since the claim's
// old value shows E -- two dependent accesses instead of three.
template <bool WAVE_PUSH, bool PRECHECK = true>
__device__ inline void expand_fit(const ReachArgs &a, Stage &st, uint32_t *cnt, int d, int32_t v,
                                  uint32_t c, int4 nb, bool &found, unsigned long long &cx,
                                  unsigned long long &ce, unsigned long long &cv) {
    const uint32_t bit = dbit(d);
    const int32_t u[4] = {nb.x, nb.y, nb.z, nb.w};
    const uint32_t deg = (nb.x >= 0 ? 1u : 0u) + (nb.y >= 0 ? 1u : 0u) + (nb.z >= 0 ? 1u : 0u) +
                         (nb.w >= 0 ? 1u : 0u);
    if (deg >= a.huge_thr) {
        push_huge(a, d, v);
        found = true;
        return;
    }
    cx++;
    ce += deg;
    uint32_t old[4];
    if (PRECHECK) {
        uint32_t wu[4];
#pragma unroll
        for (int j = 0; j < 4; j++) wu[j] = u[j] >= 0 ? a.word[u[j]] : ST_E;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            old[j] = bit;
            if (!(wu[j] & (bit | ST_E)) && colour_of(wu[j]) == c) {
                old[j] = atomicOr(&a.word[u[j]], bit);
                mark_vis(a, d, u[j]);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            old[j] = u[j] >= 0 ? atomicOr(&a.word[u[j]], bit) : bit;
            if (u[j] >= 0) mark_vis(a, d, u[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (old[j] & (bit | ST_E)) continue;
        found = true;
        cv++;
        if (WAVE_PUSH) st_push_wave(st, d, u[j], a);
        else st_push_at(st, cnt, d, u[j], a);
    }
}

constexpr int32_t LOCAL_MAX_DEG = 64;       // longer rows leave the local rounds for the next level
constexpr uint32_t LOCAL_ITEMS_PER_WG = 8;  // local rounds only when a level's frontier is this thin

// a vertex found in a local round: into the next round's items, with its row
// head when the row fits it, else (past the round's capacity) to the level's staging
__device__ inline void lq_push(const ReachArgs &a, Stage &st, LocalQ &lq, uint32_t *out, int pn, int d, int32_t u,
                               int4 h) {
    const uint32_t i = atomicAdd(out, 1u);
    if (i < LOCAL_CAP) {
        lq.v[pn][i] = (int32_t)((uint32_t)u | ((uint32_t)d << 31));
        lq.hd[pn][i] = h;
    } else {
        st_push(st, d, u, a);
    }
}

// Expands a local item whose row fits its head nb.  The claims go out with the
// loads of the targets' own heads and class bytes, so the next round starts
// from them without another round trip: one dependent access per round when
// every arc stays inside a colour (step one), two otherwise (target words, then
// claims -- see expand_fit).
__device__ inline void expand_fit_local(const ReachArgs &a, Stage &st, LocalQ &lq, uint32_t *out, int pn, int d,
                                        int32_t v, uint32_t c, int4 nb, bool &found, unsigned long long &cx,
                                        unsigned long long &ce, unsigned long long &cv) {
    const uint32_t bit = dbit(d);
    const int32_t u[4] = {nb.x, nb.y, nb.z, nb.w};
    const uint32_t deg = (nb.x >= 0 ? 1u : 0u) + (nb.y >= 0 ? 1u : 0u) + (nb.z >= 0 ? 1u : 0u) +
                         (nb.w >= 0 ? 1u : 0u);
    if (deg >= a.huge_thr) {
        push_huge(a, d, v);
        found = true;
        return;
    }
    cx++;
    ce += deg;
    const int4 *ell = dell(a, d);
    const uint8_t *cls = ddcls(a, d);
    // every access is issued for all four entries, an absent one at v itself
    // (and its claim with an empty mask), with no branch between them: a branch
    // would make the compiler wait for each access before the next
    int32_t s[4];
#pragma unroll
    for (int j = 0; j < 4; j++) s[j] = u[j] >= 0 ? u[j] : v;
    int4 hn[4];
    uint32_t cn[4], old[4];
    if (a.one_colour) {  // the claims (and their marks) first, then the heads: one round trip
#pragma unroll
        for (int j = 0; j < 4; j++) old[j] = atomicOr(&a.word[s[j]], u[j] >= 0 ? bit : 0u);
        uint32_t *bm = dvis(a, d);
        if (bm) {
#pragma unroll
            for (int j = 0; j < 4; j++) atomicOr(&bm[s[j] >> 5], u[j] >= 0 ? 1u << (s[j] & 31) : 0u);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) hn[j] = ell[s[j]];
#pragma unroll
        for (int j = 0; j < 4; j++) cn[j] = cls[s[j]];
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (u[j] < 0) old[j] = bit;
    } else {  // heads and target words, then the claims: two
        uint32_t wu[4];
#pragma unroll
        for (int j = 0; j < 4; j++) hn[j] = ell[s[j]];
#pragma unroll
        for (int j = 0; j < 4; j++) cn[j] = cls[s[j]];
#pragma unroll
        for (int j = 0; j < 4; j++) wu[j] = a.word[s[j]];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const bool claim = u[j] >= 0 && !(wu[j] & (bit | ST_E)) && colour_of(wu[j]) == c;
            old[j] = atomicOr(&a.word[s[j]], claim ? bit : 0u);
            if (claim) mark_vis(a, d, u[j]);
            if (!claim) old[j] = bit;
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (old[j] & (bit | ST_E)) continue;
        found = true;
        cv++;
        int4 h = hn[j];
        if (!(cn[j] & CLS_FITS)) h.x = HEAD_UNKNOWN;
        lq_push(a, st, lq, out, pn, d, u[j], h);
    }
}

// Serial expansion of a local item whose row does not fit its head (at most
// LOCAL_MAX_DEG arcs); its finds go to the next round without heads.
__device__ inline void expand_serial_local(const ReachArgs &a, Stage &st, LocalQ &lq, uint32_t *out, int pn, int d,
                                           int32_t v, uint32_t c, bool &found, unsigned long long &cx,
                                           unsigned long long &ce, unsigned long long &cv) {
    const uint32_t bit = dbit(d);
    const int32_t *row = drow(a, d), *col = dcol(a, d);
    const int32_t e0 = row[v], e1 = row[v + 1];
    cx++;
    ce += (unsigned long long)(e1 - e0);
    constexpr int A = 4;
    for (int32_t e = e0; e < e1; e += A) {
        int32_t u[A];
        uint32_t wu[A], old[A];
#pragma unroll
        for (int j = 0; j < A; j++) u[j] = e + j < e1 ? col[e + j] : -1;
#pragma unroll
        for (int j = 0; j < A; j++) wu[j] = u[j] >= 0 ? a.word[u[j]] : ST_E;
#pragma unroll
        for (int j = 0; j < A; j++) {
            old[j] = bit;
            if (!(wu[j] & (bit | ST_E)) && colour_of(wu[j]) == c) {
                old[j] = atomicOr(&a.word[u[j]], bit);
                mark_vis(a, d, u[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < A; j++) {
            if (old[j] & bit) continue;
            found = true;
            cv++;
            lq_push(a, st, lq, out, pn, d, u[j], make_int4(HEAD_UNKNOWN, 0, 0, 0));
        }
    }
}

// Workgroup-local rounds on the visits staged by this workgroup: round 0
// takes them from the staging buffers, every later round the vertices the
// round before found (with their row heads, loaded alongside the claims that
// found them); rows longer than LOCAL_MAX_DEG go to the next level's queue
// instead.  One barrier per round: the items alternate between two buffers
// and the counts rotate over three slots -- round r reads slot r, pushes into
// slot r + 1 and clears slot r + 2, which round r - 1 read before round r's
// barrier.  A find past the round's capacity is staged for the next level, as
// are the items left when the rounds stop.  All work-items call this.
__device__ inline void local_rounds(const ReachArgs &a, uint32_t level_items, Stage &st, LocalQ &lq,
                                    bool &found, unsigned long long &cx, unsigned long long &ce,
                                    unsigned long long &cv) {
    // a thick frontier keeps every workgroup busy without them; a thin one
    // (high-diameter graphs) would leave most workgroups idle per level.
    if (a.iter < a.local_first || level_items > LOCAL_ITEMS_PER_WG * gridDim.x) return;
    __syncthreads();  // the level's pushes are staged; the items sharing lq's memory are done with
    const uint32_t s0 = st.cnt[0], s1 = st.cnt[1];
    // block-uniform: nothing to expand, or the workgroup's own frontier is wider
    // than one item per work-item -- a wide frontier is the next level's work
    // (balanced over all workgroups)
    if (s0 + s1 == 0 || s0 + s1 > a.nt) return;
    for (uint32_t k = threadIdx.x; k < s0 + s1; k += a.nt) {
        const int d = k < s0 ? 0 : 1;
        lq.v[0][k] = (int32_t)((uint32_t)st.buf[d][d == 0 ? k : k - s0] | ((uint32_t)d << 31));
        lq.hd[0][k].x = HEAD_UNKNOWN;
    }
    __syncthreads();  // st.cnt and st.buf are read: the rounds may stage again
    if (threadIdx.x == 0) {
        st.cnt[0] = st.cnt[1] = 0;
        lq.rc[0] = s0 + s1;
        lq.rc[1] = lq.rc[2] = 0;
    }
    int ci = 0, p = 0;  // slot of this round's count, buffer of its items
    for (int r = 0; r < a.local_rounds; r++) {
        __syncthreads();  // the finds of the round before are complete
        const uint32_t cnt = lq.rc[ci];
        if (cnt == 0 || cnt > a.nt) break;  // block-uniform (read after the barrier)
        const int co = ci == 2 ? 0 : ci + 1, cz = co == 2 ? 0 : co + 1;
        if (threadIdx.x == 0) lq.rc[cz] = 0;
        uint32_t *out = &lq.rc[co];
        const uint32_t k = threadIdx.x;
        if (k < cnt) {
            const int32_t x = lq.v[p][k];
            const int d = x < 0 ? 1 : 0;
            const int32_t v = x & 0x7FFFFFFF;
            int4 nb = lq.hd[p][k];
            uint32_t c = 0;
            bool fits = true;
            if (nb.x == HEAD_UNKNOWN) {
                // class byte, own word and row head in one round trip (the empty
                // asm needs the head before the branch, so its load is not sunk
                // into the branch behind the other two)
                const uint32_t cls = ddcls(a, d)[v];
                c = colour_of(a.word[v]);
                nb = dell(a, d)[v];
                asm volatile("" ::"v"(nb.x));
                fits = (cls & CLS_FITS) != 0;
            } else if (!a.one_colour) {
                c = colour_of(a.word[v]);
            }
            if (fits) {
                expand_fit_local(a, st, lq, out, p ^ 1, d, v, c, nb, found, cx, ce, cv);
            } else {
                const int32_t *row = drow(a, d);
                if (row[v + 1] - row[v] > LOCAL_MAX_DEG) {
                    atomicAdd(&st.arcs[d], class_arcs(ddcls(a, d)[v]));
                    uint32_t g = atomicAdd(&a.ctl->qsize[d][(a.iter + 1) & 3], 1u);
                    doutq(a, d)[g] = v;
                } else {
                    expand_serial_local(a, st, lq, out, p ^ 1, d, v, c, found, cx, ce, cv);
                }
            }
        }
        ci = co;
        p ^= 1;
    }
    // the items of the last round read (or the finds of the last round run)
    // are claimed but not expanded: the next level's work
    __syncthreads();
    const uint32_t left = lq.rc[ci] < LOCAL_CAP ? lq.rc[ci] : LOCAL_CAP;
    for (uint32_t k = threadIdx.x; k < left; k += a.nt) {
        const int32_t x = lq.v[p][k];
        st_push(st, x < 0 ? 1 : 0, x & 0x7FFFFFFF, a);
    }
}

__global__ void k_reach_queue(ReachArgs a) {
    __shared__ ItemsLocal ul;
    __shared__ Stage st;
    __shared__ unsigned long long sh[16];
    Items &it = ul.it;
    LocalQ &lq = ul.lq;
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
    for (uint32_t base = b0; base < b1; base += a.nt) {
        uint32_t i = base + threadIdx.x;
        if (i < b1) {
            int d = i < qs0 ? 0 : 1;
            int32_t v = d == 0 ? a.inq[0][i] : a.inq[1][i - qs0];
            const uint32_t cls = ddcls(a, d)[v];
            const uint32_t c = colour_of(a.word[v]);
            const int4 nb = dell(a, d)[v];
            asm volatile("" ::"v"(nb.x));
            if (cls & CLS_FITS) expand_fit<true>(a, st, st.cnt, d, v, c, nb, found, cx, ce, cv);
            else expand_serial<true>(a, st, st.cnt, d, v, c, found, cx, ce, cv);
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
    // a level thin enough for the local rounds keeps its finds staged for them
    const bool keep = a.local_rounds > 0 && a.iter >= a.local_first && tot <= LOCAL_ITEMS_PER_WG * gridDim.x;
    for (uint32_t base = b0; base < b1; base += a.nt) {
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
        uint32_t off = block_excl_scan(deg, it.shw, total, a.nt);
        it.off[threadIdx.x] = off;
        it.start[threadIdx.x] = start;
        it.cd[threadIdx.x] = cd;
        __syncthreads();
        if (total > 0) expand_items<true>(a, it, a.nt, total, &st, found, cv, keep);
        st_maybe_flush(st, a);
    }
}

// ---------------------------------------------------------------------------
// l: frontier queues with balanced expansion (reach_l.cu:151-253)
// ---------------------------------------------------------------------------
__global__ void k_reach_lb(ReachArgs a) {
    __shared__ ItemsLocal ul;
    __shared__ Stage st;
    __shared__ unsigned long long sh[16];
    Items &it = ul.it;
    LocalQ &lq = ul.lq;
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
// h: a direction whose frontier holds fewer than pull_thr vertices and fewer
// than pull_arcs arcs (estimated from degree classes: few vertices may hold
// hubs with most of the graph's arcs) expands its queue (balanced push); a
// larger one runs bottom-up: every live vertex of
// the direction that is not reached scans its reverse row for a reached
// same-colour neighbour and stops at the first.  Pulled visits are queued too,
// so the next level may push again.
// ---------------------------------------------------------------------------
// the directions level a.iter pulls (read once the previous level is complete)
__device__ inline void pull_dirs(const ReachArgs &a, uint32_t &qs0, uint32_t &qs1, bool &pull0,
                                 bool &pull1) {
    qs0 = (a.dirs & 1u) ? a.ctl->qsize[0][a.iter & 3] : 0u;
    qs1 = (a.dirs & 2u) ? a.ctl->qsize[1][a.iter & 3] : 0u;
    if (!a.pull_cost) {
        pull0 = qs0 > 0 && (qs0 >= a.pull_thr || a.ctl->qarcs[0][a.iter & 3] >= a.pull_arcs);
        pull1 = qs1 > 0 && (qs1 >= a.pull_thr || a.ctl->qarcs[1][a.iter & 3] >= a.pull_arcs);
        return;
    }
    if (a.pull_cost == 2) {
        // Direction-optimising switch (the BEST schedule's first step, reach_b.cu:540-596
        // and :645-700 with alpha = 15, beta = 18): top-down until the frontier's arcs
        // exceed the arcs not yet explored top-down over alpha; bottom-up while the
        // frontier grows or holds more than n / beta vertices.  Every input below was
        // written by an earlier level, so all workgroups take the same decision.
        bool p[2];
        for (int d = 0; d < 2; d++) {
            const uint32_t qs = d == 0 ? qs0 : qs1;
            const int pi = (a.iter - 1) & 3;
            const bool was_pull = a.iter > 0 && ((a.ctl->pmask[pi] >> d) & 1u);
            if (was_pull) {
                p[d] = qs > 0 && (qs >= a.ctl->qsize[d][pi] || qs > a.pull_thr);
            } else {
                const unsigned long long done = a.iter > 0 ? a.ctl->carcs[d][pi] : 0ull;
                const unsigned long long todo = a.pull_arcs > done ? a.pull_arcs - done : 0ull;
                p[d] = qs > 0 && a.ctl->qarcs[d][a.iter & 3] * a.alpha > todo;
            }
        }
        pull0 = p[0];
        pull1 = p[1];
        return;
    }
    // Cost model: a push level costs about the arcs of its frontier, a pull
    // level about the live vertices the direction has not reached (most of
    // them scan a short row; in a batch of many graphs most never find a
    // reached neighbour).  Reached so far = the queue inputs of this and all
    // earlier levels (every claim enters a queue exactly once).
    const uint32_t live = a.ctl->live;
    bool p[2];
    for (int d = 0; d < 2; d++) {
        const uint32_t qs = d == 0 ? qs0 : qs1;
        const uint32_t reached = (a.iter > 0 ? a.ctl->creach[d][(a.iter - 1) & 3] : 0u) + qs;
        const unsigned long long unreached = live > reached ? live - reached : 0u;
        p[d] = qs > 0 && a.ctl->qarcs[d][a.iter & 3] >= 2ull * unreached;
    }
    pull0 = p[0];
    pull1 = p[1];
}

// ---------------------------------------------------------------------------
// Row scans with early exit, several per work-item in lockstep.  Task t scans
// col[t][e[t] - 1 .. e1[t]) for a neighbour u != self[t] whose word w has
// (w & mask[t]) == val[t] and colour c[t]; bit t of the result says it found
// one.  The first entry is not gathered from the column array: the caller
// passes it as u0[t], read from the head array (one coalesced word per vertex)
// together with the row bounds, e[t] being the row start + 1.  One lane's
// scans are independent, so their gathers overlap: a lane waits for one round
// trip per arc position instead of one per arc.  Tasks to scan are set in act
// (their rows non-empty).
// ---------------------------------------------------------------------------
constexpr int SCAN_ILP = 4;  // vertices per work-item per pass of the sweeps below

template <int T>
__device__ inline uint32_t scan_rows(const uint32_t *word, const int32_t *const (&col)[T], int32_t (&e)[T],
                                     const int32_t (&e1)[T], const int32_t (&u0)[T], const uint32_t (&mask)[T],
                                     const uint32_t (&val)[T], const uint32_t (&c)[T],
                                     const int32_t (&self)[T], uint32_t act, unsigned long long &ce,
                                     uint32_t &steps) {
    uint32_t hit = 0;
    int s = 0;
    for (; act; s++) {
        uint32_t wu[T];
        int32_t u[T];
#pragma unroll
        for (int t = 0; t < T; t++) u[t] = ((act >> t) & 1u) ? (s == 0 ? u0[t] : col[t][e[t]]) : -1;
#pragma unroll
        for (int t = 0; t < T; t++) wu[t] = u[t] >= 0 ? word[u[t]] : 0u;
#pragma unroll
        for (int t = 0; t < T; t++) {
            if (!((act >> t) & 1u)) continue;
            ce++;
            if (u[t] != self[t] && (wu[t] & mask[t]) == val[t] && colour_of(wu[t]) == c[t]) {
                hit |= 1u << t;
                act &= ~(1u << t);
            } else if (s > 0 ? ++e[t] >= e1[t] : e[t] >= e1[t]) {
                act &= ~(1u << t);
            }
        }
    }
    steps = (uint32_t)s > steps ? (uint32_t)s : steps;
    return hit;
}

// The same scans against the visited bitmaps (bm[t]: the bitmap of row t's
// direction): a neighbour hits when its bit is set.  Valid in the first FB step,
// where every arc joins two vertices of one colour and an eliminated vertex is
// never marked, so the bit is exactly "reached in that direction and live".
template <int T>
__device__ inline uint32_t scan_rows_vis(const uint32_t *const (&bm)[T], const int32_t *const (&col)[T],
                                         int32_t (&e)[T], const int32_t (&e1)[T], const int32_t (&u0)[T],
                                         const int32_t (&self)[T], uint32_t act, unsigned long long &ce,
                                         uint32_t &steps) {
    uint32_t hit = 0;
    int s = 0;
    for (; act; s++) {
        uint32_t bw[T];
        int32_t u[T];
#pragma unroll
        for (int t = 0; t < T; t++) u[t] = ((act >> t) & 1u) ? (s == 0 ? u0[t] : col[t][e[t]]) : -1;
#pragma unroll
        for (int t = 0; t < T; t++) bw[t] = u[t] >= 0 ? bm[t][u[t] >> 5] : 0u;
#pragma unroll
        for (int t = 0; t < T; t++) {
            if (!((act >> t) & 1u)) continue;
            ce++;
            if (u[t] != self[t] && ((bw[t] >> (u[t] & 31)) & 1u)) {
                hit |= 1u << t;
                act &= ~(1u << t);
            } else if (s > 0 ? ++e[t] >= e1[t] : e[t] >= e1[t]) {
                act &= ~(1u << t);
            }
        }
    }
    steps = (uint32_t)s > steps ? (uint32_t)s : steps;
    return hit;
}

// vertices per work-item per round of a pull sweep: 8 rounds ahead of 16 and 32 (shorter waits
// for each round's claims to be queued; R-MAT s24 1.03 -> 1.01 ms, the batch of 4,096 graphs
// 8.31 -> 8.01 ms) and of 4 (the words of a round no longer cover the load latency)
constexpr int PULL_V = 8;

// Queues the vertices one sweep round claimed (pull visits, trims): bit k of
// claimed[d] is vertex base + k * blockDim.x + threadIdx.x.  One prefix sum
// and one atomicAdd per direction and workgroup; the only barriers of the
// round.  All work-items call this.
__device__ inline void commit_round(const ReachArgs &a, int64_t base, const uint32_t (&claimed)[2],
                                   const unsigned long long (&arcs)[2], uint32_t *shw, uint32_t *sbase,
                                   unsigned long long *sh, bool trims = false, uint32_t newp = 0,
                                   const int32_t *lst = nullptr) {
    const uint32_t c0 = __popc(claimed[0]), c1 = __popc(claimed[1]);
    uint32_t total;  // packed counts: a workgroup claims at most 16 * 1024 < 2^16 per direction
    const uint32_t pre = block_excl_scan(c0 | (c1 << 16), shw, total, a.nt);
    unsigned long long ar0 = block_sum(arcs[0], sh, a.nt);
    unsigned long long ar1 = block_sum(arcs[1], sh, a.nt);
    unsigned long long np = trims ? block_sum((unsigned long long)newp, sh, a.nt) : 0ull;
    if (threadIdx.x == 0) {
        const int slot = (a.iter + 1) & 3;
        sbase[0] = (total & 0xFFFFu) ? atomicAdd(&a.ctl->qsize[0][slot], total & 0xFFFFu) : 0u;
        sbase[1] = (total >> 16) ? atomicAdd(&a.ctl->qsize[1][slot], total >> 16) : 0u;
        if (ar0) atomicAdd(&a.ctl->qarcs[0][slot], ar0);
        if (ar1) atomicAdd(&a.ctl->qarcs[1][slot], ar1);
        if (trims && (total & 0xFFFFu)) atomicSub(&a.ctl->live, total & 0xFFFFu);
        if (np) atomicAdd(&a.ctl->pcount, (uint32_t)np);
    }
    __syncthreads();
    uint32_t o[2] = {sbase[0] + (pre & 0xFFFFu), sbase[1] + (pre >> 16)};
#pragma unroll
    for (int d = 0; d < 2; d++) {
        for (uint32_t m = claimed[d]; m; m &= m - 1) {
            const int64_t i = base + (int64_t)(__ffs(m) - 1) * a.nt + threadIdx.x;
            a.outq[d][o[d]++] = lst ? lst[i] : (int32_t)i;  // lst: bits index a vertex list
        }
    }
    __syncthreads();  // sbase is rewritten by the next round
}

// ---------------------------------------------------------------------------
// Pull sweep of one level (the directions in pull0 / pull1): every live vertex
// not yet reached in a pulling direction looks for a reached neighbour over its
// reverse row.  Rounds of PULL_V vertices per work-item (vertex base + k * nt +
// lane for bit k); each round loads its PULL_V words, giving the candidates per
// direction, then settles them in batches of PULL_T (vertex, direction) tasks:
//   heads -- the reverse row's first entry (from the head array) and its
//            visited bit (first FB step) or word; hits are claimed, and a miss
//            whose row holds more than its head is left for
//   rows  -- the rest of those rows, a batch in lockstep.
// A round with many candidates (at least PULL_DENSE per lane over its wave)
// takes the bits k in order, every lane the same k, so that a batch's loads and
// claims hit consecutive vertices (a few cache lines per wave); a sparse round
// packs each lane's own candidates into its batches instead.  Every load of a
// batch is at an address valid whatever the branch (an idle task reads a
// vertex of the round or beg, an empty head vertex 0, an idle scan the column
// array's first entry) and is selected afterwards, so the loads of a batch
// leave together.  No barrier until the round's claims are queued.
// ---------------------------------------------------------------------------
constexpr int PULL_T = 4;      // tasks per batch (more spill registers)
constexpr int PULL_DENSE = 4;  // candidates per lane (wave average) for the in-order batches

struct PullTasks {
    int kk[PULL_T];      // bit of the task's vertex in the round (-1: idle)
    int32_t v[PULL_T];   // its vertex (a valid id when idle)
    uint32_t dm;         // bit t: task t pulls B (else F)
};

// Claims of the tasks in hit: the words' atomics and the class bytes of the
// claimed rows (the next level's arc estimate) leave together.
__device__ inline void pull_claims(const ReachArgs &a, uint32_t hit, const PullTasks &q, uint32_t (&claimed)[2],
                                   unsigned long long (&arcs)[2]) {
    uint32_t old[PULL_T], cs[PULL_T];
#pragma unroll
    for (int t = 0; t < PULL_T; t++) {
        const bool b = (q.dm >> t) & 1u;
        old[t] = ST_F | ST_B;
        cs[t] = 0u;
        if ((hit >> t) & 1u) {
            old[t] = atomicOr(&a.word[q.v[t]], b ? ST_B : ST_F);
            cs[t] = (b ? a.dcls[1] : a.dcls[0])[q.v[t]];
        }
    }
#pragma unroll
    for (int t = 0; t < PULL_T; t++) {
        const bool b = (q.dm >> t) & 1u;
        if (old[t] & (b ? ST_B : ST_F)) continue;
        if (b) {
            claimed[1] |= 1u << q.kk[t];
            arcs[1] += class_arcs(cs[t]);
        } else {
            claimed[0] |= 1u << q.kk[t];
            arcs[0] += class_arcs(cs[t]);
        }
    }
}

// The heads of a batch; misses with more entries join pend0 / pend1.
__device__ inline void pull_heads(const ReachArgs &a, bool vis, const PullTasks &q, uint32_t &pend0, uint32_t &pend1,
                                  uint32_t (&claimed)[2], unsigned long long (&arcs)[2], unsigned long long &cx,
                                  unsigned long long &ce, uint32_t &smax) {
    int32_t h[PULL_T];
    uint32_t cr[PULL_T], wv[PULL_T], x[PULL_T];
#pragma unroll
    for (int t = 0; t < PULL_T; t++) {
        const bool b = (q.dm >> t) & 1u;
        h[t] = (b ? a.head[0] : a.head[1])[q.v[t]];   // head of the reverse row
        cr[t] = (b ? a.dcls[0] : a.dcls[1])[q.v[t]];  // the reverse row's length class
        wv[t] = vis ? 0u : a.word[q.v[t]];           // the candidate's colour
    }
#pragma unroll
    for (int t = 0; t < PULL_T; t++) {
        const int32_t u = h[t] >= 0 ? h[t] : 0;
        x[t] = vis ? (((q.dm >> t) & 1u) ? a.vis[1] : a.vis[0])[u >> 5] : a.word[u];
    }
    uint32_t hit = 0u;
#pragma unroll
    for (int t = 0; t < PULL_T; t++) {
        if (q.kk[t] < 0 || h[t] < 0) continue;  // an empty reverse row: never reached by a pull
        cx++;
        ce++;
        smax = smax > 1u ? smax : 1u;  // a scan of one entry
        const uint32_t bit = ((q.dm >> t) & 1u) ? ST_B : ST_F;
        const bool ok = vis ? ((x[t] >> (h[t] & 31)) & 1u) != 0u
                            : (x[t] & (bit | ST_E)) == bit && colour_of(x[t]) == colour_of(wv[t]);
        if (ok) {
            hit |= 1u << t;
        } else if ((cr[t] & 0x3Fu) >= 2u) {  // at least two entries
            if ((q.dm >> t) & 1u) pend1 |= 1u << q.kk[t];
            else pend0 |= 1u << q.kk[t];
        }
    }
    pull_claims(a, hit, q, claimed, arcs);
}

// Entries 1.. of a batch's rows (each of at least two entries): entries 1 .. 3 of
// every row at once from the four-entry heads (one 16-byte load per row, then
// their visited bits or words together: most short rows settle in one round
// trip), then entries 4.. of the longer rows in lockstep.
__device__ inline void pull_rows(const ReachArgs &a, bool vis, const PullTasks &q, uint32_t (&claimed)[2],
                                 unsigned long long (&arcs)[2], unsigned long long &ce, uint32_t &smax) {
    int32_t e[PULL_T], e1[PULL_T];
    uint32_t wv[PULL_T], act = 0u, hit = 0u;
    {
        int4 h4[PULL_T];
#pragma unroll
        for (int t = 0; t < PULL_T; t++) {
            h4[t] = (((q.dm >> t) & 1u) ? a.ell[0] : a.ell[1])[q.v[t]];  // B pulls over the out-row
            wv[t] = vis ? 0u : a.word[q.v[t]];
            if (q.kk[t] >= 0) act |= 1u << t;
        }
        uint32_t x[PULL_T][3];
#pragma unroll
        for (int t = 0; t < PULL_T; t++) {
            const int32_t u3[3] = {h4[t].y, h4[t].z, h4[t].w};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int32_t u = u3[k] >= 0 ? u3[k] : 0;  // -1 past the row's end
                x[t][k] = vis ? (((q.dm >> t) & 1u) ? a.vis[1] : a.vis[0])[u >> 5] : a.word[u];
            }
        }
#pragma unroll
        for (int t = 0; t < PULL_T; t++) {
            if (!((act >> t) & 1u)) continue;
            const uint32_t bit = ((q.dm >> t) & 1u) ? ST_B : ST_F;
            const int32_t u3[3] = {h4[t].y, h4[t].z, h4[t].w};
#pragma unroll
            for (int k = 0; k < 3; k++) {  // counted as a scan in order counts them: up to the first hit
                if (u3[k] < 0 || ((hit >> t) & 1u)) continue;
                ce++;
                const bool ok = vis ? ((x[t][k] >> (u3[k] & 31)) & 1u) != 0u
                                    : (x[t][k] & (bit | ST_E)) == bit && colour_of(x[t][k]) == colour_of(wv[t]);
                if (ok) hit |= 1u << t;
            }
        }
        act &= ~hit;
    }
#pragma unroll
    for (int t = 0; t < PULL_T; t++) {
        const int32_t *rw = ((q.dm >> t) & 1u) ? a.row : a.rowT;
        e[t] = rw[q.v[t]] + 4;  // the first four entries are checked
        e1[t] = rw[q.v[t] + 1];
        if (e[t] >= e1[t]) act &= ~(1u << t);
    }
    int s = 0;
    for (; act; s++) {
        int32_t u[PULL_T];
        uint32_t x[PULL_T];
#pragma unroll
        for (int t = 0; t < PULL_T; t++) u[t] = (((q.dm >> t) & 1u) ? a.col : a.colT)[((act >> t) & 1u) ? e[t] : 0];
#pragma unroll
        for (int t = 0; t < PULL_T; t++)
            x[t] = vis ? (((q.dm >> t) & 1u) ? a.vis[1] : a.vis[0])[u[t] >> 5] : a.word[u[t]];
#pragma unroll
        for (int t = 0; t < PULL_T; t++) {
            if (!((act >> t) & 1u)) continue;
            ce++;
            const uint32_t bit = ((q.dm >> t) & 1u) ? ST_B : ST_F;
            const bool ok = vis ? ((x[t] >> (u[t] & 31)) & 1u) != 0u
                                : (x[t] & (bit | ST_E)) == bit && colour_of(x[t]) == colour_of(wv[t]);
            if (ok) {
                hit |= 1u << t;
                act &= ~(1u << t);
            } else if (++e[t] >= e1[t]) {
                act &= ~(1u << t);
            }
        }
    }
    smax = (uint32_t)s + 4u > smax ? (uint32_t)s + 4u : smax;  // the head, entries 1 .. 3, then s more
    pull_claims(a, hit, q, claimed, arcs);
}

// The heads of bits k0 .. k0 + PULL_KB - 1 of a dense round, every lane the
// same bits (consecutive vertices per wave), both pulling directions at once:
// one class byte per row direction serves as the reverse row's length in one
// direction and the claimed row's arcs in the other, and a vertex reached both
// ways is claimed by one atomic.
constexpr int PULL_KB = 4;

__device__ inline void pull_dense(const ReachArgs &a, bool vis, bool pull0, bool pull1, int64_t base, int64_t beg,
                                  int64_t end, int k0, uint32_t cm0, uint32_t cm1, uint32_t &pend0, uint32_t &pend1,
                                  uint32_t (&claimed)[2], unsigned long long (&arcs)[2], unsigned long long &cx,
                                  unsigned long long &ce, uint32_t &smax) {
    int32_t v[PULL_KB], hf[PULL_KB], hb[PULL_KB];
    uint32_t c0[PULL_KB], c1[PULL_KB], wv[PULL_KB], xf[PULL_KB], xb[PULL_KB];
#pragma unroll
    for (int j = 0; j < PULL_KB; j++) {
        const int64_t vv = base + (int64_t)(k0 + j) * a.nt + threadIdx.x;
        v[j] = (int32_t)(vv < end ? vv : beg);
        hf[j] = pull0 ? a.head[1][v[j]] : -1;  // F pulls over the in-row
        hb[j] = pull1 ? a.head[0][v[j]] : -1;  // B over the out-row
        c0[j] = a.dcls[0][v[j]];
        c1[j] = a.dcls[1][v[j]];
        wv[j] = vis ? 0u : a.word[v[j]];
    }
#pragma unroll
    for (int j = 0; j < PULL_KB; j++) {
        const int32_t uf = hf[j] >= 0 ? hf[j] : 0, ub = hb[j] >= 0 ? hb[j] : 0;
        xf[j] = pull0 ? (vis ? a.vis[0][uf >> 5] : a.word[uf]) : 0u;
        xb[j] = pull1 ? (vis ? a.vis[1][ub >> 5] : a.word[ub]) : 0u;
    }
    uint32_t bits[PULL_KB], old[PULL_KB];
#pragma unroll
    for (int j = 0; j < PULL_KB; j++) {
        const int k = k0 + j;
        bits[j] = 0u;
        const bool cf = ((cm0 >> k) & 1u) && hf[j] >= 0, cb = ((cm1 >> k) & 1u) && hb[j] >= 0;
        const bool okf = cf && (vis ? ((xf[j] >> (hf[j] & 31)) & 1u) != 0u
                                    : (xf[j] & (ST_F | ST_E)) == ST_F && colour_of(xf[j]) == colour_of(wv[j]));
        const bool okb = cb && (vis ? ((xb[j] >> (hb[j] & 31)) & 1u) != 0u
                                    : (xb[j] & (ST_B | ST_E)) == ST_B && colour_of(xb[j]) == colour_of(wv[j]));
        cx += (cf ? 1u : 0u) + (cb ? 1u : 0u);
        ce += (cf ? 1u : 0u) + (cb ? 1u : 0u);
        if (cf || cb) smax = smax > 1u ? smax : 1u;  // a scan of one entry
        if (okf) bits[j] |= ST_F;
        else if (cf && (c1[j] & 0x3Fu) >= 2u) pend0 |= 1u << k;  // at least two entries
        if (okb) bits[j] |= ST_B;
        else if (cb && (c0[j] & 0x3Fu) >= 2u) pend1 |= 1u << k;
        old[j] = bits[j] ? atomicOr(&a.word[v[j]], bits[j]) : 0u;
    }
#pragma unroll
    for (int j = 0; j < PULL_KB; j++) {
        const int k = k0 + j;
        const uint32_t won = bits[j] & ~old[j];
        if (won & ST_F) {
            claimed[0] |= 1u << k;
            arcs[0] += class_arcs(c0[j]);
        }
        if (won & ST_B) {
            claimed[1] |= 1u << k;
            arcs[1] += class_arcs(c1[j]);
        }
    }
}

// Packs the next PULL_T tasks of one lane's masks (F first).
__device__ inline void pull_pack(PullTasks &q, uint32_t &m0, uint32_t &m1, int64_t base, uint32_t nt, int64_t beg) {
    q.dm = 0u;
#pragma unroll
    for (int t = 0; t < PULL_T; t++) {
        q.kk[t] = -1;
        if (m0) {
            q.kk[t] = __ffs(m0) - 1;
            m0 &= m0 - 1u;
        } else if (m1) {
            q.kk[t] = __ffs(m1) - 1;
            m1 &= m1 - 1u;
            q.dm |= 1u << t;
        }
        q.v[t] = q.kk[t] >= 0 ? (int32_t)(base + (int64_t)q.kk[t] * nt + threadIdx.x) : (int32_t)beg;
    }
}

__device__ inline void pull_sweep(const ReachArgs &a, bool pull0, bool pull1, Items &it, Stage &st,
                                  unsigned long long *sh, bool &found, unsigned long long &cx,
                                  unsigned long long &ce, unsigned long long &cv, uint32_t &smax) {
    const bool vis = a.vis[0] != nullptr;  // kernel-uniform
    int64_t beg, end;
    my_range(a.n, beg, end);
    for (int64_t base = beg; base < end; base += (int64_t)a.nt * PULL_V) {  // block-uniform
        uint32_t cm0 = 0u, cm1 = 0u;  // candidates per pulling direction
#pragma unroll
        for (int k = 0; k < PULL_V; k++) {
            const int64_t vv = base + (int64_t)k * a.nt + threadIdx.x;
            const uint32_t w = a.word[vv < end ? vv : beg];
            if (vv < end && !(w & ST_E)) {
                if (pull0 && !(w & ST_F)) cm0 |= 1u << k;
                if (pull1 && !(w & ST_B)) cm1 |= 1u << k;
            }
        }
        uint32_t claimed[2] = {0u, 0u};
        unsigned long long arcs[2] = {0ull, 0ull};
        uint32_t pend0 = 0u, pend1 = 0u;  // misses whose rows go on past the head
        uint32_t wc = (uint32_t)(__popc(cm0) + __popc(cm1));
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) wc += (uint32_t)__shfl_xor((int)wc, o, WAVE);
        PullTasks q;
        if (wc >= (uint32_t)(PULL_DENSE * WAVE)) {  // wave-uniform: bits in order, lanes in step
            for (int k0 = 0; k0 < PULL_V; k0 += PULL_KB) {
                const uint32_t sel = ((1u << PULL_KB) - 1u) << k0;
                if (!__any(((cm0 | cm1) & sel) != 0u)) continue;  // wave-uniform
                pull_dense(a, vis, pull0, pull1, base, beg, end, k0, cm0, cm1, pend0, pend1, claimed, arcs, cx, ce,
                           smax);
            }
        } else {
            while (cm0 | cm1) {  // per work-item
                pull_pack(q, cm0, cm1, base, a.nt, beg);
                pull_heads(a, vis, q, pend0, pend1, claimed, arcs, cx, ce, smax);
            }
        }
        while (pend0 | pend1) {  // per work-item
            pull_pack(q, pend0, pend1, base, a.nt, beg);
            pull_rows(a, vis, q, claimed, arcs, ce, smax);
        }
        if (claimed[0] | claimed[1]) {
            found = true;
            cv += (unsigned long long)(__popc(claimed[0]) + __popc(claimed[1]));
        }
        if (vis) {
            // the round's claims into the visited bitmaps: vertex base + k * nt + lane
            // of a wave is one of 64 consecutive vertices starting at a multiple of 64
            // (beg and nt are), two bitmap words, one atomic each
            const uint32_t lane = threadIdx.x & (WAVE - 1);
            const int64_t w0 = base + (int64_t)(threadIdx.x - lane);
#pragma unroll
            for (int d = 0; d < 2; d++) {
                uint32_t *bm = dvis(a, d);
#pragma unroll
                for (int k = 0; k < PULL_V; k++) {
                    const unsigned long long m = __ballot((claimed[d] >> k) & 1u);
                    if (m == 0) continue;  // wave-uniform
                    const int64_t v0 = w0 + (int64_t)k * a.nt;
                    if (lane == 0 && (uint32_t)m) atomicOr(&bm[v0 >> 5], (uint32_t)m);
                    if (lane == 32 && (uint32_t)(m >> 32)) atomicOr(&bm[(v0 >> 5) + 1], (uint32_t)(m >> 32));
                }
            }
        }
        commit_round(a, base, claimed, arcs, it.shw, st.base, sh);
    }
}

__global__ void k_reach_hybrid(ReachArgs a) {
    __shared__ ItemsLocal ul;
    __shared__ Stage st;
    __shared__ unsigned long long sh[16];
    Items &it = ul.it;
    LocalQ &lq = ul.lq;
    if (!level_begin(a, true)) return;
    st_init(st);
    __syncthreads();
    uint32_t qs0, qs1;
    bool pull0, pull1;
    pull_dirs(a, qs0, qs1, pull0, pull1);
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    uint32_t smax = 0;  // longest row scan of this work-item (instrumented)
    if (blockIdx.x == 0 && threadIdx.x < 2) {
        const uint32_t pm = (pull0 ? 1u : 0u) | (pull1 ? 2u : 0u);
        if (threadIdx.x == 0) {
            a.ctl->pull_mask = pm;
            a.ctl->pmask[a.iter & 3] = pm;
        }
        const int d = (int)threadIdx.x;  // vertices reached through this level's input
        const int pi = (a.iter - 1) & 3;
        a.ctl->creach[d][a.iter & 3] = (a.iter > 0 ? a.ctl->creach[d][pi] : 0u) + (d == 0 ? qs0 : qs1);
        // arcs explored top-down: the pushed frontiers' estimated arcs
        a.ctl->carcs[d][a.iter & 3] = (a.iter > 0 ? a.ctl->carcs[d][pi] : 0ull) +
                                      (((pm >> d) & 1u) ? 0ull : a.ctl->qarcs[d][a.iter & 3]);
    }
    huge_phase<true>(a, it, &st, found, cx, ce, cv);
    st_flush(st, a);
    // push part: queue items of the directions that do not pull
    uint32_t p0 = pull0 ? 0u : qs0, p1 = pull1 ? 0u : qs1;
    if (p0 + p1 > 0) queue_balanced(a, p0, p0 + p1, it, st, found, cx, ce, cv);
    if (pull0 || pull1) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            a.ctl->pull_levels += (pull0 ? 1u : 0u) + (pull1 ? 1u : 0u);
        pull_sweep(a, pull0, pull1, it, st, sh, found, cx, ce, cv, smax);
    }
    local_rounds(a, (pull0 || pull1) ? 0xFFFFFFFFu : qs0 + qs1, st, lq, found, cx, ce, cv);
    st_flush(st, a);
    chain_progress(a.ctl, a.iter, found);
    if (a.instr) {
        instr_flush(a.ctl, cx, ce, cv, sh);
        instr_maxscan(a, smax);
    }
}

// ---------------------------------------------------------------------------
// data-driven trim within colours (trim_kernel, scc_routines.cu:336-394, run to
// its fixpoint by trim(), :616-643): the fixpoint is unique, so instead of
// re-sweeping every live vertex per iteration, level 0 sweeps once and a later
// level re-checks only the live same-colour neighbours of the vertices trimmed
// in the level before (each at most once per level, by stamp) -- unless that
// level trimmed so many that sweeping again is cheaper.
// ---------------------------------------------------------------------------
__device__ inline bool live_nb(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                               const int32_t *__restrict__ head, const uint32_t *word, int32_t v, uint32_t c,
                               uint32_t &steps) {
    const int32_t h = head[v];  // the row's first entry, loaded with the row bounds
    const int32_t e0 = row[v], e1 = row[v + 1];
    if (h < 0) return false;
    steps++;
    if (h != v) {
        uint32_t wh = word[h];
        if (!(wh & ST_E) && colour_of(wh) == c) return true;
    }
    for (int32_t e = e0 + 1; e < e1; e++) {
        steps++;
        int32_t u = col[e];
        if (u == v) continue;
        uint32_t wu = word[u];
        if (!(wu & ST_E) && colour_of(wu) == c) return true;
    }
    return false;
}

// trims v (colour c) when it has no live same-colour in- or out-neighbour
__device__ inline bool trim_check(const ReachArgs &a, int32_t v, uint32_t c, uint32_t &smax) {
    uint32_t steps = 0;
    const bool keep = live_nb(a.row, a.col, a.head[0], a.word, v, c, steps) &&
                      live_nb(a.rowT, a.colT, a.head[1], a.word, v, c, steps);
    smax = steps > smax ? steps : smax;
    if (keep) return false;
    a.word[v] = make_word(c, 28);  // P|T|E (scc_routines.cu:381)
    a.root[v] = orig_id(a.oid, v);
    return true;
}

constexpr int TRIM_V = 16;           // vertices per work-item per round of a trim sweep
constexpr int32_t TRIM_SWEEP_DIV = 16;  // a level after one that trimmed >= n/16 vertices sweeps

// General words: the live vertices of the round scan their out- and in-rows
// for a live same-colour neighbour (other than themselves), SCAN_ILP vertices
// at a time in lockstep; those missing one in either direction are trimmed.
__device__ inline uint32_t trim_round_scan(const ReachArgs &a, int64_t base, int64_t end,
                                           unsigned long long &ce, uint32_t &newp, uint32_t &smax) {
    uint32_t cand = 0;
#pragma unroll
    for (int k = 0; k < TRIM_V; k++) {
        int64_t vv = base + (int64_t)k * a.nt + threadIdx.x;  // list index when a.live is given
        uint32_t w = vv < end ? a.word[a.live ? a.live[vv] : vv] : ST_E;
        if (!(w & ST_E) && !(a.skip_fb && (w & (ST_P | ST_F | ST_B)) == (ST_F | ST_B))) cand |= 1u << k;
    }
    constexpr int T = 2 * SCAN_ILP;  // task d * SCAN_ILP + j: candidate j, out-row (d = 0) or in-row
    uint32_t trimmed = 0;
    while (cand) {  // per work-item
        int32_t v[SCAN_ILP];
        int kk[SCAN_ILP];
#pragma unroll
        for (int j = 0; j < SCAN_ILP; j++) {
            kk[j] = cand ? __ffs(cand) - 1 : -1;
            cand &= cand ? cand - 1 : 0u;
            const int64_t i = base + (int64_t)kk[j] * a.nt + threadIdx.x;
            v[j] = kk[j] >= 0 ? (a.live ? a.live[i] : (int32_t)i) : -1;
        }
        // the words (colours) load together with the heads and row bounds: only
        // this work-item writes a candidate's word in a sweep level
        uint32_t w[SCAN_ILP];
#pragma unroll
        for (int j = 0; j < SCAN_ILP; j++) w[j] = v[j] >= 0 ? a.word[v[j]] : ST_E;
        const int32_t *col[T];
        int32_t e[T], e1[T], u0[T], self[T];
        uint32_t mask[T], val[T], c[T], act = 0;
#pragma unroll
        for (int t = 0; t < T; t++) {
            const int d = t / SCAN_ILP, j = t % SCAN_ILP;
            const int32_t *row = d == 0 ? a.row : a.rowT;
            col[t] = d == 0 ? a.col : a.colT;
            mask[t] = ST_E;
            val[t] = 0;
            self[t] = v[j];
            u0[t] = -1;
            e[t] = e1[t] = 0;
            if (v[j] >= 0) {
                u0[t] = a.head[d][v[j]];
                e[t] = row[v[j]] + 1;
                e1[t] = row[v[j] + 1];
                if (u0[t] >= 0) act |= 1u << t;
            }
        }
#pragma unroll
        for (int t = 0; t < T; t++) c[t] = colour_of(w[t % SCAN_ILP]);
        // live same-colour neighbours: bit j (out-row), bit SCAN_ILP + j (in-row)
        uint32_t hit = scan_rows<T>(a.word, col, e, e1, u0, mask, val, c, self, act, ce, smax);
#pragma unroll
        for (int j = 0; j < SCAN_ILP; j++) {
            if (v[j] >= 0 && (((hit >> j) & (hit >> (SCAN_ILP + j))) & 1u) == 0) {
                a.word[v[j]] = make_word(c[j], 28);  // P|T|E (scc_routines.cu:381)
                a.root[v[j]] = orig_id(a.oid, v[j]);
                trimmed |= 1u << kk[j];
                newp += (w[j] & ST_P) ? 0u : 1u;  // a trimmed pivot holds P already
            }
        }
    }
    return trimmed;
}

__global__ void k_trim_level(ReachArgs a) {
    __shared__ Items it;
    __shared__ Stage st;
    __shared__ unsigned long long sh[16];
    if (!level_begin(a, true)) return;
    st_init(st);
    __syncthreads();
    bool found = false;
    uint32_t smax = 0;  // longest row scan of this work-item (instrumented)
    const uint32_t qs = a.iter == 0 && !a.init_list ? 0u : a.ctl->qsize[0][a.iter & 3];
    if (a.init_list) {
        // the first level of trim1: only the vertices k_init listed (a first out- or
        // in-entry not known to stay live) are checked, each with full row scans
        unsigned long long ntrim = 0;
        const uint32_t chunk = (qs + gridDim.x - 1) / gridDim.x;
        const uint32_t b0 = blockIdx.x * chunk, b1 = b0 + chunk < qs ? b0 + chunk : qs;
        for (uint32_t base = b0; base < b1; base += a.nt) {  // block-uniform
            const uint32_t i = base + threadIdx.x;
            if (i < b1) {
                const int32_t v = a.inq[0][i];
                const uint32_t w = a.word[v];
                if (!(w & ST_E) && trim_check(a, v, colour_of(w), smax)) {
                    found = true;
                    ntrim++;
                    st_push(st, 0, v, a);
                }
            }
            st_maybe_flush(st, a);
        }
        ntrim = block_sum(ntrim, sh, a.nt);
        if (threadIdx.x == 0 && ntrim) {  // no pivots exist yet: every trim is a new P
            atomicSub(&a.ctl->live, (uint32_t)ntrim);
            atomicAdd(&a.ctl->pcount, (uint32_t)ntrim);
        }
    } else if (a.iter == 0 || qs >= (uint32_t)(a.n / TRIM_SWEEP_DIV)) {
        // sweep: rounds of TRIM_V vertices per work-item, no barrier until the
        // round's trims are queued; over the live-vertex list when there is one
        int64_t beg, end;
        my_range(a.live ? (int64_t)*a.live_cnt : (int64_t)a.n, beg, end);
        unsigned long long ce = 0;
        const unsigned long long none[2] = {0ull, 0ull};
        for (int64_t base = beg; base < end; base += (int64_t)a.nt * TRIM_V) {  // block-uniform
            uint32_t newp = 0;
            const uint32_t m = trim_round_scan(a, base, end, ce, newp, smax);
            if (m) found = true;
            const uint32_t claimed[2] = {m, 0u};
            commit_round(a, base, claimed, none, it.shw, st.base, sh, true, newp, a.live);
        }
    } else {
        // data-driven: re-check the live same-colour neighbours of the vertices
        // the previous level trimmed, each at most once (stamp = this level's tag)
        unsigned long long ntrim = 0, newp = 0;
        uint32_t chunk = (qs + gridDim.x - 1) / gridDim.x;
        uint32_t b0 = blockIdx.x * chunk, b1 = b0 + chunk < qs ? b0 + chunk : qs;
        for (uint32_t base = b0; base < b1; base += a.nt) {
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
            uint32_t off = block_excl_scan(deg[0] + deg[1], it.shw, total, a.nt);
            it.off[2 * threadIdx.x] = off;
            it.off[2 * threadIdx.x + 1] = off + deg[0];
            it.start[2 * threadIdx.x] = start[0];
            it.start[2 * threadIdx.x + 1] = start[1];
            it.cd[2 * threadIdx.x] = c;
            it.cd[2 * threadIdx.x + 1] = c | 0x80000000u;
            __syncthreads();
            for (uint32_t pb = 0; pb < total; pb += a.nt) {  // block-uniform
                uint32_t e = pb + threadIdx.x;
                if (e < total) {
                    uint32_t k = find_item(it.off, 2 * a.nt, e);
                    uint32_t cd = it.cd[k];
                    const int32_t *col = (cd >> 31) ? a.colT : a.col;
                    int32_t u = col[it.start[k] + (int32_t)(e - it.off[k])];
                    uint32_t cc = cd & 0x7FFFFFFFu;
                    uint32_t wu = a.word[u];
                    if (!(wu & ST_E) && colour_of(wu) == cc &&
                        !(a.skip_fb && (wu & (ST_P | ST_F | ST_B)) == (ST_F | ST_B)) &&
                        a.stamp[u] != a.tag &&  // plain load first: hubs are reached by many arcs
                        atomicExch(&a.stamp[u], a.tag) != a.tag && trim_check(a, u, cc, smax)) {
                        found = true;
                        ntrim++;
                        newp += (wu & ST_P) ? 0u : 1u;  // a trimmed pivot holds P already
                        st_push(st, 0, u, a);
                    }
                }
                st_maybe_flush(st, a);
            }
        }
        ntrim = block_sum(ntrim, sh, a.nt);
        newp = block_sum(newp, sh, a.nt);
        if (threadIdx.x == 0 && ntrim) atomicSub(&a.ctl->live, (uint32_t)ntrim);
        if (threadIdx.x == 0 && newp) atomicAdd(&a.ctl->pcount, (uint32_t)newp);
    }
    st_flush(st, a);
    chain_progress(a.ctl, a.iter, found);
    if (a.instr) instr_maxscan(a, smax);
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
static void launch_reach_kernel(K kernel, const Launch &l, const ReachArgs &args) {
    ReachArgs a = args;
    a.nt = l.block;
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

hipError_t preload_reach() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(k_reach_hybrid));
}

}  // namespace gpuscc
