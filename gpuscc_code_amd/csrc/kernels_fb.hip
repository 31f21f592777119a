// FB bookkeeping kernels for gfx950: vertex-word init, trim (trim-1 within
// colours), colour update + pivot election, trim-2, the WCC relabel pass,
// canonical labels and the small reductions.  All are grid-stride kernels on a
// persistent grid (a few workgroups per CU), so a level that has no work costs
// one short launch.
#include "fb_device.hpp"
#include "kernels.hpp"

namespace gpuscc {

// ---------------------------------------------------------------------------
// init: colour from init_colour (or 1, thrust::fill scc.cu:214), state 0,
// root = self
// ---------------------------------------------------------------------------
// graph of vertex v in a batch: largest g with off[g] <= v
__device__ inline uint32_t graph_of(const int32_t *off, int32_t G, int32_t v) {
    int32_t lo = 0, hi = G;
    while (hi - lo > 1) {
        int32_t mid = (lo + hi) >> 1;
        if (off[mid] <= v) lo = mid; else hi = mid;
    }
    return (uint32_t)lo;
}

// init: colour 1 (thrust::fill, scc.cu:214) or, for a batch, graph id + 1;
// state 0; root = self
// A row of degree class k (bit length of its length, 0x80: it holds v itself)
// has no entry other than v iff k is 0, or 1 with the self-loop flag.
__device__ inline bool row_has_other(uint32_t k) { return (k & 0x3Fu) != 0 && (k & 0xBFu) != 0x81u; }

// Fresh words: colour 1, or graph id + 1 in a batch; root = v.  With the
// degree classes given, the first level of trim1 (trim1_kernel,
// scc_routines.cu:276-332) is folded in: every arc lies inside one colour and
// every vertex is live, so v has a live same-colour neighbour in a direction
// iff its row there holds a vertex other than v -- a test of two class bytes.
// Vertices failing it start trimmed (P|T|E, scc_routines.cu:381).  Adds the
// live and trimmed counts to ctl->live and ctl->pcount (zeroed before).
// Rounds of INIT_V vertices per work-item: the class bytes of a round load
// together (a loop of one vertex per iteration waits for each load in turn).
// In a batch the graph of a work-item's first vertex is found by binary search
// and later vertices (increasing ids) advance from it.
constexpr int INIT_V = 8;

// With hlive given, the live vertices whose out- or in-row does not start with
// another vertex that stays live (hlive, k_head_live) are appended to unk (count
// in ctl->qsize[0][0]): the first level of trim1 checks only those, since every
// other live vertex has a live same-colour neighbour both ways.  VEC = 4 (n a
// multiple of 4): a work-item takes four consecutive vertices at a time, their
// class bytes in one 4-byte load each and their words and roots in one 16-byte
// store each.
template <int VEC>
__global__ void k_init(uint32_t *word, int32_t *root, int32_t n, const int32_t *off, int32_t G,
                       const uint8_t *cls_out, const uint8_t *cls_in, const uint8_t *hlive, int32_t *unk,
                       Ctl *ctl, uint32_t *vis, int64_t vis_words, int32_t *mn, const int32_t *oid) {
    constexpr int R = INIT_V / VEC;  // groups of VEC vertices per work-item and round
    __shared__ unsigned long long sh[16];
    __shared__ uint32_t shw[16];
    __shared__ uint32_t sbase;
    int64_t beg, end;
    my_range(n, beg, end);  // beg is a multiple of 64, end of 4 when n is
    const uint32_t bs = blockDim.x;
    unsigned long long nt = 0;
    uint32_t g = 0;  // batch: graph of the current vertex
    if (off && beg + (int64_t)threadIdx.x * VEC < end) g = graph_of(off, G, (int32_t)(beg + threadIdx.x * VEC));
    for (int64_t base = beg; base < end; base += (int64_t)bs * INIT_V) {  // block-uniform
        uint32_t co[R], ci[R], ho[R], hi[R];  // VEC bytes each
#pragma unroll
        for (int q = 0; q < R; q++) {
            const int64_t v0 = base + ((int64_t)q * bs + threadIdx.x) * VEC;
            const bool in = cls_out && v0 < end;
            if (VEC == 4) {
                co[q] = in ? *reinterpret_cast<const uint32_t *>(cls_out + v0) : 0u;
                ci[q] = in ? *reinterpret_cast<const uint32_t *>(cls_in + v0) : 0u;
                ho[q] = in && hlive ? *reinterpret_cast<const uint32_t *>(hlive + v0) : 0u;
                hi[q] = in && hlive ? *reinterpret_cast<const uint32_t *>(hlive + (int64_t)n + v0) : 0u;
            } else {
                co[q] = in ? cls_out[v0] : 0u;
                ci[q] = in ? cls_in[v0] : 0u;
                ho[q] = in && hlive ? hlive[v0] : 0u;
                hi[q] = in && hlive ? hlive[(int64_t)n + v0] : 0u;
            }
        }
        uint32_t um = 0;  // live vertices of this round the first trim level must check
#pragma unroll
        for (int q = 0; q < R; q++) {
            const int64_t v0 = base + ((int64_t)q * bs + threadIdx.x) * VEC;
            if (v0 >= end) continue;
            uint32_t w[VEC];
#pragma unroll
            for (int j = 0; j < VEC; j++) {
                const int64_t v = v0 + j;
                uint32_t c = 1u;
                if (off) {
                    while (off[g + 1] <= v) g++;
                    c = (g + 1u) & COLOUR_MASK;
                }
                const uint32_t bo = (co[q] >> (8 * j)) & 0xFFu, bi = (ci[q] >> (8 * j)) & 0xFFu;
                const bool trim = cls_out && !(row_has_other(bo) && row_has_other(bi));
                w[j] = make_word(c, trim ? 28u : 0u);
                nt += trim ? 1u : 0u;
                if (!trim && !(((ho[q] >> (8 * j)) & 0xFFu) && ((hi[q] >> (8 * j)) & 0xFFu)))
                    um |= 1u << (q * VEC + j);
            }
            if (vis && (v0 & 31) == 0) {  // the visited bitmaps: one word per 32 vertices
                vis[v0 >> 5] = 0u;
                vis[vis_words + (v0 >> 5)] = 0u;
            }
            if (VEC == 4) {
                *reinterpret_cast<uint4 *>(word + v0) = make_uint4(w[0], w[1 % VEC], w[2 % VEC], w[3 % VEC]);
                const int4 self = make_int4((int32_t)v0, (int32_t)v0 + 1, (int32_t)v0 + 2, (int32_t)v0 + 3);
                // roots in the caller's ids; mn is indexed by them, its identity is the same either way
                *reinterpret_cast<int4 *>(root + v0) = oid ? *reinterpret_cast<const int4 *>(oid + v0) : self;
                if (mn) *reinterpret_cast<int4 *>(mn + v0) = self;
            } else {
                word[v0] = w[0];
                root[v0] = orig_id(oid, (int32_t)v0);
                if (mn) mn[v0] = (int32_t)v0;
            }
        }
        if (hlive) {  // one prefix sum and one atomicAdd per round and workgroup
            uint32_t total;
            const uint32_t pre = block_excl_scan((uint32_t)__popc(um), shw, total, bs);
            if (total == 0) continue;  // block-uniform
            if (threadIdx.x == 0) sbase = atomicAdd(&ctl->qsize[0][0], total);
            __syncthreads();
            uint32_t o = sbase + pre;
            for (uint32_t m = um; m; m &= m - 1) {
                const int b = __ffs(m) - 1;
                unk[o++] = (int32_t)(base + ((int64_t)(b / VEC) * bs + threadIdx.x) * VEC + b % VEC);
            }
            __syncthreads();  // sbase is rewritten by the next round
        }
    }
    nt = block_sum(nt, sh);
    if (threadIdx.x == 0) {
        const unsigned long long cnt = end > beg ? (unsigned long long)(end - beg) : 0ull;
        if (cnt > nt) atomicAdd(&ctl->live, (uint32_t)(cnt - nt));
        if (nt) atomicAdd(&ctl->pcount, (uint32_t)nt);
    }
}

// counts arcs whose ends lie in different graphs of a batch
__global__ void k_batch_check(const int32_t *row, const int32_t *col, int32_t n, const int32_t *off,
                              int32_t G, Ctl *ctl) {
    __shared__ unsigned long long sh[16];
    int64_t beg, end;
    my_range(n, beg, end);
    unsigned long long bad = 0;
    for (int64_t v = beg + threadIdx.x; v < end; v += blockDim.x) {
        uint32_t g = graph_of(off, G, (int32_t)v);
        int32_t lo = off[g], hi = off[g + 1];
        for (int32_t e = row[v]; e < row[v + 1]; e++) bad += (col[e] < lo || col[e] >= hi);
    }
    bad = block_sum(bad, sh);
    if (threadIdx.x == 0 && bad) atomicAdd(&ctl->count, bad);
}

// calibration copy of known traffic: 16-byte loads and stores
__global__ void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_fill_u64(unsigned long long *p, int64_t count, unsigned long long val) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

__global__ void k_fill_i32(int32_t *p, int64_t count, int32_t val) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

// bits 0-5: bit length of the row's length; bit 6 (CLS_FITS): at most four
// entries, all held in ell[v]; bit 7: the row holds v itself
__global__ void k_degree_class(const int32_t *__restrict__ row, const int32_t *__restrict__ col, int32_t n,
                               uint8_t *out, int32_t *head, int4 *ell) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x) {
        int32_t e0 = row[v], e1 = row[v + 1];
        uint32_t deg = (uint32_t)(e1 - e0);
        bool self = false;
        for (int32_t e = e0; e < e1 && !self; e += 8) {  // eight column ids in flight per step
            int32_t u[8];
#pragma unroll
            for (int j = 0; j < 8; j++) u[j] = e + j < e1 ? col[e + j] : -1;
#pragma unroll
            for (int j = 0; j < 8; j++) self = self || u[j] == (int32_t)v;
        }
        out[v] = (uint8_t)((deg ? 32 - __clz(deg) : 0) | (deg <= 4 ? CLS_FITS : 0) | (self ? 0x80 : 0));
        head[v] = e0 < e1 ? col[e0] : -1;
        int4 q;
        q.x = e0 < e1 ? col[e0] : -1;
        q.y = e0 + 1 < e1 ? col[e0 + 1] : -1;
        q.z = e0 + 2 < e1 ? col[e0 + 2] : -1;
        q.w = e0 + 3 < e1 ? col[e0 + 3] : -1;
        ell[v] = q;
    }
}

// Upload-time index for the first sweep of trim1: whether the first entry of a
// row is a vertex that k_init leaves live.  Every arc joins two vertices of one
// colour when a run starts, so such a row already holds a live same-colour
// neighbour and the sweep needs no gather for it.
__global__ void k_head_live(const uint8_t *__restrict__ cls, const int32_t *__restrict__ head, int32_t n,
                            uint8_t *hlive) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * (int64_t)n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = (int32_t)(i < n ? i : i - n), h = head[i];
        hlive[i] = (h >= 0 && h != v && row_has_other(cls[h]) && row_has_other(cls[(int64_t)n + h])) ? 1 : 0;
    }
}

// Chain / counter resets (a kernel, not hipMemsetAsync, so it is safe in
// graphs), preceded by a chain guard when guard_stage > 0: the chain that
// ended with level guard_last records its depth in depth[guard_stage - 1] and
// sets abort = guard_stage when that level still made progress (the chain may
// need more levels).  With abort set the resets are skipped, unless bit 512
// clears it.  Bit order of application: 1024, 64, 128, 1, 2, 4, 8, 16, 32
// (seeds before the queue rings read them, npiv before it is cleared).
__device__ inline void ctl_reset_body(Ctl *ctl, int what, int guard_last, int guard_stage) {
    if (guard_stage > 0 && !ctl->abort) {
        uint32_t ran = ctl->ran;
        ctl->depth[guard_stage - 1] = ran;
        if (ran == (uint32_t)(guard_last + 1) && ctl->changed[guard_last & 3] != 0)
            ctl->abort = (uint32_t)guard_stage;
    }
    if (what & 512) ctl->abort = 0;  // the host has handled the stage that stopped
    else if (ctl->abort) return;
    if (what & 1024) {  // before k_init
        ctl->live = 0;
        ctl->pcount = 0;
        ctl->fbelim = 0;
        ctl->terr = 0;
    }
    if (what & 64) {  // the pivot list seeds both directions
        ctl->seeds[0] = ctl->npiv;
        ctl->seeds[1] = ctl->npiv;
    }
    if (what & 128) {
        ctl->seeds[0] = 0;
        ctl->seeds[1] = 0;
    }
    if (what & 1) {
        for (int i = 0; i < 4; i++) ctl->changed[i] = 0;
        for (int d = 0; d < 2; d++)
            for (int i = 0; i < 4; i++) ctl->hsize[d][i] = 0;
        ctl->ran = 0;
        ctl->pull_levels = 0;
        for (int i = 0; i < 4; i++) ctl->st_maxscan[i] = 0;
    }
    if (what & 2) {
        // queue chains: level 0 reads the seed lists
        for (int d = 0; d < 2; d++) {
            ctl->qsize[d][0] = ctl->seeds[d];
            for (int i = 1; i < 4; i++) ctl->qsize[d][i] = 0;
            for (int i = 0; i < 4; i++) ctl->qarcs[d][i] = 0;
            for (int i = 0; i < 4; i++) ctl->creach[d][i] = 0;
            for (int i = 0; i < 4; i++) ctl->carcs[d][i] = 0;
        }
        for (int i = 0; i < 4; i++) ctl->pmask[i] = 0;
    }
    if (what & 4) {
        ctl->npiv_prev = ctl->npiv;
        ctl->first_piv_prev = ctl->first_piv;
        ctl->npiv = 0;
        ctl->first_piv = 0xFFFFFFFFu;
    }
    if (what & 2048) ctl->lcnt[0] = 0;  // live-vertex list 0 is about to be rebuilt
    if (what & 4096) ctl->lcnt[1] = 0;
    if (what & 8) ctl->count = 0;
    if (what & 16) ctl->maxc = 0;
    if (what & 32) {
        ctl->st_expanded = 0;
        ctl->st_edges = 0;
        ctl->st_visited = 0;
        ctl->st_truns = 0;
        ctl->st_trounds = 0;
        ctl->st_tbusy = ctl->st_tidle = 0;
        for (int i = 0; i < 4; i++) ctl->st_tailclk[i] = 0;
        ctl->tail_L = ctl->tail_lstep = 0;
        for (int i = 0; i < 4; i++) {
            ctl->st_tclk[i] = 0;
            ctl->tlog[i][0] = ctl->tlog[i][1] = ctl->tlog[i][2] = 0;
        }
    }
    if (what & 8192) {  // before a batch launch: work counter, statistics, first pivot
        ctl->wnext = 0;
        for (int i = 0; i < 4; i++) ctl->tail[i] = 0;
        ctl->tail_pivots = 0;
        ctl->first_piv = 0xFFFFFFFFu;
    }
}

// With publish (the host's pinned mirror of the block), the block is copied
// there after the resets, so a host wait needs no separate copy.
__global__ void k_ctl_reset(Ctl *ctl, int what, int guard_last, int guard_stage, Ctl *publish) {
    if (threadIdx.x == 0) ctl_reset_body(ctl, what, guard_last, guard_stage);
    if (!publish) return;
    __syncthreads();
    const uint32_t *src = reinterpret_cast<const uint32_t *>(ctl);
    uint32_t *dst = reinterpret_cast<uint32_t *>(publish);
    for (uint32_t i = threadIdx.x; i < sizeof(Ctl) / 4; i += blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// colour update + election: update_kernel (scc_routines.cu:187-224).
// recolour: live v with F and B joins the SCC of the pivot elected for its
// colour last time (E set, root = that pivot); other live v take colour
// 3c+[F]+2[B] and state 0.  Every live v then offers its key to the slot of
// its colour with a 64-bit atomicMin (lowest key wins, deterministic).
// ---------------------------------------------------------------------------
constexpr int UPD_V = 8;  // vertices per work-item per round: their loads are issued together

// Live-vertex lists: with lin, the vertices processed are lin[0 .. *lin_cnt)
// instead of all of [0, n); with lout, the vertices still live after the
// update are appended to lout (count *lout_cnt, zeroed before).  After the
// first FB step few vertices stay live (a mesh keeps 0.4 %), and each later
// step's update, confirm and first trim level then cost that many, not n.
template <int K>
__device__ inline void list_append(uint32_t keep, const int32_t (&vid)[K], int32_t *out, uint32_t *out_cnt,
                                   uint32_t *shw, uint32_t &sbase) {
    uint32_t total;  // all work-items call this (two barriers)
    const uint32_t pre = block_excl_scan((uint32_t)__popc(keep), shw, total);
    if (total == 0) return;  // block-uniform
    if (threadIdx.x == 0) sbase = atomicAdd(out_cnt, total);
    __syncthreads();
    uint32_t o = sbase + pre;
#pragma unroll
    for (int k = 0; k < K; k++)  // static indices keep vid in registers
        if ((keep >> k) & 1u) out[o++] = vid[k];
    __syncthreads();  // sbase is rewritten by the next call
}

// mn (may be null): an FB-eliminated
// vertex lowers mn[its pivot] to its own id (mn[v] = v before the run), so that
// mn[root] is the minimum of each SCC whatever pivot the election chose.
__global__ void k_update(uint32_t *word, int32_t *root, int32_t n,
                         const unsigned long long *__restrict__ prev, unsigned long long *cur,
                         int slot_bits, unsigned long long tag, int recolour, int elect,
                         const uint16_t *__restrict__ score, Ctl *ctl, const int32_t *lin,
                         const uint32_t *lin_cnt, int32_t *lout, uint32_t *lout_cnt, int32_t *mn,
                         const int32_t *oid) {
    __shared__ unsigned long long sh[16];
    __shared__ uint32_t wslot[16];
    __shared__ unsigned long long wmin[16];
    __shared__ uint32_t shw[16];
    __shared__ uint32_t sbase;
    if (ctl && __atomic_load_n(&ctl->abort, __ATOMIC_RELAXED)) return;
    int64_t beg, end;
    my_range(lin ? (int64_t)*lin_cnt : (int64_t)n, beg, end);
    const int lane = threadIdx.x & (WAVE - 1);
    unsigned long long gone = 0;  // F and B: eliminated into the pivot's SCC
    // wave-uniform pending minimum of one slot, carried across rounds
    uint32_t pslot = 0xFFFFFFFFu;
    unsigned long long pmin = ~0ull;
    for (int64_t base = beg; base < end; base += (int64_t)blockDim.x * UPD_V) {  // block-uniform
        uint32_t w[UPD_V];
        int32_t vid[UPD_V];
#pragma unroll
        for (int k = 0; k < UPD_V; k++) {
            int64_t i = base + (int64_t)k * blockDim.x + threadIdx.x;
            vid[k] = i < end ? (lin ? lin[i] : (int32_t)i) : -1;
        }
#pragma unroll
        for (int k = 0; k < UPD_V; k++) w[k] = vid[k] >= 0 ? word[vid[k]] : ST_E;
        uint32_t cmask = 0, keep = 0, sl[UPD_V];
        unsigned long long key[UPD_V];
#pragma unroll
        for (int k = 0; k < UPD_V; k++) {
            sl[k] = 0;
            key[k] = ~0ull;
            if (w[k] & ST_E) continue;
            int32_t v = vid[k];
            const int32_t ov = orig_id(oid, v);
            uint32_t c = colour_of(w[k]);
            if (recolour) {
                uint32_t f = ((w[k] & ST_F) ? 1u : 0u) + ((w[k] & ST_B) ? 2u : 0u);
                if (f == 3) {
                    word[v] = w[k] | ST_E;
                    uint32_t pv = (uint32_t)prev[slot_of(c, slot_bits)];
                    root[v] = pv < (uint32_t)n ? (int32_t)pv : ov;
                    if (mn && pv < (uint32_t)n && ov < mn[pv]) atomicMin(&mn[pv], ov);
                    gone++;
                    continue;
                }
                c = (3u * c + f) & COLOUR_MASK;
                word[v] = make_word(c, 0);
            }
            keep |= 1u << k;
            if (elect) {
                cmask |= 1u << k;
                sl[k] = slot_of(c, slot_bits);
                uint32_t sc = score ? (uint32_t)(0xFFFFu - score[v]) : 0u;
                key[k] = tag | ((unsigned long long)sc << 32) | (uint32_t)ov;
            }
        }
        if (lout) list_append(keep, vid, lout, lout_cnt, shw, sbase);
        // When every candidate of the wave targets one slot (the first election
        // sends all live vertices there), their minimum joins the wave's
        // pending minimum (combined over the workgroup at the end: one
        // atomicMin per workgroup); otherwise one atomicMin per candidate.
        uint32_t sl_lane = cmask ? sl[__ffs(cmask) - 1] : 0u;
        bool lane_same = true;
        unsigned long long lane_min = ~0ull;
#pragma unroll
        for (int k = 0; k < UPD_V; k++) {
            if (!((cmask >> k) & 1u)) continue;
            lane_same = lane_same && sl[k] == sl_lane;
            lane_min = key[k] < lane_min ? key[k] : lane_min;
        }
        unsigned long long bm = __ballot(cmask != 0);
        if (bm == 0) continue;  // wave-uniform (the loop's trip count is block-uniform)
        int first = __ffsll((long long)bm) - 1;
        uint32_t sl0 = __shfl(sl_lane, first, WAVE);
        if (__all(cmask == 0 || (lane_same && sl_lane == sl0))) {
            unsigned long long m = lane_min;
            for (int o = WAVE / 2; o > 0; o >>= 1) {
                unsigned long long x = __shfl_xor(m, o, WAVE);
                m = x < m ? x : m;
            }
            if (sl0 != pslot) {
                if (pslot != 0xFFFFFFFFu && lane == 0 && pmin < cur[pslot]) atomicMin(&cur[pslot], pmin);
                pslot = sl0;
                pmin = m;
            } else {
                pmin = m < pmin ? m : pmin;
            }
        } else {
#pragma unroll
            for (int k = 0; k < UPD_V; k++)
                if (((cmask >> k) & 1u) && key[k] < cur[sl[k]]) atomicMin(&cur[sl[k]], key[k]);
        }
    }
    // the waves' pending minima: one atomicMin when they share their slot
    const int wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    if (lane == 0) {
        wslot[wid] = pslot;
        wmin[wid] = pmin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s0 = 0xFFFFFFFFu;
        unsigned long long m0 = ~0ull;
        bool one = true;
        for (int i = 0; i < nw; i++) {
            if (wslot[i] == 0xFFFFFFFFu) continue;
            if (s0 == 0xFFFFFFFFu) s0 = wslot[i];
            one = one && wslot[i] == s0;
            m0 = wmin[i] < m0 ? wmin[i] : m0;
        }
        if (one && s0 != 0xFFFFFFFFu && m0 < cur[s0]) atomicMin(&cur[s0], m0);
        wslot[0] = one ? 0xFFFFFFFFu : 0u;  // 0: the waves flush their own
    }
    __syncthreads();
    if (wslot[0] == 0u && lane == 0 && pslot != 0xFFFFFFFFu && pmin < cur[pslot]) atomicMin(&cur[pslot], pmin);
    if (recolour && ctl) {
        gone = block_sum(gone, sh);
        if (threadIdx.x == 0 && gone) {
            atomicSub(&ctl->live, (uint32_t)gone);
            atomicAdd(&ctl->fbelim, (uint32_t)gone);
        }
    }
}

// Winners become pivots (state 19 = P|F|B, scc_routines.cu:219) and are
// appended to the pivot list through an LDS-staged block queue.
constexpr int CONF_V = 16;  // vertices per work-item per round

__global__ void k_confirm(uint32_t *word, int32_t n, const unsigned long long *__restrict__ cur,
                          int slot_bits, int32_t *piv, Ctl *ctl, const int32_t *lin, const uint32_t *lin_cnt,
                          uint32_t *vis, int64_t vis_words, const int32_t *oid) {
    __shared__ uint32_t shw[16];
    __shared__ uint32_t sbase;
    if (__atomic_load_n(&ctl->abort, __ATOMIC_RELAXED)) return;
    int64_t beg, end;
    my_range(lin ? (int64_t)*lin_cnt : (int64_t)n, beg, end);
    for (int64_t base = beg; base < end; base += (int64_t)blockDim.x * CONF_V) {  // block-uniform
        uint32_t w[CONF_V];
        int32_t vid[CONF_V];
#pragma unroll
        for (int k = 0; k < CONF_V; k++) {
            int64_t i = base + (int64_t)k * blockDim.x + threadIdx.x;
            vid[k] = i < end ? (lin ? lin[i] : (int32_t)i) : -1;
        }
#pragma unroll
        for (int k = 0; k < CONF_V; k++) w[k] = vid[k] >= 0 ? word[vid[k]] : ST_E;
        uint32_t won = 0;
#pragma unroll
        for (int k = 0; k < CONF_V; k++) {
            if (w[k] & ST_E) continue;
            uint32_t v = (uint32_t)vid[k];
            uint32_t c = colour_of(w[k]);
            const uint32_t ov = (uint32_t)orig_id(oid, (int32_t)v);
            if ((uint32_t)cur[slot_of(c, slot_bits)] == ov) {
                word[v] = make_word(c, 19);
                won |= 1u << k;
                atomicMin(&ctl->first_piv, ov);
                if (vis) {  // a pivot is reached both ways
                    atomicOr(&vis[v >> 5], 1u << (v & 31));
                    atomicOr(&vis[vis_words + (v >> 5)], 1u << (v & 31));
                }
            }
        }
        // the round's winners join the pivot list: one prefix sum, one atomicAdd
        uint32_t total;
        uint32_t pre = block_excl_scan((uint32_t)__popc(won), shw, total);
        if (total == 0) continue;  // block-uniform
        if (threadIdx.x == 0) {
            sbase = atomicAdd(&ctl->npiv, total);
            atomicAdd(&ctl->pcount, total);  // live vertices hold no P before they win
        }
        __syncthreads();
        uint32_t o = sbase + pre;
#pragma unroll
        for (int k = 0; k < CONF_V; k++)
            if ((won >> k) & 1u) piv[o++] = vid[k];
        __syncthreads();  // sbase is rewritten by the next round
    }
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// what 0: FB-eliminated not trimmed ((s&12)==4, countElimmed scc.cu:42-51)
// what 1: P bits (#SCC, scc.cu:914-919)
__global__ void k_count_state(const uint32_t *word, int32_t n, int what, Ctl *ctl) {
    __shared__ unsigned long long sh[16];
    int64_t beg, end;
    my_range(n, beg, end);
    unsigned long long c = 0;
#pragma unroll 8
    for (int64_t v = beg + threadIdx.x; v < end; v += blockDim.x) {
        uint32_t s = word[v] & 0xFFu;
        if (what == 0) c += ((s & 12u) == 4u);
        else c += ((s & ST_P) != 0);
    }
    c = block_sum(c, sh);
    if (threadIdx.x == 0 && c) atomicAdd(&ctl->count, c);
}

// max colour over all vertices (thrust::reduce maximum, scc.cu:802)
__global__ void k_max_colour(const uint32_t *word, int32_t n, Ctl *ctl) {
    __shared__ uint32_t sh[16];
    int64_t beg, end;
    my_range(n, beg, end);
    uint32_t m = 0;
#pragma unroll 8
    for (int64_t v = beg + threadIdx.x; v < end; v += blockDim.x) {
        uint32_t c = colour_of(word[v]);
        m = c > m ? c : m;
    }
    m = block_max(m, sh);
    if (threadIdx.x == 0) atomicMax(&ctl->maxc, m);
}

// state 19 -> 0: unsetPivots_kernel (scc_routines.cu:10-19)
// With piv (the pivot list of the last election, ctl->npiv entries) only those
// vertices are visited: after an election they are the only ones in state 19.
__global__ void k_unset_pivots(uint32_t *word, int32_t n, Ctl *ctl, const int32_t *piv) {
    __shared__ unsigned long long sh[16];
    int64_t beg, end;
    my_range(piv ? (int64_t)ctl->npiv : (int64_t)n, beg, end);
    unsigned long long unset = 0;
    for (int64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const int32_t v = piv ? piv[i] : (int32_t)i;
        uint32_t w = word[v];
        if ((w & 0xFFu) == 19u) {
            word[v] = w & ~0xFFu;
            unset++;
        }
    }
    unset = block_sum(unset, sh);
    if (threadIdx.x == 0 && unset) atomicSub(&ctl->pcount, (uint32_t)unset);
}

// ---------------------------------------------------------------------------
// trim-2, one pass: trim2_kernel (scc_routines.cu:397-523).  Pass A marks
// pairs with ST_PAIR (not E), so every test sees the states as they were
// before the pass; pass B commits 20 (P|E) to the lower id and 4 (E) to the
// other.
// ---------------------------------------------------------------------------
__device__ inline int32_t only_nb(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                  const uint32_t *word, int32_t v, uint32_t c) {
    int32_t found = -1;
    for (int32_t e = row[v]; e < row[v + 1]; e++) {
        int32_t u = col[e];
        if (u == v) continue;
        uint32_t wu = word[u];
        if ((wu & ST_E) || colour_of(wu) != c) continue;
        if (found >= 0) return -2;
        found = u;
    }
    return found;
}

__global__ void k_trim2_mark(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                             const int32_t *__restrict__ rowT, const int32_t *__restrict__ colT,
                             uint32_t *word, int32_t *root, int32_t n, const int32_t *oid) {
    int64_t beg, end;
    my_range(n, beg, end);
    for (int64_t vv = beg + threadIdx.x; vv < end; vv += blockDim.x) {
        int32_t v = (int32_t)vv;
        uint32_t w = word[v];
        if (w & ST_E) continue;
        uint32_t c = colour_of(w);
        int32_t u = -1;
        int32_t a = only_nb(row, col, word, v, c);
        if (a >= 0 && only_nb(row, col, word, a, c) == v) u = a;
        if (u < 0) {
            a = only_nb(rowT, colT, word, v, c);
            if (a >= 0 && only_nb(rowT, colT, word, a, c) == v) u = a;
        }
        if (u < 0) continue;
        const int32_t ov = orig_id(oid, v), ou = orig_id(oid, u);  // the pair's lower id by the caller's ids
        word[v] = make_word(c, ST_PAIR | (ov < ou ? ST_P : 0u));
        root[v] = ov < ou ? ov : ou;
    }
}

__global__ void k_trim2_commit(uint32_t *word, int32_t n, Ctl *ctl) {
    __shared__ unsigned long long sh[16];
    int64_t beg, end;
    my_range(n, beg, end);
    unsigned long long pairs = 0;
    for (int64_t v = beg + threadIdx.x; v < end; v += blockDim.x) {
        uint32_t w = word[v];
        if (!(w & ST_PAIR)) continue;
        uint32_t p = w & ST_P;
        word[v] = make_word(colour_of(w), p | ST_E);
        pairs += p ? 1 : 0;
    }
    pairs = block_sum(pairs, sh);
    if (threadIdx.x == 0 && pairs) {
        atomicAdd(&ctl->count, pairs);
        atomicSub(&ctl->live, (uint32_t)(2 * pairs));
        atomicAdd(&ctl->pcount, (uint32_t)pairs);
    }
}

// ---------------------------------------------------------------------------
// one WCC pass: computeWCC (wcc.cu:7-52) to its fixpoint gives each live v
// the minimum id among live vertices it reaches over same-colour out-edges;
// pointer jumping (wcc.cu:42-46) shortcuts chains.  With a live-vertex list
// (lin, *lin_cnt; null: all n) only the listed vertices are visited: labels are
// ids of live vertices, so no other entry of lab is ever read.
// ---------------------------------------------------------------------------
// Labels are the caller's ids (oid, nid: the renumbering, see orig_id), so the
// minima and the colours they give are those of the graph as given; a pointer
// jump reads the label of the vertex a label names through nid.
__global__ void k_wcc_init(int32_t *lab, int32_t n, const int32_t *lin, const uint32_t *lin_cnt,
                           const int32_t *oid) {
    int64_t beg, end;
    my_range(lin ? (int64_t)*lin_cnt : (int64_t)n, beg, end);
    for (int64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const int32_t v = lin ? lin[i] : (int32_t)i;
        lab[v] = orig_id(oid, v);
    }
}

__global__ void k_wcc_iter(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                           const uint32_t *__restrict__ word, int32_t *lab, int32_t n, Ctl *ctl,
                           int iter, const int32_t *lin, const uint32_t *lin_cnt, const int32_t *nid) {
    if (chain_exit(ctl, iter)) return;
    int64_t beg, end;
    my_range(lin ? (int64_t)*lin_cnt : (int64_t)n, beg, end);
    bool found = false;
    for (int64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        int32_t v = lin ? lin[i] : (int32_t)i;
        uint32_t w = word[v];
        if (w & ST_E) continue;
        uint32_t c = colour_of(w);
        int32_t cur = lab[v], m = cur;
        for (int32_t e = row[v]; e < row[v + 1]; e++) {
            int32_t u = col[e];
            uint32_t wu = word[u];
            if ((wu & ST_E) || colour_of(wu) != c) continue;
            int32_t lu = lab[u];
            m = lu < m ? lu : m;
        }
        int32_t j = lab[nid ? nid[m] : m];  // pointer jump: the label of a reached vertex is reached too
        m = j < m ? j : m;
        if (m < cur) {
            lab[v] = m;
            found = true;
        }
    }
    chain_progress(ctl, iter, found);
}

// fresh colours per class (update_ranges2/3, wcc.cu:79-144): every member of
// the class rooted at r takes (maxc + 1 + r) & mask, state 0.
__global__ void k_wcc_recolour(uint32_t *word, const int32_t *lab, int32_t n, const Ctl *ctl, const int32_t *lin,
                               const uint32_t *lin_cnt) {
    uint32_t maxc = ctl->maxc;
    int64_t beg, end;
    my_range(lin ? (int64_t)*lin_cnt : (int64_t)n, beg, end);
    for (int64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const int32_t v = lin ? lin[i] : (int32_t)i;
        uint32_t w = word[v];
        if (w & ST_E) continue;
        word[v] = make_word((maxc + 1u + (uint32_t)lab[v]) & COLOUR_MASK, 0);
    }
}

// root[v] = mn[root[v]], in place: each work-item reads and writes only its own
// root, and writes only where the minimum differs from the pivot.  Trimmed
// vertices (T) are left out: every trim sets the root to the vertex itself, and
// mn of a trimmed vertex is its own id -- only FB-eliminated vertices lower the
// entry of their pivot, and a pivot that is trimmed has no such vertex (one
// reached both ways would share a cycle with it and keep it live) -- so the
// gather, scattered in the caller's numbering, is needed only for the rest.
__global__ void k_root_to_min(int32_t *root, const int32_t *__restrict__ mn, const uint32_t *__restrict__ word,
                              int32_t n) {
    int64_t beg, end;
    my_range(n, beg, end);
#pragma unroll 8
    for (int64_t v = beg + threadIdx.x; v < end; v += blockDim.x) {
        if (word[v] & ST_T) continue;
        const int32_t r = root[v], m = mn[r];
        if (m != r) root[v] = m;
    }
}

// out[i] = in[idx[i]]: labels (or any per-vertex array) from internal to the
// caller's order, idx being nid
__global__ void k_gather_i32(const int32_t *__restrict__ in, const int32_t *__restrict__ idx, int32_t *out,
                             int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[idx[i]];
}

// Renumbering of a CSR: row v of (row, col) becomes row map[v] of (nrow, ncol),
// its entries renamed through map, in the same order (nrow from the host: the
// prefix sums of the lengths in the new order).  One wave per row, its lanes
// copying 64 entries at a time.
__global__ void k_permute_rows(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                               const int32_t *__restrict__ map, const int32_t *__restrict__ nrow,
                               int32_t *__restrict__ ncol, int32_t n) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t v = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; v < n; v += waves) {
        const int32_t e0 = row[v], e1 = row[v + 1], o = nrow[map[v]];
        for (int32_t e = e0 + lane; e < e1; e += WAVE) ncol[o + (e - e0)] = map[col[e]];
    }
}

// seeds of a reach given as vertex words: live vertices holding `bit`
// (seed list length accumulates in ctl->seeds[bit == ST_F ? 0 : 1])
__global__ void k_collect(const uint32_t *word, int32_t n, uint32_t bit, int32_t *q, Ctl *ctl) {
    __shared__ int32_t sq[1024];
    __shared__ uint32_t scount, sbase;
    int64_t beg, end;
    my_range(n, beg, end);
    for (int64_t base = beg; base < end; base += blockDim.x) {
        if (threadIdx.x == 0) scount = 0;
        __syncthreads();
        int64_t v = base + threadIdx.x;
        if (v < end) {
            uint32_t w = word[v];
            if ((w & (bit | ST_E)) == bit) sq[atomicAdd(&scount, 1u)] = (int32_t)v;
        }
        __syncthreads();
        if (scount > 0) {
            if (threadIdx.x == 0) sbase = atomicAdd(&ctl->seeds[bit == ST_F ? 0 : 1], scount);
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < scount; i += blockDim.x) q[sbase + i] = sq[i];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define GRID(l) dim3((l).grid), dim3((l).block), 0, (l).stream

void launch_init(const Launch &l, uint32_t *word, int32_t *root, int32_t n, const int32_t *off,
                 int32_t G, const uint8_t *cls, const uint8_t *hlive, int32_t *unk, Ctl *ctl, uint32_t *vis,
                 int64_t vis_words, int32_t *mn, const int32_t *oid) {
    if ((n & 3) == 0)
        hipLaunchKernelGGL(k_init<4>, GRID(l), word, root, n, off, G, cls, cls ? cls + n : nullptr,
                           cls ? hlive : nullptr, unk, ctl, vis, vis_words, mn, oid);
    else
        hipLaunchKernelGGL(k_init<1>, GRID(l), word, root, n, off, G, cls, cls ? cls + n : nullptr,
                           cls ? hlive : nullptr, unk, ctl, vis, vis_words, mn, oid);
}
void launch_batch_check(const Launch &l, const Graph &g, const int32_t *off, int32_t G, Ctl *ctl) {
    hipLaunchKernelGGL(k_batch_check, GRID(l), g.row, g.col, g.n, off, G, ctl);
}
void launch_stream_copy(const Launch &l, const void *src, void *dst, int64_t bytes) {
    hipLaunchKernelGGL(k_stream_copy, GRID(l), (const uint4 *)src, (uint4 *)dst, bytes / 16);
}
void launch_fill_u64(const Launch &l, unsigned long long *p, int64_t count, unsigned long long v) {
    hipLaunchKernelGGL(k_fill_u64, GRID(l), p, count, v);
}
void launch_fill_i32(const Launch &l, int32_t *p, int64_t count, int32_t v) {
    hipLaunchKernelGGL(k_fill_i32, GRID(l), p, count, v);
}
void launch_degree_class(const Launch &l, const int32_t *row, const int32_t *col, int32_t n, uint8_t *out,
                         int32_t *head, int4 *ell) {
    hipLaunchKernelGGL(k_degree_class, GRID(l), row, col, n, out, head, ell);
}
void launch_head_live(const Launch &l, const uint8_t *cls, const int32_t *head, int32_t n, uint8_t *hlive) {
    hipLaunchKernelGGL(k_head_live, GRID(l), cls, head, n, hlive);
}
void launch_ctl_reset(hipStream_t s, Ctl *ctl, int what, int guard_last, int guard_stage, Ctl *publish) {
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, s, ctl, what, guard_last, guard_stage, publish);
}
void launch_update(const Launch &l, uint32_t *word, int32_t *root, int32_t n,
                   const unsigned long long *prev, unsigned long long *cur, int slot_bits,
                   unsigned long long tag, int recolour, int elect, const uint16_t *score, Ctl *ctl,
                   const int32_t *lin, const uint32_t *lin_cnt, int32_t *lout, uint32_t *lout_cnt, int32_t *mn,
                   const int32_t *oid) {
    hipLaunchKernelGGL(k_update, GRID(l), word, root, n, prev, cur, slot_bits, tag, recolour, elect,
                       score, ctl, lin, lin_cnt, lout, lout_cnt, mn, oid);
}
void launch_confirm(const Launch &l, uint32_t *word, int32_t n, const unsigned long long *cur,
                    int slot_bits, int32_t *piv, Ctl *ctl, const int32_t *lin, const uint32_t *lin_cnt,
                    uint32_t *vis, int64_t vis_words, const int32_t *oid) {
    hipLaunchKernelGGL(k_confirm, GRID(l), word, n, cur, slot_bits, piv, ctl, lin, lin_cnt, vis, vis_words, oid);
}
void launch_count_state(const Launch &l, const uint32_t *word, int32_t n, int what, Ctl *ctl) {
    hipLaunchKernelGGL(k_count_state, GRID(l), word, n, what, ctl);
}
void launch_max_colour(const Launch &l, const uint32_t *word, int32_t n, Ctl *ctl) {
    hipLaunchKernelGGL(k_max_colour, GRID(l), word, n, ctl);
}
void launch_unset_pivots(const Launch &l, uint32_t *word, int32_t n, Ctl *ctl, const int32_t *piv) {
    hipLaunchKernelGGL(k_unset_pivots, GRID(l), word, n, ctl, piv);
}
void launch_trim2(const Launch &l, const Graph &g, uint32_t *word, int32_t *root, Ctl *ctl, const int32_t *oid) {
    hipLaunchKernelGGL(k_trim2_mark, GRID(l), g.row, g.col, g.rowT, g.colT, word, root, g.n, oid);
    hipLaunchKernelGGL(k_trim2_commit, GRID(l), word, g.n, ctl);
}
void launch_wcc_init(const Launch &l, int32_t *lab, int32_t n, const int32_t *lin, const uint32_t *lin_cnt,
                     const int32_t *oid) {
    hipLaunchKernelGGL(k_wcc_init, GRID(l), lab, n, lin, lin_cnt, oid);
}
void launch_wcc_iter(const Launch &l, const Graph &g, const uint32_t *word, int32_t *lab, Ctl *ctl,
                     int iter, const int32_t *lin, const uint32_t *lin_cnt, const int32_t *nid) {
    hipLaunchKernelGGL(k_wcc_iter, GRID(l), g.row, g.col, word, lab, g.n, ctl, iter, lin, lin_cnt, nid);
}
void launch_wcc_recolour(const Launch &l, uint32_t *word, const int32_t *lab, int32_t n,
                         const Ctl *ctl, const int32_t *lin, const uint32_t *lin_cnt) {
    hipLaunchKernelGGL(k_wcc_recolour, GRID(l), word, lab, n, ctl, lin, lin_cnt);
}
void launch_root_to_min(const Launch &l, int32_t *root, const int32_t *mn, const uint32_t *word, int32_t n) {
    hipLaunchKernelGGL(k_root_to_min, GRID(l), root, mn, word, n);
}
void launch_gather_i32(const Launch &l, const int32_t *in, const int32_t *idx, int32_t *out, int64_t count) {
    hipLaunchKernelGGL(k_gather_i32, GRID(l), in, idx, out, count);
}
void launch_permute_rows(const Launch &l, const int32_t *row, const int32_t *col, const int32_t *map,
                         const int32_t *nrow, int32_t *ncol, int32_t n) {
    hipLaunchKernelGGL(k_permute_rows, GRID(l), row, col, map, nrow, ncol, n);
}
void launch_collect(const Launch &l, const uint32_t *word, int32_t n, uint32_t bit, int32_t *q,
                    Ctl *ctl) {
    hipLaunchKernelGGL(k_collect, GRID(l), word, n, bit, q, ctl);
}

hipError_t preload_fb() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(k_fill_i32));
}

}  // namespace gpuscc
