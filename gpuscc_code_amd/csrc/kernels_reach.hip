// Forward / backward reachability inside colours for gfx950: one kernel per
// BFS level, chained on the device (a level whose predecessor found nothing
// exits at once), in the four schedules of the reference plus its default:
//   q  sweep, one lane per vertex                 (reach_q.cu:12-79)
//   t  sweep, workgroup-balanced edge expansion   (reach_t.cu:17-270)
//   d  frontier queue, one lane per item          (reach_d.cu:11-64)
//   l  frontier queue, workgroup-balanced edges   (reach_l.cu:53-253)
//   h  per level: queue expansion while the frontier is small, balanced sweep
//      once it is large (scc.cu:493-563 switches by eliminated count instead)
//
// Every visit is claimed with one atomicOr on the target's word, whose old
// value says who set the bit first, so a vertex enters a queue exactly once:
// queues of n entries cannot overflow (the reference's can, worklistc.h:80).
// Pushes go through an LDS staging buffer, one global atomicAdd per flush.
#include "fb_device.hpp"
#include "kernels.hpp"

namespace gpuscc {

constexpr uint32_t SQ_CAP = 2048;  // LDS staging entries per workgroup (8 KiB)

__device__ inline bool try_claim(uint32_t *word, int32_t u, uint32_t c, uint32_t bit) {
    uint32_t wu = word[u];
    if ((wu & (bit | ST_E)) != 0 || colour_of(wu) != c) return false;
    uint32_t old = atomicOr(&word[u], bit);
    return !(old & bit);
}

__device__ inline void stage_push(int32_t *sbuf, uint32_t *scount, int32_t v, int32_t *outq,
                                  Ctl *ctl, int iter) {
    uint32_t i = atomicAdd(scount, 1u);
    if (i < SQ_CAP) {
        sbuf[i] = v;
    } else {
        uint32_t g = atomicAdd(&ctl->qsize[(iter + 1) & 3], 1u);
        outq[g] = v;
    }
}

// all threads of the workgroup call this
__device__ inline void stage_flush(int32_t *sbuf, uint32_t *scount, uint32_t *sbase, int32_t *outq,
                                   Ctl *ctl, int iter) {
    __syncthreads();
    uint32_t c = *scount;
    if (c > SQ_CAP) c = SQ_CAP;
    if (c > 0) {
        if (threadIdx.x == 0) *sbase = atomicAdd(&ctl->qsize[(iter + 1) & 3], c);
        __syncthreads();
        uint32_t b = *sbase;
        for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) outq[b + i] = sbuf[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) *scount = 0;
    __syncthreads();
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
// q: one lane per vertex; a live vertex with the bit and without the expanded
// bit sets the expanded bit and scans its row (reach_q.cu:44-79)
// ---------------------------------------------------------------------------
__global__ void k_reach_sweep(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                              uint32_t *word, int32_t n, uint32_t bit, uint32_t xbit, Ctl *ctl,
                              int iter, int instr) {
    __shared__ unsigned long long sh[16];
    if (chain_exit(ctl, iter)) return;
    int64_t beg, end;
    my_range(n, beg, end);
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    for (int64_t vv = beg + threadIdx.x; vv < end; vv += blockDim.x) {
        int32_t v = (int32_t)vv;
        uint32_t w = word[v];
        if ((w & (bit | xbit | ST_E)) != bit) continue;
        word[v] = w | xbit;
        uint32_t c = colour_of(w);
        int32_t e0 = row[v], e1 = row[v + 1];
        cx++;
        ce += (unsigned long long)(e1 - e0);
        for (int32_t e = e0; e < e1; e++) {
            if (try_claim(word, col[e], c, bit)) {
                found = true;
                cv++;
            }
        }
    }
    chain_progress(ctl, iter, found);
    if (instr) instr_flush(ctl, cx, ce, cv, sh);
}

// ---------------------------------------------------------------------------
// balanced expansion of up to blockDim items held in LDS: item k owns edges
// [start[k], start[k]+deg[k]); work-item t takes edges t, t+blockDim, ... of
// the concatenation (the CTA gather of reach_t.cu:195-265, without its
// 256-edge windows)
// ---------------------------------------------------------------------------
struct Items {
    uint32_t off[1024];
    int32_t start[1024];
    uint32_t colour[1024];
    uint32_t shw[16];
};

template <bool PUSH>
__device__ inline void expand_items(const int32_t *__restrict__ col, uint32_t *word, uint32_t bit,
                                    Items &it, uint32_t total, int32_t *sbuf, uint32_t *scount,
                                    uint32_t *sbase, int32_t *outq, Ctl *ctl, int iter, bool &found,
                                    unsigned long long &cv) {
    uint32_t cnt = blockDim.x;
    // block-uniform trip count: every pass adds at most blockDim entries to the
    // staging buffer, which is flushed before it can overflow
    for (uint32_t base = 0; base < total; base += blockDim.x) {
        uint32_t e = base + threadIdx.x;
        if (e < total) {
            uint32_t k = find_item(it.off, cnt, e);
            int32_t u = col[it.start[k] + (int32_t)(e - it.off[k])];
            if (try_claim(word, u, it.colour[k], bit)) {
                found = true;
                cv++;
                if (PUSH) stage_push(sbuf, scount, u, outq, ctl, iter);
            }
        }
        if (PUSH) {
            __syncthreads();
            bool full = *scount > SQ_CAP - blockDim.x;  // read by all before anyone pushes again
            __syncthreads();
            if (full) stage_flush(sbuf, scount, sbase, outq, ctl, iter);
        }
    }
}

// t: balanced sweep.  PUSH=true (hybrid sweep levels) also queues every new
// visit so a later queue level can continue from it.
template <bool PUSH>
__device__ inline void topo_level(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                  uint32_t *word, int32_t n, uint32_t bit, uint32_t xbit, Ctl *ctl,
                                  int iter, Items &it, int32_t *sbuf, uint32_t *scount,
                                  uint32_t *sbase, int32_t *outq, bool &found,
                                  unsigned long long &cx, unsigned long long &ce,
                                  unsigned long long &cv) {
    int64_t beg, end;
    my_range(n, beg, end);
    for (int64_t base = beg; base < end; base += blockDim.x) {
        int64_t vv = base + threadIdx.x;
        uint32_t deg = 0, c = 0;
        int32_t start = 0;
        if (vv < end) {
            int32_t v = (int32_t)vv;
            uint32_t w = word[v];
            if ((w & (bit | xbit | ST_E)) == bit) {
                word[v] = w | xbit;
                start = row[v];
                deg = (uint32_t)(row[v + 1] - start);
                c = colour_of(w);
                cx++;
                ce += deg;
            }
        }
        uint32_t total;
        uint32_t off = block_excl_scan(deg, it.shw, total);
        it.off[threadIdx.x] = off;
        it.start[threadIdx.x] = start;
        it.colour[threadIdx.x] = c;
        __syncthreads();
        if (total > 0)
            expand_items<PUSH>(col, word, bit, it, total, sbuf, scount, sbase, outq, ctl, iter, found,
                               cv);
        if (PUSH) stage_flush(sbuf, scount, sbase, outq, ctl, iter);
        else __syncthreads();
    }
}

__global__ void k_reach_topo(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                             uint32_t *word, int32_t n, uint32_t bit, uint32_t xbit, Ctl *ctl,
                             int iter, int instr) {
    __shared__ Items it;
    __shared__ unsigned long long sh[16];
    if (chain_exit(ctl, iter)) return;
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    topo_level<false>(row, col, word, n, bit, xbit, ctl, iter, it, nullptr, nullptr, nullptr,
                      nullptr, found, cx, ce, cv);
    chain_progress(ctl, iter, found);
    if (instr) instr_flush(ctl, cx, ce, cv, sh);
}

// ---------------------------------------------------------------------------
// d: frontier queue, one lane per item, serial row scan (reach_d.cu:36-64)
// ---------------------------------------------------------------------------
__global__ void k_reach_queue(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                              uint32_t *word, uint32_t bit, const int32_t *__restrict__ inq,
                              int32_t *outq, Ctl *ctl, int iter, int instr) {
    __shared__ int32_t sbuf[SQ_CAP];
    __shared__ uint32_t scount, sbase;
    __shared__ unsigned long long sh[16];
    if (chain_exit(ctl, iter)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->qsize[(iter + 2) & 3] = 0;
    if (threadIdx.x == 0) scount = 0;
    __syncthreads();
    uint32_t qs = ctl->qsize[iter & 3];
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < qs; base += stride) {
        uint32_t i = base + threadIdx.x;
        if (i < qs) {
            int32_t v = inq[i];
            uint32_t c = colour_of(word[v]);
            int32_t e0 = row[v], e1 = row[v + 1];
            cx++;
            ce += (unsigned long long)(e1 - e0);
            for (int32_t e = e0; e < e1; e++) {
                int32_t u = col[e];
                if (try_claim(word, u, c, bit)) {
                    found = true;
                    cv++;
                    stage_push(sbuf, &scount, u, outq, ctl, iter);
                }
            }
        }
        stage_flush(sbuf, &scount, &sbase, outq, ctl, iter);
    }
    chain_progress(ctl, iter, found);
    if (instr) instr_flush(ctl, cx, ce, cv, sh);
}

// ---------------------------------------------------------------------------
// l: frontier queue with balanced edge expansion (reach_l.cu:151-253).
// XBIT != 0 also marks each item expanded (hybrid queue levels).
// ---------------------------------------------------------------------------
__device__ inline void lb_level(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                uint32_t *word, uint32_t bit, uint32_t xbit,
                                const int32_t *__restrict__ inq, int32_t *outq, Ctl *ctl, int iter,
                                uint32_t qs, Items &it, int32_t *sbuf, uint32_t *scount,
                                uint32_t *sbase, bool &found, unsigned long long &cx,
                                unsigned long long &ce, unsigned long long &cv) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < qs; base += stride) {
        uint32_t i = base + threadIdx.x;
        uint32_t deg = 0, c = 0;
        int32_t start = 0;
        if (i < qs) {
            int32_t v = inq[i];
            uint32_t w = word[v];
            if (xbit) word[v] = w | xbit;
            start = row[v];
            deg = (uint32_t)(row[v + 1] - start);
            c = colour_of(w);
            cx++;
            ce += deg;
        }
        uint32_t total;
        uint32_t off = block_excl_scan(deg, it.shw, total);
        it.off[threadIdx.x] = off;
        it.start[threadIdx.x] = start;
        it.colour[threadIdx.x] = c;
        __syncthreads();
        if (total > 0)
            expand_items<true>(col, word, bit, it, total, sbuf, scount, sbase, outq, ctl, iter, found,
                               cv);
        stage_flush(sbuf, scount, sbase, outq, ctl, iter);
    }
}

__global__ void k_reach_lb(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                           uint32_t *word, uint32_t bit, const int32_t *__restrict__ inq,
                           int32_t *outq, Ctl *ctl, int iter, int instr) {
    __shared__ Items it;
    __shared__ int32_t sbuf[SQ_CAP];
    __shared__ uint32_t scount, sbase;
    __shared__ unsigned long long sh[16];
    if (chain_exit(ctl, iter)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->qsize[(iter + 2) & 3] = 0;
    if (threadIdx.x == 0) scount = 0;
    __syncthreads();
    uint32_t qs = ctl->qsize[iter & 3];
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    lb_level(row, col, word, bit, 0u, inq, outq, ctl, iter, qs, it, sbuf, &scount, &sbase, found,
             cx, ce, cv);
    chain_progress(ctl, iter, found);
    if (instr) instr_flush(ctl, cx, ce, cv, sh);
}

// ---------------------------------------------------------------------------
// h: per level, the frontier size in the queue ring picks the expansion: below
// `threshold` items the level expands the queue (balanced); at or above it the
// level sweeps all vertices with the bit and without the expanded bit (the
// same set), which reads the vertex words once, coalesced, instead of
// gathering them through the queue.  Both forms mark expanded vertices and
// queue new visits, so consecutive levels may use either.
// ---------------------------------------------------------------------------
__global__ void k_reach_hybrid(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                               uint32_t *word, int32_t n, uint32_t bit, uint32_t xbit,
                               const int32_t *__restrict__ inq, int32_t *outq, Ctl *ctl, int iter,
                               uint32_t threshold, int instr) {
    __shared__ Items it;
    __shared__ int32_t sbuf[SQ_CAP];
    __shared__ uint32_t scount, sbase;
    __shared__ unsigned long long sh[16];
    if (chain_exit(ctl, iter)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->qsize[(iter + 2) & 3] = 0;
    if (threadIdx.x == 0) scount = 0;
    __syncthreads();
    uint32_t qs = ctl->qsize[iter & 3];
    bool found = false;
    unsigned long long cx = 0, ce = 0, cv = 0;
    if (qs >= threshold) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->mode_sweeps++;
        topo_level<true>(row, col, word, n, bit, xbit, ctl, iter, it, sbuf, &scount, &sbase, outq,
                         found, cx, ce, cv);
    } else {
        lb_level(row, col, word, bit, xbit, inq, outq, ctl, iter, qs, it, sbuf, &scount, &sbase,
                 found, cx, ce, cv);
    }
    chain_progress(ctl, iter, found);
    if (instr) instr_flush(ctl, cx, ce, cv, sh);
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
#define GRID(l) dim3((l).grid), dim3((l).block), 0, (l).stream
static inline uint32_t bit_of(int dir) { return dir == 0 ? ST_F : ST_B; }
static inline uint32_t xbit_of(int dir) { return dir == 0 ? ST_FX : ST_BX; }

void launch_reach_sweep(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                        int32_t n, int dir, Ctl *ctl, int iter, int instr) {
    hipLaunchKernelGGL(k_reach_sweep, GRID(l), row, col, word, n, bit_of(dir), xbit_of(dir), ctl,
                       iter, instr);
}
void launch_reach_topo(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                       int32_t n, int dir, Ctl *ctl, int iter, int instr) {
    hipLaunchKernelGGL(k_reach_topo, GRID(l), row, col, word, n, bit_of(dir), xbit_of(dir), ctl,
                       iter, instr);
}
void launch_reach_queue(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                        const int32_t *inq, int32_t *outq, int dir, Ctl *ctl, int iter, int instr) {
    hipLaunchKernelGGL(k_reach_queue, GRID(l), row, col, word, bit_of(dir), inq, outq, ctl, iter,
                       instr);
}
void launch_reach_lb(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                     const int32_t *inq, int32_t *outq, int dir, Ctl *ctl, int iter, int instr) {
    hipLaunchKernelGGL(k_reach_lb, GRID(l), row, col, word, bit_of(dir), inq, outq, ctl, iter,
                       instr);
}
void launch_reach_hybrid(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                         int32_t n, const int32_t *inq, int32_t *outq, int dir, Ctl *ctl, int iter,
                         uint32_t threshold, int instr) {
    hipLaunchKernelGGL(k_reach_hybrid, GRID(l), row, col, word, n, bit_of(dir), xbit_of(dir), inq,
                       outq, ctl, iter, threshold, instr);
}
void launch_clear_bits(const Launch &l, uint32_t *word, int32_t n, uint32_t bits) {
    hipLaunchKernelGGL(k_clear_bits, GRID(l), word, n, bits);
}

}  // namespace gpuscc
