// Device-side definitions shared by the FB kernels (gfx950, wave64).
//
// Vertex word (one u32 per vertex in HBM): bits 0..7 hold the reference's
// state byte (scc.h:48-93: B=1 F=2 E=4 T=8 P=16 BExt=32 FExt=64; bit 128 marks
// a trim-2 pair awaiting commit), bits 8..31 the colour.  Packing colour and
// state into one word makes the per-edge test of reachability -- "live, not
// yet visited, same colour" (reach_q.cu:32-33,69-70) -- a single 4-byte gather
// instead of the reference's two (states[] u8 + ranges[] u32).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>

namespace gpuscc {

constexpr uint32_t ST_B = 1, ST_F = 2, ST_E = 4, ST_T = 8, ST_P = 16, ST_BX = 32, ST_FX = 64;
constexpr uint32_t ST_PAIR = 128;
constexpr uint32_t COLOUR_MASK = 0x00FFFFFFu;
constexpr int WAVE = 64;

__host__ __device__ inline uint32_t colour_of(uint32_t w) { return w >> 8; }
__host__ __device__ inline uint32_t make_word(uint32_t c, uint32_t s) { return (c << 8) | s; }

// Election slot of a colour: top slot_bits bits of colour * 0x9E3779B1 (mod 2^32).
__host__ __device__ inline uint32_t slot_of(uint32_t c, int slot_bits) {
    return slot_bits == 0 ? 0u : (uint32_t)(c * 0x9E3779B1u) >> (32 - slot_bits);
}

// Election key, smaller wins: [63:48] inverted epoch tag, [47:32] inverted
// score (0 for the plain rule), [31:0] vertex id.  A newer epoch has a smaller
// inverted tag, so stale slots never need clearing between elections.
__host__ __device__ inline uint64_t elect_tag(uint32_t epoch) {
    return (uint64_t)(0xFFFFu - (epoch & 0xFFFFu)) << 48;
}

// Vertex ids.  An upload may renumber the vertices by decreasing degree (hubs
// first, so the words, visited bits and rows the scans touch most share cache
// lines); oid[v] is then the caller's id of internal vertex v, and nid the
// inverse.  Everything that orders vertices by id -- election keys, roots,
// trim-2's lower id, WCC labels -- uses the caller's ids, so every result is
// the one of the graph as given.  Null: no renumbering.
__device__ inline int32_t orig_id(const int32_t *oid, int32_t v) { return oid ? oid[v] : v; }

// Device control block: level-chain flags, queue sizes and reduction scratch.
struct Ctl {
    uint32_t changed[4];   // chain ring: kernel i sets [i&3] on progress, clears [(i+1)&3]
    uint32_t qsize[2][4];  // queue chain rings (forward, backward): input of level i is [d][i&3]
    uint32_t hsize[2][4];  // rings of the huge-row lists, same indexing
    uint32_t seeds[2];     // seed-list lengths of the next reach (forward, backward)
    uint32_t ran;          // chain kernels that did work (depth / iterations)
    uint32_t npiv;         // pivots elected by the last election = pivot list length
    uint32_t maxc;         // max colour (reduction scratch)
    uint32_t pull_levels;  // hybrid: direction-levels that ran bottom-up (pull)
    uint32_t pull_mask;    // hybrid: directions the last level pulled (bit 0 F, bit 1 B)
    uint32_t live;         // vertices not eliminated (set by k_init; trims, pair trims and
                           // FB eliminations subtract)
    uint32_t first_piv;    // lowest pivot of the last election
    uint32_t npiv_prev, first_piv_prev;  // npiv, first_piv of the election before the last
    uint32_t pcount;       // vertices holding P (class representatives: #SCC at the end)
    uint32_t fbelim;       // vertices FB-eliminated by updates (countElimmed before the WCC pass)
    uint32_t abort;        // speculative step: chain stage that did not converge (0 = none);
                           // every gated kernel exits at once while it is set
    uint32_t depth[3];     // speculative step: levels of its reach chain and trim chain; of trim-1's chain
    unsigned long long count;     // count reductions (FB-eliminated, P bits, pairs)
    unsigned long long qarcs[2][4];  // estimated arcs of the queued vertices, same rings as qsize
    uint32_t creach[2][4];  // vertices reached so far by each direction (ring: level iter writes [iter & 3])
    unsigned long long carcs[2][4];  // direction-optimising policy: estimated arcs of the frontiers each
                                     // direction has pushed so far (same ring)
    uint32_t pmask[4];      // directions each level pulled (bit 0 F, bit 1 B; ring: level iter writes [iter & 3])
    unsigned long long st_expanded, st_edges, st_visited;  // instrumented counters
    uint32_t st_maxscan[4];  // instrumented: longest serial row scan (arcs) of one work-item,
                             // ring: level iter writes [iter & 3]
    uint32_t lcnt[2];        // lengths of the two live-vertex lists (an update reads one, writes the other)
    uint32_t tail[4];        // k_fb_tail: steps run, status (0 done, 1 host continues, 2 error),
                             // reach levels and trim rounds
    unsigned long long tail_pivots;  // k_fb_tail: pivots elected
    uint32_t tail_live0;             // k_fb_tail: live vertices when it started
    uint32_t wnext;                  // k_fb_batch: the next graph to take (work counter)
    uint32_t terr;                   // tile reach: a tile of more than TILE_MAX vertices (never: the
                                     // upload cuts them smaller)
    unsigned long long st_truns, st_trounds;  // instrumented tile reach: tile runs, their rounds
    unsigned long long st_tclk[4];            // instrumented tile reach: 100 MHz ticks of the runs' phases
                                              // (set-up and seeds, fixpoint, counting, claims across)
    unsigned long long st_tbusy, st_tidle;  // instrumented tile reach: 100 MHz ticks waves spent on takes,
                                            // idle polls
    unsigned long long st_tailclk[4];  // instrumented tail: 100 MHz ticks of its steps' reach, trim, update,
                                       // confirm (thread 0's view)
    uint32_t tail_L, tail_lstep;       // instrumented tail: live vertices copied to LDS, the step it happened
    uint32_t tlog[4][3];  // instrumented tile reach, ring by iteration: tiles run, most rounds and
                          // longest run (100 MHz ticks) of one tile
};

// Persistent-grid helpers.  Workgroups are dealt to the 8 XCDs round-robin;
// remapping the block index gives each XCD one contiguous share of the vertex
// range, so neighbouring ids (mesh rows, CSR locality) share that XCD's L2.
__device__ inline uint32_t xcd_block() {
    uint32_t b = blockIdx.x, g = gridDim.x;
    if ((g & 7u) != 0) return b;
    return (b & 7u) * (g >> 3) + (b >> 3);
}

// contiguous share of [0, n) for the calling workgroup (XCD-aware block id)
__device__ inline void my_range(int64_t n, int64_t &beg, int64_t &end) {
    uint32_t g = gridDim.x, lb = xcd_block();
    int64_t chunk = ((n + g - 1) / g + 63) & ~(int64_t)63;
    beg = (int64_t)lb * chunk;
    end = beg + chunk < n ? beg + chunk : n;
}

// Gate of a chained level kernel: block 0 clears the flag of the next level;
// a level whose predecessor made no progress exits at once.
__device__ inline bool chain_exit(Ctl *ctl, int iter) {
    if (__atomic_load_n(&ctl->abort, __ATOMIC_RELAXED)) return true;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->changed[(iter + 1) & 3] = 0;
    if (iter > 0 && __atomic_load_n(&ctl->changed[(iter - 1) & 3], __ATOMIC_RELAXED) == 0)
        return true;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->ran = (uint32_t)(iter + 1);
    return false;
}

__device__ inline void chain_progress(Ctl *ctl, int iter, bool found) {
    if (__any(found) && (threadIdx.x & (WAVE - 1)) == 0) ctl->changed[iter & 3] = 1;
}

// exclusive prefix sum over the workgroup (nt = its size <= 1024, wave64).  Kernels
// in loops pass nt from their arguments: blockDim.x is a vector load from the
// dispatch packet, which the compiler may repeat in every iteration.
__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t *shw, uint32_t &total,
                                           uint32_t nt = blockDim.x) {
    int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    uint32_t inc = x;
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, WAVE);
        if (lane >= o) inc += y;
    }
    __syncthreads();
    if (lane == WAVE - 1) shw[wid] = inc;
    __syncthreads();
    int nw = (int)((nt + WAVE - 1) / WAVE);
    uint32_t pre = 0, tot = 0;
    for (int i = 0; i < nw; i++) {
        uint32_t s = shw[i];
        if (i < wid) pre += s;
        tot += s;
    }
    total = tot;
    return pre + inc - x;
}

// Block-wide sum / max through LDS (nt = workgroup size <= 1024).
template <typename T>
__device__ inline T block_sum(T v, T *sh, uint32_t nt = blockDim.x) {
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
    int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(nt + WAVE - 1) / WAVE; i++) r += sh[i];
    return r;  // valid in thread 0
}

template <typename T>
__device__ inline T block_max(T v, T *sh) {
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        T x = __shfl_down(v, o, WAVE);
        v = x > v ? x : v;
    }
    int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x + WAVE - 1) / WAVE; i++) r = sh[i] > r ? sh[i] : r;
    return r;
}

}  // namespace gpuscc
