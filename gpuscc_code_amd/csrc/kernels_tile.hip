// Reachability over tiles (first FB step of a regular graph).  On a graph of
// high diameter the reach from the first pivot takes thousands of BFS levels
// (the 2048 x 2048 mesh: about 4,100), and a level costs a global memory round
// trip in the level kernels (a workgroup barrier in their local rounds).  Here
// the upload numbers the vertices tile by tile: a tile is a region of up to
// 4,096 vertices grown by BFS, made of cells of 64 vertices, themselves grown
// by BFS inside the tile, so that most arcs stay inside a cell.  Per vertex and
// direction the upload keeps its neighbours in its own cell as a 64-bit mask
// (bit = lane), up to four neighbours in other cells of its tile as 16-bit
// indices, and flags for arcs that leave the tile.
//
// One iteration of the chain gives every tile of its list one 1024-wide
// workgroup; wave w owns cells w, w + 16, w + 32, w + 48, one vertex per lane,
// with their neighbour masks in registers.  Inside the tile the waves work on
// their own, without barriers: a wave takes the claim words of its cells
// (LDS, atomic exchange), closes the claimed vertices inside their cells -- one
// ballot: each lane tests the cell closure it holds (the cell's vertices it is
// reached from, computed for every cell by iteration 0) against the claims --
// and claims the new vertices' neighbours in other cells of the tile with LDS
// atomics.  A count of claim
// words not yet done says when the tile is at its fixpoint.  Then the tile
// writes its new bits back and claims the targets of its new vertices' arcs
// that leave the tile (global atomics, as the level kernels claim), putting
// them in their tiles' inboxes and those tiles in the next list.  So a cell
// is crossed in a few vector instructions, a step between cells costs an LDS
// round trip, and a step between tiles one iteration.  The reach is the
// same closure, so the F / B bits are the level kernels' (the reach of
// scc.cu:285-704 over reach_q.cu:12-79's rule: live, same colour, not yet
// reached; in a first FB step every live vertex has one colour, so the masks
// need no colour test).
#include "fb_device.hpp"
#include "kernels.hpp"

namespace gpuscc {

constexpr int TILE_BLOCK = 1024;
constexpr int TILE_WAVES = TILE_BLOCK / WAVE;                          // 16
constexpr int TILE_CELLS = (int)(TILE_MAX / CELL);                     // 64
constexpr int CELLS_PER_WAVE = TILE_CELLS / TILE_WAVES;                // 4
constexpr int TILE_LIST_MAX = TILE_BLOCK / 64;  // a workgroup's tiles of one chunk of flags (grid >= 64)
static_assert(CELL == WAVE, "a cell is one vertex per lane of a wave");

// claims into cell k, direction d: [cidx(k, d)], bit = lane; a wave's own entries are
// contiguous (cell k belongs to wave k % 16, as its entry k / 16)
__device__ inline uint32_t cidx(uint32_t k, int d) {
    return ((k % TILE_WAVES) * CELLS_PER_WAVE + k / TILE_WAVES) * 2u + (uint32_t)d;
}

struct TileShared {
    unsigned long long claim[TILE_CELLS * 2];  // [cidx]: claims not yet taken by the cell's wave
    unsigned long long xa[2 * TILE_MAX];       // the tile's arcs to other cells ([direction][vertex])
    uint32_t exits[2 * TILE_MAX];  // (vertex, direction) reached in this run with arcs leaving the tile
    // claim words that are non-zero or taken and not yet processed: a depositor counts a
    // word before it can turn non-zero (and takes the count back when it was non-zero
    // already), a wave uncounts a word it took once it has made all its own claims, so the
    // count never falls below the work left and zero means the tile is at its fixpoint
    uint32_t pending;
    unsigned long long sink[WAVE];  // a lane's claim slot without a target ORs nothing into its own word
    uint32_t n_exit;
    uint32_t stuck;    // a wave gave up waiting (never: the count reaches zero)
    unsigned long long ins[6];  // instrumented runs: the waves' busy ticks, idle polls, takes, most takes of
                                // a wave, vertices reached, their arcs (flushed once per run)
};

// One workgroup works on a tile: a workgroup-scope barrier orders its LDS and
// global accesses (as in the tail, agent-scope fences would write back the L2)
__device__ inline void tile_sync() { __syncthreads(); }

// a value every lane holds alike, kept in scalar registers
__device__ inline unsigned long long uni64(unsigned long long x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// lane l's 64-bit value (l uniform)
__device__ inline unsigned long long lane64(unsigned long long x, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}

// a claim of u across a tile boundary in direction d, and u's tile flagged for
// the next iteration (every tile but the last holds tsize vertices).  The claim
// sets the bit without reading the word: that tile's next run takes the live
// vertices whose bits are set and which it has not expanded, so a bit set twice,
// or on an eliminated vertex, changes nothing (the F and B bits of eliminated
// vertices are read by nothing).
__device__ inline void claim_across(const TileArgs &a, int d, int32_t u) {
    atomicOr(&a.word[u], d ? ST_B : ST_F);
    a.flags[(size_t)((a.iter + 1) % 3) * a.ntiles + u / a.tsize] = 1;
}

// orders this wave's LDS operations only (a workgroup fence on all memory would also
// wait for its global write-backs)
__device__ inline void lds_fence() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local"); }

// claim bit of word idx for another cell, counted in pending first
__device__ inline void deposit(TileShared &sh, uint32_t idx, unsigned long long bit) {
    atomicAdd(&sh.pending, 1u);
    lds_fence();  // counted before the word can turn non-zero
    if (atomicOr(&sh.claim[idx], bit) != 0ull) atomicSub(&sh.pending, 1u);
}

__device__ inline unsigned long long lds_load(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// claims into the other cells of the tile along the whole row of tile vertex i
// (vertices with more such arcs than the upload keeps)
__device__ __noinline__ void cross_row(TileShared &sh, const int32_t *row, const int32_t *col, int32_t lo,
                                       uint32_t T, uint32_t i, int d) {
    const int32_t v = lo + (int32_t)i;
    for (int32_t e = row[v]; e < row[v + 1]; e++) {
        const uint32_t u = (uint32_t)(col[e] - lo);
        if (u < T && u / CELL != i / CELL) deposit(sh, cidx(u / CELL, d), 1ull << (u % CELL));
    }
}

constexpr uint32_t TILE_SPIN_LIMIT = 1u << 20;  // idle polls before a wave gives up (never reached)

// The closure of every cell of tile t inside the cell (Warshall's algorithm on the
// cell's 64 x 64 bit matrix, one row per lane): ct[0][u] = the vertices of u's cell
// from which u is reached inside the cell, ct[1][u] = those u reaches, u included.
// The masks ignore which vertices are live: in the first FB step only trim-1 at its
// fixpoint has eliminated vertices, and no path from a live pivot to a live vertex
// passes through a trimmed one (the first trimmed vertex on it would have kept a
// live predecessor, so it was trimmed for lacking live successors, after its
// successor on the path, and so on to the end), so the live vertices of a closure
// from reached vertices are exactly those reached through live vertices.
__device__ void cell_closure(const TileArgs &a, int32_t t) {
    const uint32_t lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const int32_t lo = a.toff[t];
    const uint32_t T = (uint32_t)(a.toff[t + 1] - lo);
    for (uint32_t k = wv; k * CELL < T; k += TILE_WAVES) {
        const uint32_t i = k * CELL + lane;
        const bool ok = i < T;
        const int32_t v = lo + (int32_t)(ok ? i : 0);
#pragma unroll
        for (int d = 0; d < 2; d++) {
            // forward: u is reached from its in-neighbours; backward: from its out-neighbours
            unsigned long long m = ok ? (a.adj[d ^ 1][v] | (1ull << lane)) : 0ull;
#pragma unroll
            for (int q = 0; q < (int)CELL; q++) {
                const unsigned long long mq = lane64(m, q);
                if ((m >> q) & 1ull) m |= mq;
            }
            if (ok) a.ct[d][v] = m;
        }
    }
}

// one tile's run to its fixpoint (block-uniform control flow throughout)
__device__ void tile_run(const TileArgs &a, TileShared &sh, int32_t t, bool &found) {
    const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int32_t lo = a.toff[t], hi = a.toff[t + 1];
    const uint32_t T = (uint32_t)(hi - lo);
    if (T > TILE_MAX) {  // never: the upload cuts tiles of at most TILE_MAX vertices
        if (tid == 0) atomicAdd(&a.ctl->terr, 1u);
        return;
    }
    unsigned long long clk[5];  // phase boundaries (instrumented runs, thread 0)
    if (a.instr) clk[0] = wall_clock64();
    for (uint32_t i = tid; i < TILE_CELLS * 2u; i += TILE_BLOCK) sh.claim[i] = 0ull;
    for (uint32_t i = tid; i < T; i += TILE_BLOCK) {
        sh.xa[i] = a.xarc[0][lo + (int32_t)i];
        sh.xa[TILE_MAX + i] = a.xarc[1][lo + (int32_t)i];
    }
    if (tid == 0) {
        sh.n_exit = 0;
        sh.pending = 0;
        sh.stuck = 0;
    }
    if (tid < WAVE) sh.sink[tid] = 0ull;
    if (tid < 6) sh.ins[tid] = 0ull;
    const uint32_t base = wv * CELLS_PER_WAVE * 2;  // this wave's claim words
    unsigned long long *done = a.done + (size_t)t * (TILE_CELLS * 2);
    // lane q < 8 of a wave keeps the masks of its pair q = 2j + d (its cell j, direction d):
    // the vertices the tile has expanded so far
    unsigned long long reach_l = lane < 8 ? done[base + lane] : 0ull;
    const unsigned long long done0 = reach_l;
    // this wave's cells, one vertex per lane: word, closure masks in the cell, flags
    uint32_t w[CELLS_PER_WAVE], xfl = 0;
    unsigned long long ct[CELLS_PER_WAVE][2];
#pragma unroll
    for (int j = 0; j < CELLS_PER_WAVE; j++) {
        const uint32_t i = (wv + TILE_WAVES * j) * CELL + lane;
        const bool ok = i < T;
        const int32_t v = lo + (int32_t)(ok ? i : 0);
        w[j] = ok ? a.word[v] : ST_E;  // lanes past the tile's end: never live
        xfl |= (ok ? (uint32_t)a.xflag[v] : 0u) << (8 * j);
        ct[j][0] = ok ? a.ct[0][v] : 0ull;
        ct[j][1] = ok ? a.ct[1][v] : 0ull;
    }
    // the seeds: vertices whose words are reached but which the tile has not expanded
    // (claimed across its boundary since its last run, the pivots), live ones only; each
    // wave puts them in its own claim words, counted
    if (a.instr) clk[1] = wall_clock64();
    unsigned long long seed_l = 0ull;
#pragma unroll
    for (int j = 0; j < CELLS_PER_WAVE; j++) {
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const unsigned long long wb = __ballot((w[j] & (d ? ST_B : ST_F)) && !(w[j] & ST_E));
            if (lane == (uint32_t)(j * 2 + d)) seed_l = wb & ~reach_l;
        }
    }
    tile_sync();  // the claim words and the count are cleared
    if (lane < 8 && seed_l) sh.claim[base + lane] = seed_l;
    const uint32_t nseed = (uint32_t)__popcll(__ballot(seed_l != 0ull));
    if (lane == 0 && nseed) atomicAdd(&sh.pending, nseed);
    const unsigned long long below = (1ull << lane) - 1ull;
    tile_sync();  // every wave has counted its seeds before claims inside the tile begin
    // to the fixpoint, each wave on its own: take every claim word of its cells at once, close
    // the claims inside their cells, write the new bits back, list the new vertices whose arcs
    // leave the tile, claim into other cells (counted first), uncount the words taken; stop
    // when nothing is counted
    uint32_t polls = 0, idles = 0, n_reached = 0;
    uint32_t newbits = 0;  // bits reached in this run, 8 per cell j: written back at the end for vertices no
                           // other tile can claim (no arc crosses the tile boundary at them)
    unsigned long long busy = 0, t_take = 0;
    for (uint32_t idle = 0;;) {
        unsigned long long pl = lane < 8 ? lds_load(&sh.claim[base + lane]) : 0ull;
        const unsigned long long m = uni64(__ballot(pl != 0ull));  // pairs with claims
        if (m == 0ull) {
            const uint32_t left = __hip_atomic_load(&sh.pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (__builtin_amdgcn_readfirstlane((int)left) == 0) break;  // the tile's fixpoint
            if (++idle > TILE_SPIN_LIMIT) {  // never: the count reaches zero once the work is done
                if (lane == 0) sh.stuck = 1;
                break;
            }
            idles++;
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        idle = 0;
        polls++;
        if (a.instr) t_take = wall_clock64();
        // only this wave clears its words: an exchange returns at least what was seen
        if ((m >> lane) & 1ull) pl = atomicExch(&sh.claim[base + lane], 0ull);
        unsigned long long fr_l = 0ull;  // lane q: the vertices pair q reached in this take
        uint32_t bound = 0;              // claims the take may make into other cells
#pragma unroll
        for (int j = 0; j < CELLS_PER_WAVE; j++) {
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int q = j * 2 + d;
                if (!((m >> q) & 1ull)) continue;
                const unsigned long long live = uni64(__ballot(!(w[j] & ST_E)));
                const unsigned long long r0 = lane64(reach_l, q);
                const unsigned long long s = lane64(pl, q) & live & ~r0;
                if (s == 0ull) continue;
                // closure inside the cell: lane u joins when a claimed vertex reaches it there
                const unsigned long long r = r0 | (uni64(__ballot((ct[j][d] & s) != 0ull)) & live);
                bound += CELL_XARCS * (uint32_t)__popcll(r & ~r0);
                if (lane == (uint32_t)q) {
                    reach_l = r;
                    fr_l = r & ~r0;
                }
            }
        }
        // claims into the other cells of the tile, counted first (by a bound; the count of
        // the words they did not turn non-zero is taken back afterwards)
        if (bound) {
            if (lane == 0) atomicAdd(&sh.pending, bound);
            lds_fence();  // counted before any word turns non-zero
        }
        uint32_t made = 0;  // claims that turned a word non-zero
        unsigned long long xm_l = 0ull;  // lane q: the new vertices of pair q whose arcs leave the tile
        uint32_t nx = 0;
#pragma unroll
        for (int j = 0; j < CELLS_PER_WAVE; j++) {
            const uint32_t i = (wv + TILE_WAVES * j) * CELL + lane;
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int q = j * 2 + d;
                const unsigned long long fr = lane64(fr_l, q);
                if (fr == 0ull) continue;
                n_reached += (uint32_t)__popcll(fr);
                const bool mine = (fr >> lane) & 1ull;
                const unsigned long long xq = mine ? sh.xa[d * TILE_MAX + i] : ~0ull;
                // the four claims issued together (an unused slot ORs nothing into the lane's
                // own sink word: no two lanes meet on one address for it)
                unsigned long long old[CELL_XARCS];
#pragma unroll
                for (int k = 0; k < CELL_XARCS; k++) {
                    const uint32_t u = (uint32_t)((xq >> (16 * k)) & 0xFFFFull);
                    const bool ok = u != 0xFFFFu;
                    old[k] = atomicOr(ok ? &sh.claim[cidx(u / CELL, d)] : &sh.sink[lane], ok ? 1ull << (u % CELL) : 0ull);
                }
#pragma unroll
                for (int k = 0; k < CELL_XARCS; k++)
                    made += (uint32_t)__popcll(__ballot(((xq >> (16 * k)) & 0xFFFFull) != 0xFFFFull && old[k] == 0ull));
                if (mine && ((xfl >> (8 * j + 2 + d)) & 1u))  // more such arcs: the whole row
                    cross_row(sh, d ? a.rowT : a.row, d ? a.colT : a.col, lo, T, i, d);
                // the new bits back: at once where other tiles may claim too (no return value
                // awaited), at the end elsewhere
                if (mine) {
                    if (xfl & (3u << (8 * j))) atomicOr(&a.word[lo + (int32_t)i], d ? ST_B : ST_F);
                    else newbits |= (d ? ST_B : ST_F) << (8 * j);
                }
                const unsigned long long xm = __ballot(mine && ((xfl >> (8 * j + d)) & 1u));
                if (lane == (uint32_t)q) xm_l = xm;
                nx += (uint32_t)__popcll(xm);
            }
        }
        // the new vertices whose arcs leave the tile listed (one LDS atomic per take)
        if (nx) {
            uint32_t at = 0;
            if (lane == 0) at = atomicAdd(&sh.n_exit, nx);
            at = (uint32_t)__builtin_amdgcn_readfirstlane((int)at);
#pragma unroll
            for (int j = 0; j < CELLS_PER_WAVE; j++) {
                const uint32_t i = (wv + TILE_WAVES * j) * CELL + lane;
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    const unsigned long long xm = lane64(xm_l, j * 2 + d);
                    if ((xm >> lane) & 1ull)
                        sh.exits[at + (uint32_t)__popcll(xm & below)] =
                            i | ((uint32_t)d << 16) | (((xfl >> (8 * j + 4 + d)) & 1u) << 17);
                    at += (uint32_t)__popcll(xm);
                }
            }
        }
        if (lane == 0) atomicSub(&sh.pending, (bound - made) + (uint32_t)__popcll(m));  // the words taken are done
        if (a.instr) busy += wall_clock64() - t_take;
    }
    if (lane < 8) done[base + lane] = reach_l;
#pragma unroll
    for (int j = 0; j < CELLS_PER_WAVE; j++) {  // only this run writes these words during the reach
        const uint32_t nb = (newbits >> (8 * j)) & 0xFFu;
        if (nb) a.word[lo + (int32_t)((wv + TILE_WAVES * j) * CELL + lane)] = w[j] | nb;
    }
    tile_sync();
    if (a.instr) clk[2] = wall_clock64();
    if (tid == 0 && sh.stuck) atomicAdd(&a.ctl->terr, 1u);
    if (a.instr) {  // arcs of the vertices reached in this run (the inbox's vertices count as new)
        uint32_t arcs = 0;
#pragma unroll
        for (int j = 0; j < CELLS_PER_WAVE; j++) {
            const uint32_t i = (wv + TILE_WAVES * j) * CELL + lane;
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int q = j * 2 + d;
                const unsigned long long fresh = lane64(reach_l, q) & ~lane64(done0, q);
                if ((fresh >> lane) & 1ull) {
                    const int32_t *row = d ? a.rowT : a.row;
                    arcs += (uint32_t)(row[lo + (int32_t)i + 1] - row[lo + (int32_t)i]);
                }
            }
        }
        for (int o = WAVE / 2; o > 0; o >>= 1) arcs += (uint32_t)__shfl_xor(arcs, o, WAVE);
        if (lane == 0) {
            atomicAdd(&sh.ins[0], busy);
            atomicAdd(&sh.ins[1], (unsigned long long)idles);
            atomicAdd(&sh.ins[2], (unsigned long long)polls);
            atomicMax(&sh.ins[3], (unsigned long long)polls);
            atomicAdd(&sh.ins[4], (unsigned long long)n_reached);
            atomicAdd(&sh.ins[5], (unsigned long long)arcs);
        }
        clk[3] = wall_clock64();
    }
    // claims across the tile boundary from the listed vertices: their neighbours outside the
    // tile kept at upload, or their whole rows when they have more
    const uint32_t ne = sh.n_exit;
    for (uint32_t q = tid; q < ne; q += TILE_BLOCK) {
        const uint32_t ent = sh.exits[q];
        const uint32_t i = ent & 0xFFFFu;
        const int d = (int)((ent >> 16) & 1u);
        const int32_t v = lo + (int32_t)i;
        found = true;
        if (!((ent >> 17) & 1u)) {
            const int2 xo = a.xout[d][v];
            if (xo.x >= 0) claim_across(a, d, xo.x);
            if (xo.y >= 0) claim_across(a, d, xo.y);
        } else {  // more than two: the row
            const int32_t *row = d ? a.rowT : a.row, *col = d ? a.colT : a.col;
            for (int32_t e = row[v]; e < row[v + 1]; e++) {
                const int32_t u = col[e];
                if (u < lo || u >= hi) claim_across(a, d, u);
            }
        }
    }
    tile_sync();  // the shared claims are cleared for the next tile
    if (a.instr && tid == 0) {  // once per run (the wave counts were summed in LDS)
        clk[4] = wall_clock64();
        for (int q = 0; q < 4; q++) atomicAdd(&a.ctl->st_tclk[q], clk[q + 1] - clk[q]);
        atomicMax(&a.ctl->tlog[a.iter & 3][2], (uint32_t)(clk[4] - clk[0] - (clk[3] - clk[2])));
        atomicAdd(&a.ctl->st_tbusy, sh.ins[0]);
        atomicAdd(&a.ctl->st_tidle, sh.ins[1]);
        atomicAdd(&a.ctl->st_trounds, sh.ins[2]);
        atomicMax(&a.ctl->tlog[a.iter & 3][1], (uint32_t)sh.ins[3]);
        atomicAdd(&a.ctl->st_expanded, sh.ins[4]);
        atomicAdd(&a.ctl->st_visited, sh.ins[4]);
        atomicAdd(&a.ctl->st_edges, sh.ins[5]);
        atomicAdd(&a.ctl->st_truns, 1ull);
        atomicAdd(&a.ctl->tlog[a.iter & 3][0], 1u);
    }
}

__global__ __launch_bounds__(TILE_BLOCK) void k_tile_reach(TileArgs a) {
    __shared__ TileShared sh;
    __shared__ uint32_t wcnt[TILE_WAVES];
    __shared__ int32_t mine[TILE_LIST_MAX];
    __shared__ uint32_t n_mine;
    if (chain_exit(a.ctl, a.iter)) return;
    const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    if (a.instr && blockIdx.x == 0 && tid == 0)  // the iteration log entry after next
        a.ctl->tlog[(a.iter + 2) & 3][0] = a.ctl->tlog[(a.iter + 2) & 3][1] = a.ctl->tlog[(a.iter + 2) & 3][2] = 0;
    bool found = false;
    if (a.iter == 0) {
        // the tiles of the pivots (the seed list, ctl->qsize[0][0] entries; reached both ways
        // by their election), flagged for iteration 1
        const uint32_t np = a.ctl->qsize[0][0];
        for (uint32_t k = blockIdx.x * TILE_BLOCK + tid; k < np; k += gridDim.x * TILE_BLOCK) {
            a.flags[a.ntiles + a.piv[k] / a.tsize] = 1;
            found = true;
        }
        // every cell's closure, and nothing expanded yet
        for (int32_t t = (int32_t)blockIdx.x; t < a.ntiles; t += (int32_t)gridDim.x) {
            cell_closure(a, t);
            for (uint32_t i = tid; i < TILE_CELLS * 2u; i += TILE_BLOCK)
                a.done[(size_t)t * (TILE_CELLS * 2) + i] = 0ull;
        }
    } else {
        // the flagged tiles, ranked in tile order; rank g goes to workgroup g % grid.  The
        // flags of iteration i + 2 are cleared meanwhile (iteration i - 1 read them)
        if (blockIdx.x == 0)
            for (int32_t t = (int32_t)tid; t < a.ntiles; t += TILE_BLOCK)
                a.flags[(size_t)((a.iter + 2) % 3) * a.ntiles + t] = 0;
        const uint8_t *flags = a.flags + (size_t)(a.iter % 3) * a.ntiles;
        const unsigned long long below = (1ull << lane) - 1ull;
        uint32_t ranked = 0;  // flagged tiles before this chunk
        for (int32_t c0 = 0; c0 < a.ntiles; c0 += TILE_BLOCK) {  // block-uniform
            const int32_t t = c0 + (int32_t)tid;
            const bool f = t < a.ntiles && flags[t];
            const unsigned long long m = __ballot(f);
            if (lane == 0) wcnt[wv] = (uint32_t)__popcll(m);
            if (tid == 0) n_mine = 0;
            __syncthreads();
            uint32_t pre = 0, tot = 0;
            for (int q = 0; q < TILE_WAVES; q++) {
                const uint32_t x = wcnt[q];
                pre += q < (int)wv ? x : 0u;
                tot += x;
            }
            if (f && (ranked + pre + (uint32_t)__popcll(m & below)) % gridDim.x == blockIdx.x)
                mine[atomicAdd(&n_mine, 1u)] = t;  // at most TILE_LIST_MAX per chunk: grid >= 64
            __syncthreads();
            const uint32_t nm = n_mine;
            for (uint32_t k = 0; k < nm; k++) tile_run(a, sh, mine[k], found);  // block-uniform
            ranked += tot;
            __syncthreads();  // wcnt, mine and n_mine are rewritten for the next chunk
        }
    }
    chain_progress(a.ctl, a.iter, found);
}

void launch_tile_reach(hipStream_t s, uint32_t grid, const TileArgs &a) {
    hipLaunchKernelGGL(k_tile_reach, dim3(grid), dim3(TILE_BLOCK), 0, s, a);
}

hipError_t preload_tile() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(k_tile_reach));
}

}  // namespace gpuscc
