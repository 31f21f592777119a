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
static_assert(CELL == WAVE, "a cell is one vertex per lane of a wave");

// claims into cell k, direction d: [cidx(k, d)], bit = lane; a wave's own entries are
// contiguous (cell k belongs to wave k % 16, as its entry k / 16)
__device__ inline uint32_t cidx(uint32_t k, int d) {
    return ((k % TILE_WAVES) * CELLS_PER_WAVE + k / TILE_WAVES) * 2u + (uint32_t)d;
}

struct TileShared {
    unsigned long long claim[TILE_CELLS * 2];    // [cidx]: claims not yet taken by the cell's wave
    unsigned long long reached[TILE_CELLS * 2];  // [cidx]: reached so far (written by the cell's wave only)
    unsigned long long seeds[TILE_CELLS * 2];    // [cidx]: the inbox's vertices
    // the tile's arcs to other cells ([direction][vertex]); after the fixpoint, the list of
    // (vertex, direction) whose arcs leave the tile
    union {
        unsigned long long xa[2 * TILE_MAX];
        uint32_t exits[2 * TILE_MAX];
    };
    // claim words that are non-zero or taken and not yet processed: a depositor counts a
    // word before it can turn non-zero (and takes the count back when it was non-zero
    // already), a wave uncounts a word it took once it has made all its own claims, so the
    // count never falls below the work left and zero means the tile is at its fixpoint
    uint32_t pending;
    unsigned long long sink[WAVE];  // a lane's claim slot without a target ORs nothing into its own word
    uint32_t n_in[2];  // inbox entries per direction
    uint32_t n_exit;
    uint32_t stuck;    // a wave gave up waiting (never: the count reaches zero)
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

__device__ inline void tile_push_next(const TileArgs &a, int32_t t2) {
    // join the next iteration's list once: the stamp holds the last list the tile joined
    const unsigned long long s = (a.run << 32) | (unsigned long long)(uint32_t)(a.iter + 2);
    if (atomicMax(&a.stamp[t2], s) < s) {
        const uint32_t k = atomicAdd(&a.ctl->tcnt[(a.iter + 1) & 3], 1u);
        a.list[(a.iter + 1) & 1][k] = t2;
    }
}

// u, reached across an arc of direction d, into its tile's inbox for the next iteration
__device__ inline void tile_post(const TileArgs &a, int d, int32_t u) {
    const int32_t t2 = a.tile_of[u];
    const int par = (a.iter + 1) & 1;
    const uint32_t k = atomicAdd(&a.incnt[(par * 2 + d) * a.ntiles + t2], 1u);
    a.inbox[(size_t)(par * 2 + d) * a.n + a.toff[t2] + k] = u;
    tile_push_next(a, t2);
}

// claim bit of word idx for another cell, counted in pending first
__device__ inline void deposit(TileShared &sh, uint32_t idx, unsigned long long bit) {
    atomicAdd(&sh.pending, 1u);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // counted before the word can turn non-zero
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

// a wave's sum
__device__ inline uint32_t wave_sum(uint32_t x) {
    for (int o = WAVE / 2; o > 0; o >>= 1) x += (uint32_t)__shfl_xor(x, o, WAVE);
    return x;
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
    const int par = a.iter & 1;
    unsigned long long clk[5];  // phase boundaries (instrumented runs, thread 0)
    if (a.instr) clk[0] = wall_clock64();
    for (uint32_t i = tid; i < TILE_CELLS * 2u; i += TILE_BLOCK) sh.claim[i] = 0ull;
    for (uint32_t i = tid; i < T; i += TILE_BLOCK) {
        sh.xa[i] = a.xarc[0][lo + (int32_t)i];
        sh.xa[TILE_MAX + i] = a.xarc[1][lo + (int32_t)i];
    }
    if (tid < 2) sh.n_in[tid] = a.incnt[(par * 2 + (int)tid) * a.ntiles + t];
    if (tid == 0) {
        sh.n_exit = 0;
        sh.pending = 0;
        sh.stuck = 0;
    }
    if (tid < WAVE) sh.sink[tid] = 0ull;
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
    tile_sync();
    // the inbox (claimed into the tile across its boundary since its last run; the pivots
    // in iteration 0) as the first claims
    for (int d = 0; d < 2; d++) {
        const uint32_t ni = sh.n_in[d];
        const int32_t *in = a.inbox + (size_t)(par * 2 + d) * a.n + lo;
        for (uint32_t k0 = wv * WAVE; k0 < ni; k0 += TILE_BLOCK) {  // wave-uniform
            const uint32_t k = k0 + lane;
            const bool ok = k < ni;
            const uint32_t i = ok ? (uint32_t)(in[k] - lo) : 0u;
            // counted for the whole wave first, taken back for words that were non-zero
            const uint32_t cnt = (uint32_t)__popcll(__ballot(ok));
            if (lane == 0) atomicAdd(&sh.pending, cnt);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            const unsigned long long old =
                atomicOr(ok ? &sh.claim[cidx(i / CELL, d)] : &sh.sink[lane], ok ? 1ull << (i % CELL) : 0ull);
            const uint32_t dup = (uint32_t)__popcll(__ballot(ok && old != 0ull));
            if (lane == 0 && dup) atomicSub(&sh.pending, dup);
        }
    }
    tile_sync();
    if (tid < 2) a.incnt[(par * 2 + (int)tid) * a.ntiles + t] = 0;  // consumed
    // reached before this run: the words' bits, except the inbox's vertices, which are
    // expanded whatever their words say (a claim sets the word before the tile runs).  Lane
    // q < 8 of a wave keeps the masks of its pair q = 2j + d (its cell j, direction d).
    const uint32_t base = wv * CELLS_PER_WAVE * 2;  // this wave's claim words
    if (a.instr) clk[1] = wall_clock64();
    const unsigned long long seed_l = lane < 8 ? sh.claim[base + lane] : 0ull;
    unsigned long long reach_l = 0ull;
#pragma unroll
    for (int j = 0; j < CELLS_PER_WAVE; j++) {
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const unsigned long long wb = __ballot(w[j] & (d ? ST_B : ST_F));
            if (lane == (uint32_t)(j * 2 + d)) reach_l = wb;
        }
    }
    reach_l &= ~seed_l;
    tile_sync();  // every wave has its inbox claims before claims inside the tile begin
    // to the fixpoint, each wave on its own: take every claim word of its cells at once, close
    // the claims inside their cells, claim into other cells (counted first), uncount the words
    // taken; stop when nothing is counted
    uint32_t polls = 0, idles = 0;
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
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // counted before any word turns non-zero
        }
        uint32_t made = 0;  // claims that turned a word non-zero
#pragma unroll
        for (int j = 0; j < CELLS_PER_WAVE; j++) {
            const uint32_t i = (wv + TILE_WAVES * j) * CELL + lane;
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int q = j * 2 + d;
                const unsigned long long fr = lane64(fr_l, q);
                if (fr == 0ull) continue;
                const bool mine = (fr >> lane) & 1ull;
                const unsigned long long x = mine ? sh.xa[d * TILE_MAX + i] : ~0ull;
                // the four claims issued together (an unused slot ORs nothing into the lane's
                // own sink word: no two lanes meet on one address for it)
                unsigned long long old[CELL_XARCS];
#pragma unroll
                for (int k = 0; k < CELL_XARCS; k++) {
                    const uint32_t u = (uint32_t)((x >> (16 * k)) & 0xFFFFull);
                    const bool ok = u != 0xFFFFu;
                    old[k] = atomicOr(ok ? &sh.claim[cidx(u / CELL, d)] : &sh.sink[lane], ok ? 1ull << (u % CELL) : 0ull);
                }
#pragma unroll
                for (int k = 0; k < CELL_XARCS; k++)
                    made += (uint32_t)__popcll(__ballot(((x >> (16 * k)) & 0xFFFFull) != 0xFFFFull && old[k] == 0ull));
                if (mine && ((xfl >> (8 * j + 2 + d)) & 1u))  // more such arcs: the whole row
                    cross_row(sh, d ? a.rowT : a.row, d ? a.colT : a.col, lo, T, i, d);
            }
        }
        const uint32_t dups = bound - made;
        if (lane == 0) atomicSub(&sh.pending, dups + (uint32_t)__popcll(m));  // the words taken are done
        if (a.instr) busy += wall_clock64() - t_take;
    }
    if (lane < 8) {  // the masks, for the write-back
        sh.reached[base + lane] = reach_l;
        sh.seeds[base + lane] = seed_l;
    }
    tile_sync();
    if (tid == 0 && sh.stuck) atomicAdd(&a.ctl->terr, 1u);
    if (a.instr) {
        if (lane == 0) {
            atomicAdd(&a.ctl->st_tbusy, busy);
            atomicAdd(&a.ctl->st_tidle, (unsigned long long)idles);
            atomicAdd(&a.ctl->st_trounds, (unsigned long long)polls);
            atomicMax(&a.ctl->tlog[a.iter & 3][1], polls);
        }
        if (tid == 0) {
            atomicAdd(&a.ctl->st_truns, 1ull);
            atomicAdd(&a.ctl->tlog[a.iter & 3][0], 1u);
        }
        clk[2] = wall_clock64();
    }
    // the new bits back (others may claim the tile's vertices meanwhile: atomics), and the
    // vertices reached in this run whose arcs leave the tile listed (the arcs to other cells
    // are no longer needed)
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t n_reached = 0, n_arcs = 0;  // instrumented counts of this lane
#pragma unroll
    for (int j = 0; j < CELLS_PER_WAVE; j++) {
        const uint32_t i = (wv + TILE_WAVES * j) * CELL + lane;
        // reached in this run: reached now, and not before (the inbox's vertices count as new)
        uint32_t bits = 0;
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const uint32_t bit = d ? ST_B : ST_F;
            const unsigned long long r = sh.reached[base + j * 2 + d], sd = sh.seeds[base + j * 2 + d];
            if ((((r & (sd | ~__ballot(w[j] & bit))) >> lane) & 1ull) != 0) bits |= bit;
        }
        if (bits) atomicOr(&a.word[lo + (int32_t)i], bits);
#pragma unroll
        for (int d = 0; d < 2; d++) {  // one LDS atomic per wave and list
            const bool x = (bits & (d ? ST_B : ST_F)) && ((xfl >> (8 * j + d)) & 1u);
            const unsigned long long m = __ballot(x);
            if (m == 0ull) continue;
            uint32_t at = 0;
            if (lane == (uint32_t)__builtin_ctzll(m)) at = atomicAdd(&sh.n_exit, (uint32_t)__popcll(m));
            at = (uint32_t)__shfl(at, __builtin_ctzll(m), WAVE);
            if (x) sh.exits[at + (uint32_t)__popcll(m & below)] = i | ((uint32_t)d << 16);
        }
        if (a.instr && bits) {  // each vertex reached in this run expands its row in each direction
            const int32_t v = lo + (int32_t)i;
            n_reached += (bits & ST_F ? 1u : 0u) + (bits & ST_B ? 1u : 0u);
            n_arcs += (uint32_t)((bits & ST_F ? a.row[v + 1] - a.row[v] : 0) +
                                 (bits & ST_B ? a.rowT[v + 1] - a.rowT[v] : 0));
        }
    }
    if (a.instr) {
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            n_reached += (uint32_t)__shfl_xor(n_reached, o, WAVE);
            n_arcs += (uint32_t)__shfl_xor(n_arcs, o, WAVE);
        }
        if (lane == 0 && n_reached) {
            atomicAdd(&a.ctl->st_expanded, (unsigned long long)n_reached);
            atomicAdd(&a.ctl->st_visited, (unsigned long long)n_reached);
            atomicAdd(&a.ctl->st_edges, (unsigned long long)n_arcs);
        }
    }
    tile_sync();
    if (a.instr) clk[3] = wall_clock64();
    // claims across the tile boundary: the level kernels' test (live, not yet reached; in a
    // first FB step every live vertex has the one colour), the claimed vertex into its tile's
    // inbox.  The target's tile and the tile's first vertex are loaded with its word, ahead
    // of the claim.
    const uint32_t ne = sh.n_exit;
    for (uint32_t q = tid; q < ne; q += TILE_BLOCK) {
        const uint32_t i = sh.exits[q] & 0xFFFFu;
        const int d = (int)(sh.exits[q] >> 16);
        const int32_t v = lo + (int32_t)i;
        const uint32_t bit = d ? ST_B : ST_F;
        const int32_t *row = d ? a.rowT : a.row, *col = d ? a.colT : a.col;
        const int32_t e1 = row[v + 1];
        for (int32_t e = row[v]; e < e1; e++) {
            const int32_t u = col[e];
            if (u >= lo && u < hi) continue;
            const uint32_t wu = a.word[u];
            const int32_t t2 = a.tile_of[u];
            if (wu & (bit | ST_E)) continue;
            const int32_t base2 = a.toff[t2];
            if (atomicOr(&a.word[u], bit) & bit) continue;
            const int par2 = (a.iter + 1) & 1;
            const uint32_t k = atomicAdd(&a.incnt[(par2 * 2 + d) * a.ntiles + t2], 1u);
            a.inbox[(size_t)(par2 * 2 + d) * a.n + base2 + k] = u;
            tile_push_next(a, t2);
            found = true;
        }
    }
    tile_sync();  // the shared claims are cleared for the next tile
    if (a.instr && tid == 0) {
        clk[4] = wall_clock64();
        for (int q = 0; q < 4; q++) atomicAdd(&a.ctl->st_tclk[q], clk[q + 1] - clk[q]);
        atomicMax(&a.ctl->tlog[a.iter & 3][2], (uint32_t)(clk[4] - clk[0]));
    }
}

__global__ __launch_bounds__(TILE_BLOCK) void k_tile_reach(TileArgs a) {
    __shared__ TileShared sh;
    if (chain_exit(a.ctl, a.iter)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctl->tcnt[(a.iter + 2) & 3] = 0;  // the list after next
        if (a.instr) a.ctl->tlog[(a.iter + 2) & 3][0] = a.ctl->tlog[(a.iter + 2) & 3][1] =
                         a.ctl->tlog[(a.iter + 2) & 3][2] = 0;
    }
    bool found = false;
    if (a.iter == 0) {
        // the pivots (the seed list, ctl->qsize[0][0] entries), reached both ways by their
        // election, into their tiles' inboxes
        const uint32_t np = a.ctl->qsize[0][0];
        for (uint32_t k = blockIdx.x * TILE_BLOCK + threadIdx.x; k < np; k += gridDim.x * TILE_BLOCK) {
            tile_post(a, 0, a.piv[k]);
            tile_post(a, 1, a.piv[k]);
            found = true;
        }
        // and every cell's closure, for the iterations that follow
        for (int32_t t = (int32_t)blockIdx.x; t < a.ntiles; t += (int32_t)gridDim.x) cell_closure(a, t);
    } else {
        const uint32_t cnt = a.ctl->tcnt[a.iter & 3];
        const int32_t *list = a.list[a.iter & 1];
        for (uint32_t k = blockIdx.x; k < cnt; k += gridDim.x)  // block-uniform
            tile_run(a, sh, list[k], found);
    }
    chain_progress(a.ctl, a.iter, found);
}

void launch_tile_reach(hipStream_t s, uint32_t grid, const TileArgs &a) {
    hipLaunchKernelGGL(k_tile_reach, dim3(grid), dim3(TILE_BLOCK), 0, s, a);
}

}  // namespace gpuscc
