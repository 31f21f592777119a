// Reachability over tiles (first FB step of a regular graph).  On a graph of
// high diameter the reach from the first pivot takes thousands of BFS levels
// (the 2048 x 2048 mesh: about 4,100), and the level kernels pay a global
// memory round trip per level even with workgroup-local rounds.  Here the
// vertices are numbered tile by tile at upload (regions of up to TILE_MAX
// vertices grown by BFS, so a tile is a compact piece of the graph) and the
// arcs inside each tile are kept as a local CSR of 16-bit indices.  One
// iteration of the chain gives every tile of its list one 1024-wide workgroup,
// which copies the tile's words and local CSR into LDS, starts from the
// vertices claimed into the tile across its boundary since its last run (the
// tile's inbox; the pivots in iteration 0), reaches F and B inside the tile to
// the fixpoint with LDS atomics and barriers only, writes the new bits back,
// and claims the targets of the reached vertices' arcs that leave the tile
// (global atomics, as the level kernels claim), putting them in their tiles'
// inboxes and those tiles in the next list.  The number of iterations is the
// number of tile boundaries on the longest path (45 on the mesh), and each BFS
// level inside a tile costs LDS operations and two barriers instead of a
// global round trip.  The reach is the same closure, so the F / B bits are the
// level kernels' (the reach of scc.cu:285-704 over reach_q.cu:12-79's rule:
// live, same colour, not yet reached).
#include "fb_device.hpp"
#include "kernels.hpp"

namespace gpuscc {

constexpr int TILE_BLOCK = 1024;

__shared__ __align__(16) uint8_t tile_arena[TILE_ARENA];

// One workgroup works on a tile: a workgroup-scope barrier orders its LDS and
// global accesses (as in the tail, agent-scope fences would write back the L2)
__device__ inline void tile_sync() { __syncthreads(); }

struct TileShared {
    uint32_t cnt[2];   // vertices queued per direction (all reached in this run, in order)
    uint32_t head[2];  // start of the current level
    uint32_t n_in[2];  // inbox entries taken
};

// LDS layout of a tile of T vertices and A arcs (both directions):
// words u32[T] | offsets u16[T + 1] per direction | queues u16[T] per direction
// | flags u8[T] | arcs u16[A].  tile_lds_bytes (kernels.hpp) gives its size.
struct TileLds {
    uint32_t T;
    __device__ uint32_t *w() const { return (uint32_t *)tile_arena; }
    __device__ uint16_t *off(int d) const { return (uint16_t *)(tile_arena + 4u * T) + (d ? T + 1 : 0); }
    __device__ uint16_t *q(int d) const { return (uint16_t *)(tile_arena + 4u * T + 4u * (T + 1)) + (d ? T : 0); }
    __device__ uint8_t *fl() const { return tile_arena + 8u * T + 4u * (T + 1); }
    __device__ uint16_t *arcs() const { return (uint16_t *)(tile_arena + ((9u * T + 4u * (T + 1) + 1u) & ~1u)); }
};

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

// one tile's run; false when it does not fit (never, by the upload's check)
__device__ bool tile_run(const TileArgs &a, TileShared &sh, int32_t t, bool &found) {
    const uint32_t tid = threadIdx.x, nt = TILE_BLOCK;
    const int32_t lo = a.toff[t], hi = a.toff[t + 1];
    const uint32_t T = (uint32_t)(hi - lo);
    const int32_t p0 = a.lpos[0][lo], p1 = a.lpos[1][lo];
    const uint32_t A0 = (uint32_t)(a.lpos[0][hi] - p0), A1 = (uint32_t)(a.lpos[1][hi] - p1);
    if (T > TILE_MAX || tile_lds_bytes(T, A0 + A1) > TILE_ARENA) return false;  // block-uniform
    const TileLds L{T};
    uint32_t *w = L.w();
    uint16_t *o0 = L.off(0), *o1 = L.off(1), *q0 = L.q(0), *q1 = L.q(1), *ar = L.arcs();
    uint8_t *fl = L.fl();
    const int par = a.iter & 1;
    // the tile: words, flags, local CSR (offsets relative to the tile's first arc)
    for (uint32_t i = tid; i < T; i += nt) {
        w[i] = a.word[lo + (int32_t)i];
        fl[i] = a.xflag[lo + (int32_t)i];
    }
    for (uint32_t i = tid; i <= T; i += nt) {
        o0[i] = (uint16_t)(a.lpos[0][lo + (int32_t)i] - p0);
        o1[i] = (uint16_t)(A0 + (uint32_t)(a.lpos[1][lo + (int32_t)i] - p1));
    }
    for (uint32_t i = tid; i < A0; i += nt) ar[i] = a.larc[0][p0 + (int32_t)i];
    for (uint32_t i = tid; i < A1; i += nt) ar[A0 + i] = a.larc[1][p1 + (int32_t)i];
    if (tid < 2) {
        sh.n_in[tid] = a.incnt[(par * 2 + (int)tid) * a.ntiles + t];
        sh.cnt[tid] = 0;
        sh.head[tid] = 0;
    }
    tile_sync();
    // seeds: the inbox (claimed into the tile last iteration; the pivots in iteration 0),
    // each queued once per direction whatever its copied word says
    for (int d = 0; d < 2; d++) {
        const uint32_t ni = sh.n_in[d];
        const int32_t *in = a.inbox + (size_t)(par * 2 + d) * a.n + lo;
        uint16_t *q = d ? q1 : q0;
        const uint32_t bit = d ? ST_B : ST_F;
        for (uint32_t k = tid; k < ni; k += nt) {
            const uint32_t i = (uint32_t)(in[k] - lo);
            atomicOr(&w[i], bit);  // a seed is queued even when the copy held its bit already
            q[atomicAdd(&sh.cnt[d], 1u)] = (uint16_t)i;
        }
    }
    tile_sync();
    if (tid < 2) a.incnt[(par * 2 + (int)tid) * a.ntiles + t] = 0;  // consumed
    // F and B inside the tile to the fixpoint, a level per pass
    for (;;) {
        const uint32_t h0 = sh.head[0], h1 = sh.head[1], e0 = sh.cnt[0], e1 = sh.cnt[1];
        const uint32_t n0 = e0 - h0, n1 = e1 - h1;
        if (n0 + n1 == 0) break;  // block-uniform (read after the barrier)
        for (uint32_t k = tid; k < n0 + n1; k += nt) {
            const int d = k < n0 ? 0 : 1;
            const uint32_t v = d == 0 ? q0[h0 + k] : q1[h1 + k - n0];
            const uint32_t bit = d ? ST_B : ST_F, c = colour_of(w[v]);
            const uint16_t *o = d ? o1 : o0;
            uint16_t *q = d ? q1 : q0;
            for (uint32_t e = o[v]; e < o[v + 1]; e++) {
                const uint32_t u = ar[e];
                const uint32_t wu = w[u];
                if ((wu & (bit | ST_E)) || colour_of(wu) != c) continue;
                if (atomicOr(&w[u], bit) & bit) continue;
                q[atomicAdd(&sh.cnt[d], 1u)] = (uint16_t)u;
            }
        }
        tile_sync();
        if (tid == 0) {
            sh.head[0] = e0;
            sh.head[1] = e1;
        }
        tile_sync();
    }
    // the new bits back (others may claim the tile's vertices meanwhile: atomics), and the
    // arcs that leave the tile from every vertex reached in this run
    const uint32_t c0 = sh.cnt[0], c1 = sh.cnt[1];
    for (uint32_t k = tid; k < c0 + c1; k += nt) {
        const int d = k < c0 ? 0 : 1;
        const uint32_t i = d == 0 ? q0[k] : q1[k - c0];
        const uint32_t bit = d ? ST_B : ST_F;
        const int32_t v = lo + (int32_t)i;
        atomicOr(&a.word[v], bit);
        if (!((fl[i] >> d) & 1u)) continue;
        const uint32_t c = colour_of(w[i]);
        const int32_t *row = d ? a.rowT : a.row, *col = d ? a.colT : a.col;
        for (int32_t e = row[v]; e < row[v + 1]; e++) {
            const int32_t u = col[e];
            if (u >= lo && u < hi) continue;
            const uint32_t wu = a.word[u];
            if ((wu & (bit | ST_E)) || colour_of(wu) != c) continue;
            if (atomicOr(&a.word[u], bit) & bit) continue;
            tile_post(a, d, u);
            found = true;
        }
    }
    tile_sync();  // LDS is rebuilt for the next tile
    return true;
}

__global__ __launch_bounds__(TILE_BLOCK) void k_tile_reach(TileArgs a) {
    __shared__ TileShared sh;
    if (chain_exit(a.ctl, a.iter)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl->tcnt[(a.iter + 2) & 3] = 0;  // the list after next
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
    } else {
        const uint32_t cnt = a.ctl->tcnt[a.iter & 3];
        const int32_t *list = a.list[a.iter & 1];
        for (uint32_t k = blockIdx.x; k < cnt; k += gridDim.x)  // block-uniform
            if (!tile_run(a, sh, list[k], found) && threadIdx.x == 0) atomicAdd(&a.ctl->terr, 1u);
    }
    chain_progress(a.ctl, a.iter, found);
}

void launch_tile_reach(hipStream_t s, uint32_t grid, const TileArgs &a) {
    hipLaunchKernelGGL(k_tile_reach, dim3(grid), dim3(TILE_BLOCK), 0, s, a);
}

}  // namespace gpuscc
