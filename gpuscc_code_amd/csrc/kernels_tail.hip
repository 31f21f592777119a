// The tail of an FB run in one workgroup.  Once few vertices are live (after
// the first FB steps of a high-diameter graph only a fraction of a percent
// stay live, spread over many small colour sets), an FB step is about a dozen
// kernel launches that have almost nothing to do, and their launch and
// inter-kernel costs, not their work, set the run time: the plain-variant mesh
// takes 387 such steps.  k_fb_tail runs the remaining steps in one launch of
// one workgroup, with barriers in place of kernel boundaries.  Each step does
// what the multi-kernel path does, so the final colours, states and roots are
// the same (the reach is a closure, the trim fixpoint is unique, and the
// election keys are the same):
//   reach F and B from the pivot list inside colours (reach_q.cu:12-79),
//   trim to the fixpoint, skipping F-and-B non-pivots (trim_kernel,
//   scc_routines.cu:336-394), update + election + confirm
//   (update_kernel, scc_routines.cu:187-224, lowest key per slot).
// The steps first run over global memory and the live list of the last
// election (GlobalSpace).  As soon as the live subgraph fits in LDS -- the
// words and global ids of the live vertices, and the arcs that join two live
// vertices of one colour, the only arcs any later step follows, since colours
// only split and eliminated vertices stay eliminated -- it is copied there once
// and the remaining steps touch global memory only for roots and the election
// slots (LocalSpace); the words, the live list and the pivot list are written
// back at the end.
#include "fb_device.hpp"
#include "kernels.hpp"

namespace gpuscc {

constexpr int TAIL_BLOCK = 1024;
// LDS for the local copy: words and ids (8 B per vertex), out and in row
// offsets (8 B), four uint16 queues (8 B), two uint16 live lists (4 B) and uint16 arcs
constexpr uint32_t TAIL_ARENA = 150u * 1024u;

__host__ __device__ inline uint32_t tail_fixed_bytes(uint32_t L) { return 28u * L + 16u; }

// The local copy lives here; its parts are addressed by offsets from this
// array, so every access compiles to an LDS instruction.
__shared__ __align__(16) uint8_t tail_arena[TAIL_ARENA];

// Every access of the tail comes from this one workgroup, whose waves share
// one CU and its L1, so a workgroup-scope barrier orders them (the release /
// acquire fences of __syncthreads).  Agent-scope fences would write back and
// invalidate the XCD's L2 at every barrier (the L2s of the XCDs are not
// coherent with each other), several microseconds each; the kernel boundaries
// order the tail with the kernels before and after it.
__device__ inline void tail_sync() { __syncthreads(); }

struct TailShared {
    uint32_t cnt[4];  // queue sizes: F in, B in, F out, B out
    uint32_t npiv, lcnt, first, flag;
    unsigned long long red[2];  // this step: vertices leaving the live set, new P by trims
    uint32_t scan[TAIL_BLOCK / WAVE];
};

// The state of one run of steps.  The buffers a run works in (queues, live
// lists, pivot list: a batch graph's slices of them) and its step bound are
// kept here rather than in a modified copy of the kernel's TailArgs: a copy
// whose pointer arrays are indexed at run time would live in scratch memory,
// and every access to it would be a global load.  Two-element choices are
// selects for the same reason.
struct TailRun {
    int lsel, pt;
    uint32_t epoch, npiv;
    unsigned long long depths, trims;
    uint32_t lc0, lc1;  // a batch graph's live-list lengths (the tail keeps them in ctl->lcnt)
    uint32_t first;     // a batch graph's lowest pivot of its last election
    int keep_colour;    // the step from no pivots: its election keeps the colours (the first
                        // election, scc.cu:235-245, elects before any split)
    int32_t *q0, *q1, *q2, *q3, *l0, *l1, *piv;  // queues, live lists, pivot list
    int32_t vbase;       // the vertex the slices start at (a batch graph's first), else 0
    uint32_t max_steps;  // bounds the steps, and each step's levels and trim rounds
    __device__ int32_t *list(int k) const { return k ? l1 : l0; }
    __device__ uint32_t lcnt(int k) const { return k ? lc1 : lc0; }
    __device__ void set_lcnt(int k, uint32_t v) {
        if (k) lc1 = v;
        else lc0 = v;
    }
};

// The election through the global slot tables (the multi-kernel path's):
// prev = slots[pt] holds the last election's keys, cur = slots[pt ^ 1] this one's.
struct GlobalSlots {
    __device__ void clear_cur(const TailArgs &, int, uint32_t) const {}
    __device__ uint32_t prev_winner(const TailArgs &t, int pt, uint32_t s) const { return (uint32_t)t.slots[pt][s]; }
    __device__ void offer(const TailArgs &t, int pt, uint32_t s, unsigned long long key) const {
        // read the slot first: most vertices of a colour lose, and their atomics
        // would queue at one address
        unsigned long long *slot = &t.slots[pt ^ 1][s];
        if (key < __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) atomicMin(slot, key);
    }
    __device__ uint32_t winner(const TailArgs &t, int pt, uint32_t s) const {
        return (uint32_t)__hip_atomic_load(&t.slots[pt ^ 1][s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
};

constexpr uint32_t TAB_EMPTY = 0xFFFFFFFFu;  // no slot id (slot ids are below 2^24)

// The election of one graph of a batch through tables of its own in global
// memory: open addressing keyed by slot id, H entries (a power of two above 1.5
// times the graph's vertices, so every distinct slot has a place) of (slot id,
// lowest id), cleared each step.  Colours of one graph collide only with each
// other, as they would if the graph were solved alone.
struct GraphSlots {
    uint32_t *tab[2];  // 2H words each: slot ids, then lowest ids
    uint32_t H;
    __device__ uint32_t *tb(int k) const { return (k & 1) ? tab[1] : tab[0]; }  // a select, not an indexed load
    __device__ void clear_cur(const TailArgs &, int pt, uint32_t tid) const {
        uint32_t *e = tb(pt ^ 1);
        for (uint32_t i = tid; i < 2u * H; i += TAIL_BLOCK) e[i] = TAB_EMPTY;
    }
    __device__ uint32_t find(int k, uint32_t s) const {
        const uint32_t *e = tb(k);
        for (uint32_t j = 0, h = s & (H - 1); j < H; j++, h = (h + 1) & (H - 1)) {
            const uint32_t x = __hip_atomic_load(&e[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (x == s) return __hip_atomic_load(&e[H + h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (x == TAB_EMPTY) break;
        }
        return TAB_EMPTY;
    }
    __device__ void insert(int k, uint32_t s, uint32_t g) const {
        uint32_t *e = tb(k);
        for (uint32_t j = 0, h = s & (H - 1); j < H; j++, h = (h + 1) & (H - 1)) {  // H > distinct slots
            uint32_t x = __hip_atomic_load(&e[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (x == TAB_EMPTY) x = atomicCAS(&e[h], TAB_EMPTY, s);
            if (x == TAB_EMPTY || x == s) {
                if (g < __hip_atomic_load(&e[H + h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
                    atomicMin(&e[H + h], g);
                return;
            }
        }
    }
    __device__ uint32_t prev_winner(const TailArgs &, int pt, uint32_t s) const { return find(pt, s); }
    __device__ void offer(const TailArgs &, int pt, uint32_t s, unsigned long long key) const {
        insert(pt ^ 1, s, (uint32_t)key);
    }
    __device__ uint32_t winner(const TailArgs &, int pt, uint32_t s) const { return find(pt ^ 1, s); }
};

// Vertices are global ids; the trim and update go over the live list of the
// last election, the confirm over the list the update writes.
template <class Slots>
struct GlobalSpaceT : Slots {
    const TailArgs *t;
    const TailRun *r;
    const int32_t *lst;
    uint32_t lc;
    int32_t *out;
    __device__ uint32_t size(int) const { return lc; }
    __device__ int32_t item(int, uint32_t i) const { return lst[i]; }
    __device__ uint32_t csize(uint32_t ln) const { return ln; }
    __device__ int32_t citem(int, uint32_t i) const { return out[i]; }
    __device__ void keep(int, uint32_t k, int32_t v) const { out[k] = v; }
    __device__ void kept(int, uint32_t, uint32_t) const {}  // the host reads the length from the control block
    __device__ uint32_t ld(int32_t v) const { return t->word[v]; }
    __device__ void st(int32_t v, uint32_t w) const { t->word[v] = w; }
    __device__ uint32_t orw(int32_t v, uint32_t b) const { return atomicOr(&t->word[v], b); }
    __device__ int32_t gid(int32_t v) const { return v; }
    __device__ int32_t *qptr(int b) const {  // select, not an indexed load of the pointer
        int32_t *q0 = r->q0, *q1 = r->q1, *q2 = r->q2, *q3 = r->q3;
        return b == 0 ? q0 : b == 1 ? q1 : b == 2 ? q2 : q3;
    }
    __device__ int32_t qget(int b, uint32_t k) const { return qptr(b)[k]; }
    __device__ void qput(int b, uint32_t k, int32_t v) const { qptr(b)[k] = v; }
    __device__ void start(uint32_t np, uint32_t tid) const {
        for (uint32_t i = tid; i < np; i += TAIL_BLOCK) r->q0[i] = r->q1[i] = r->piv[i];
    }
    __device__ void pivot(uint32_t k, int32_t v) const { r->piv[k] = v; }
    template <class F>
    __device__ bool each_nb(int d, int32_t v, F f) const {  // true once f is
        const int32_t *row = d == 0 ? t->row : t->rowT, *col = d == 0 ? t->col : t->colT;
        for (int32_t e = row[v]; e < row[v + 1]; e++)
            if (f(col[e])) return true;
        return false;
    }
};

using GlobalSpace = GlobalSpaceT<GlobalSlots>;

// Vertices are local indexes 0..L-1 into the LDS copy; the pivot list is the
// first two queues (a step's reach starts from them).  tail_arena holds, for L
// live vertices and `to` out-arcs: words[L], global ids[L], out offsets[L+1],
// in offsets[L+1], four uint16 queues[L], two uint16 live lists[L] and their
// lengths (as the global path, the trim and update go over the list of the last
// update, which writes the other: a step's sweeps cover the vertices still live,
// not all L), out arcs[to], in arcs (uint16), and,
// when there is room (H > 0, from byte t0), two election tables of H entries
// (slot id, lowest global id), open addressing: the slot tables of the steps
// in LDS, cleared every step instead of tagged.  Without the room, the global
// slot tables serve.
template <class Slots>
struct LocalSpaceT : Slots {
    uint32_t L, to, H, t0;
    __device__ uint32_t *tab(int k) const { return (uint32_t *)(tail_arena + t0) + (k & 1) * 2u * H; }
    __device__ void clear_cur(const TailArgs &t, int pt, uint32_t tid) const {
        if (!H) {
            Slots::clear_cur(t, pt, tid);
            return;
        }
        uint32_t *e = tab(pt ^ 1);
        for (uint32_t i = tid; i < 2u * H; i += TAIL_BLOCK) e[i] = TAB_EMPTY;
    }
    __device__ uint32_t find(int k, uint32_t s) const {
        const uint32_t *e = tab(k);
        for (uint32_t j = 0, h = s & (H - 1); j < H; j++, h = (h + 1) & (H - 1)) {
            const uint32_t x = e[h];
            if (x == s) return e[H + h];
            if (x == TAB_EMPTY) break;
        }
        return TAB_EMPTY;
    }
    __device__ void insert(int k, uint32_t s, uint32_t g) const {
        uint32_t *e = tab(k);
        for (uint32_t j = 0, h = s & (H - 1); j < H; j++, h = (h + 1) & (H - 1)) {  // H > distinct slots
            uint32_t x = e[h];
            if (x == TAB_EMPTY) x = atomicCAS(&e[h], TAB_EMPTY, s);
            if (x == TAB_EMPTY || x == s) {
                if (g < e[H + h]) atomicMin(&e[H + h], g);
                return;
            }
        }
    }
    __device__ uint32_t prev_winner(const TailArgs &t, int pt, uint32_t s) const {
        return H ? find(pt, s) : Slots::prev_winner(t, pt, s);
    }
    __device__ void offer(const TailArgs &t, int pt, uint32_t s, unsigned long long key) const {
        if (H)
            insert(pt ^ 1, s, (uint32_t)key);
        else
            Slots::offer(t, pt, s, key);
    }
    __device__ uint32_t winner(const TailArgs &t, int pt, uint32_t s) const {
        return H ? find(pt ^ 1, s) : Slots::winner(t, pt, s);
    }
    __device__ uint32_t *w() const { return (uint32_t *)tail_arena; }
    __device__ int32_t *gidv() const { return (int32_t *)(tail_arena + 4u * L); }
    __device__ uint32_t *off(int d) const { return (uint32_t *)(tail_arena + 8u * L) + (d ? L + 1 : 0); }
    __device__ uint16_t *q(int b) const { return (uint16_t *)(tail_arena + 16u * L + 8u) + b * L; }
    __device__ uint16_t *lst(int k) const { return (uint16_t *)(tail_arena + 24u * L + 8u) + (k ? L : 0); }
    __device__ uint32_t *lcount() const { return (uint32_t *)(tail_arena + 28u * L + 8u); }
    __device__ uint16_t *adj(int d) const { return (uint16_t *)(tail_arena + tail_fixed_bytes(L)) + (d ? to : 0); }
    // ls: the step's live list (TailRun::lsel, the global path's choice)
    __device__ uint32_t size(int ls) const { return lcount()[ls]; }
    __device__ int32_t item(int ls, uint32_t i) const { return (int32_t)lst(ls)[i]; }
    __device__ uint32_t csize(uint32_t ln) const { return ln; }
    __device__ int32_t citem(int ls, uint32_t i) const { return (int32_t)lst(ls ^ 1)[i]; }
    __device__ void keep(int ls, uint32_t k, int32_t v) const { lst(ls ^ 1)[k] = (uint16_t)v; }
    __device__ void kept(int ls, uint32_t ln, uint32_t tid) const {  // read after the confirm's barrier
        if (tid == 0) lcount()[ls ^ 1] = ln;
    }
    __device__ uint32_t ld(int32_t v) const { return w()[v]; }
    __device__ void st(int32_t v, uint32_t x) const { w()[v] = x; }
    __device__ uint32_t orw(int32_t v, uint32_t b) const { return atomicOr(&w()[v], b); }
    __device__ int32_t gid(int32_t v) const { return gidv()[v]; }
    __device__ int32_t qget(int b, uint32_t k) const { return q(b)[k]; }
    __device__ void qput(int b, uint32_t k, int32_t v) const { q(b)[k] = (uint16_t)v; }
    __device__ void start(uint32_t, uint32_t) const {}
    __device__ void pivot(uint32_t k, int32_t v) const { q(0)[k] = q(1)[k] = (uint16_t)v; }
    template <class F>
    __device__ bool each_nb(int d, int32_t v, F f) const {
        const uint32_t *o = off(d);
        const uint16_t *a = adj(d);
        for (uint32_t e = o[v]; e < o[v + 1]; e++)
            if (f((int32_t)a[e])) return true;
        return false;
    }
};


// End-of-step bookkeeping.  The tail of a run publishes each step in the control
// block (the host goes on from it); a graph of a batch keeps its list lengths
// and pivot count in TailRun and adds its classes (trimmed vertices, pivots) to
// the device's count, which other workgroups update too.
struct TailBook {
    __device__ void end(const TailArgs &t, TailShared &sh, TailRun &r, uint32_t ln) const {
        if (threadIdx.x == 0) {
            Ctl *ctl = t.ctl;
            ctl->live -= (uint32_t)sh.red[0];
            ctl->pcount += (uint32_t)sh.red[1] + r.npiv;
            ctl->lcnt[r.lsel ^ 1] = ln;
            ctl->npiv_prev = ctl->npiv;
            ctl->first_piv_prev = ctl->first_piv;
            ctl->npiv = r.npiv;
            ctl->first_piv = sh.first;
        }
    }
};
struct BatchBook {
    __device__ void end(const TailArgs &t, TailShared &sh, TailRun &r, uint32_t ln) const {
        if (threadIdx.x == 0 && sh.red[1] + r.npiv) atomicAdd(&t.ctl->pcount, (uint32_t)sh.red[1] + r.npiv);
        r.set_lcnt(r.lsel ^ 1, ln);
        r.first = sh.first;
    }
};

// One FB step; false when a bound that every correct run stays within is passed.
template <class Sp, class Bk = TailBook>
__device__ bool tail_step(const TailArgs &t, const Sp &sp, TailShared &sh, TailRun &r, const Bk &bk = Bk()) {
    const uint32_t tid = threadIdx.x, nt = TAIL_BLOCK;
    unsigned long long clk = 0;  // phase clocks of an instrumented tail (thread 0)
    auto phase = [&](int q) {
        if (t.instr && tid == 0) {
            const unsigned long long now = wall_clock64();
            if (q >= 0) atomicAdd(&t.ctl->st_tailclk[q], now - clk);
            clk = now;
        }
    };
    phase(-1);
    // ---- reach: F over out-arcs, B over in-arcs, both from the pivot list
    if (tid == 0) {
        sh.cnt[0] = sh.cnt[1] = r.npiv;
        sh.cnt[2] = sh.cnt[3] = 0;
    }
    sp.start(r.npiv, tid);
    sp.clear_cur(t, r.pt, tid);  // this step's election table (the reach does not touch it)
    tail_sync();
    int qi = 0;  // input queues: 2 * qi + d
    for (uint32_t level = 0;; level++) {
        const uint32_t n0 = sh.cnt[0], n1 = sh.cnt[1];
        if (n0 + n1 == 0) break;                     // block-uniform (read after the sync)
        if (level > 2u * r.max_steps) return false;  // every level queues a new vertex
        r.depths++;
        for (uint32_t i = tid; i < n0 + n1; i += nt) {
            const int d = i < n0 ? 0 : 1;
            const int32_t v = sp.qget(2 * qi + d, d == 0 ? i : i - n0);
            const uint32_t c = colour_of(sp.ld(v)), bit = d == 0 ? ST_F : ST_B;
            // the sweep schedules mark every vertex they expand (reach_q.cu:44-79), and an
            // FB-eliminated word keeps the mark: the same words as the multi-kernel path
            if (t.xbits) sp.orw(v, d == 0 ? ST_FX : ST_BX);
            const int ob = 2 * (qi ^ 1) + d;
            sp.each_nb(d, v, [&](int32_t u) {
                const uint32_t wu = sp.ld(u);
                if ((wu & (bit | ST_E)) || colour_of(wu) != c) return false;
                if (sp.orw(u, bit) & bit) return false;
                sp.qput(ob, atomicAdd(&sh.cnt[2 + d], 1u), u);  // each vertex queued once
                return false;
            });
        }
        tail_sync();
        if (tid == 0) {
            sh.cnt[0] = sh.cnt[2];
            sh.cnt[1] = sh.cnt[3];
            sh.cnt[2] = sh.cnt[3] = 0;
        }
        qi ^= 1;
        tail_sync();
    }
    phase(0);
    // ---- trim to the fixpoint (F and B without P: an SCC of two or more); a
    // round that trims nothing ends it (the count is read between two barriers,
    // so every work-item takes the same decision)
    const uint32_t lc = sp.size(r.lsel);
    if (tid < 2) sh.red[tid] = 0;
    tail_sync();
    for (uint32_t it = 0, seen = 0;; it++) {
        for (uint32_t i = tid; i < lc; i += nt) {
            const int32_t v = sp.item(r.lsel, i);
            const uint32_t w = sp.ld(v);
            if ((w & ST_E) || (w & (ST_P | ST_F | ST_B)) == (ST_F | ST_B)) continue;
            const uint32_t c = colour_of(w);
            auto live = [&](int32_t u) {
                if (u == v) return false;
                const uint32_t wu = sp.ld(u);
                return !(wu & ST_E) && colour_of(wu) == c;
            };
            if (sp.each_nb(0, v, live) && sp.each_nb(1, v, live)) continue;
            sp.st(v, make_word(c, 28));  // P|T|E (scc_routines.cu:381)
            const int32_t g = sp.gid(v);
            t.root[g] = orig_id(t.oid, g);
            atomicAdd(&sh.red[0], 1ull);
            if (!(w & ST_P)) atomicAdd(&sh.red[1], 1ull);  // a trimmed pivot holds P already
        }
        tail_sync();
        const uint32_t now = (uint32_t)sh.red[0];
        tail_sync();
        r.trims++;
        if (now == seen) break;  // block-uniform
        seen = now;
        if (it > r.max_steps) return false;  // every round but the last trims a vertex
    }
    phase(1);
    // ---- update: eliminate F and B into the pivot's SCC, split the rest, elect
    r.epoch++;
    const unsigned long long tag = elect_tag(r.epoch);
    if (tid == 0) {
        sh.lcnt = 0;
        sh.npiv = 0;
        sh.first = 0xFFFFFFFFu;
    }
    tail_sync();
    for (uint32_t i = tid; i < lc; i += nt) {
        const int32_t v = sp.item(r.lsel, i);
        const uint32_t w = sp.ld(v);
        if (w & ST_E) continue;
        uint32_t c = colour_of(w);
        const uint32_t f = ((w & ST_F) ? 1u : 0u) + ((w & ST_B) ? 2u : 0u);
        const int32_t g = sp.gid(v), og = orig_id(t.oid, g);  // keys and roots: the caller's ids
        if (f == 3) {
            sp.st(v, w | ST_E);
            const uint32_t pv = sp.prev_winner(t, r.pt, slot_of(c, t.slot_bits));
            t.root[g] = pv < (uint32_t)t.n ? (int32_t)pv : og;
            if (t.mn && pv < (uint32_t)t.n && og < t.mn[pv]) atomicMin(&t.mn[pv], og);
            atomicAdd(&sh.red[0], 1ull);
            continue;
        }
        if (!r.keep_colour) c = (3u * c + f) & COLOUR_MASK;
        sp.st(v, make_word(c, 0));
        sp.keep(r.lsel, atomicAdd(&sh.lcnt, 1u), v);
        sp.offer(t, r.pt, slot_of(c, t.slot_bits), tag | (uint32_t)og);
    }
    tail_sync();
    phase(2);
    // ---- confirm: the lowest key of each slot becomes its colour's pivot (state 19)
    const uint32_t ln = sh.lcnt, cn = sp.csize(ln);
    sp.kept(r.lsel, ln, tid);
    for (uint32_t i = tid; i < cn; i += nt) {
        const int32_t v = sp.citem(r.lsel, i);
        const uint32_t w = sp.ld(v);
        if (w & ST_E) continue;
        const uint32_t c = colour_of(w);
        const uint32_t og = (uint32_t)orig_id(t.oid, sp.gid(v));
        const uint32_t win = sp.winner(t, r.pt, slot_of(c, t.slot_bits));
        if (win != og) continue;
        sp.st(v, make_word(c, 19));
        sp.pivot(atomicAdd(&sh.npiv, 1u), v);
        atomicMin(&sh.first, og);
    }
    tail_sync();
    phase(3);
    r.npiv = sh.npiv;
    bk.end(t, sh, r, ln);  // the counts the host reads at the end
    r.pt ^= 1;
    r.lsel ^= 1;
    tail_sync();
    return true;
}

__device__ inline bool tail_arc(const TailArgs &t, int32_t v, int32_t u, uint32_t c) {
    if (u == v) return false;
    const uint32_t wu = t.word[u];
    return !(wu & ST_E) && colour_of(wu) == c;
}

// Copies the L live vertices of list lst, and the arcs between live vertices
// of one colour, into LDS.  1: copied; 0: the arcs do not fit; -1: a live
// vertex outside the list, which the elections never leave.
template <class LS>
__device__ int tail_build(const TailArgs &t, const TailRun &r, TailShared &sh, LS &ls, uint32_t L,
                          const int32_t *lst, uint32_t npiv, int pt) {
    const uint32_t tid = threadIdx.x, nt = TAIL_BLOCK;
    ls.L = L;
    ls.to = 0;
    ls.H = 0;
    {  // the live list of the step: every vertex of the copy
        uint16_t *l = ls.lst(r.lsel);
        for (uint32_t i = tid; i < L; i += nt) l[i] = (uint16_t)i;
        if (tid == 0) ls.lcount()[r.lsel] = L;
    }
    uint32_t *lw = ls.w(), *off0 = ls.off(0), *off1 = ls.off(1);
    int32_t *gidv = ls.gidv();
    // global id -> local index of the live vertices, indexed from r.vbase (the
    // first vertex of a batch graph, whose queue slice this is)
    uint32_t *map = (uint32_t *)r.q3 - r.vbase;
    for (uint32_t i = tid; i < L; i += nt) {
        const int32_t v = lst[i];
        gidv[i] = v;
        lw[i] = t.word[v];
        map[v] = i;
    }
    if (tid == 0) sh.flag = 0;
    tail_sync();
    // each work-item counts a run of consecutive vertices, so one scan of the
    // per-work-item sums places every row
    const uint32_t per = (L + nt - 1) / nt, b = min(tid * per, L), e = min(b + per, L);
    uint32_t so = 0, si = 0;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t w = lw[i], c = colour_of(w);
        const int32_t v = gidv[i];
        uint32_t k[2] = {0, 0};
        if (!(w & ST_E)) {
            for (int d = 0; d < 2; d++) {
                const int32_t *row = d == 0 ? t.row : t.rowT, *col = d == 0 ? t.col : t.colT;
                for (int32_t x = row[v]; x < row[v + 1]; x++) {
                    const int32_t u = col[x];
                    if (!tail_arc(t, v, u, c)) continue;
                    const uint32_t j = map[u];
                    if (j >= L || gidv[j] != u) sh.flag = 1;
                    k[d]++;
                }
            }
        }
        off0[i] = k[0];
        off1[i] = k[1];
        so += k[0];
        si += k[1];
    }
    uint32_t to = 0, ti = 0;
    const uint32_t bo = block_excl_scan(so, sh.scan, to, nt);
    const uint32_t bi = block_excl_scan(si, sh.scan, ti, nt);
    tail_sync();
    if (sh.flag) return -1;                                             // block-uniform
    if (tail_fixed_bytes(L) + 2ull * (to + ti) > TAIL_ARENA) return 0;  // block-uniform
    ls.to = to;
    // election tables when they fit: H a power of two above 1.5 L (at most L distinct slots)
    ls.t0 = (tail_fixed_bytes(L) + 2u * (to + ti) + 15u) & ~15u;
    uint32_t H = 64;
    while (H < L + L / 2 + 1) H <<= 1;
    if (ls.t0 + 16ull * H <= TAIL_ARENA) ls.H = H;
    uint16_t *adj0 = ls.adj(0), *adj1 = ls.adj(1), *q0 = ls.q(0), *q1 = ls.q(1);
    uint32_t po = bo, pi = bi;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t ko = off0[i], ki = off1[i];
        off0[i] = po;
        off1[i] = pi;
        const uint32_t w = lw[i], c = colour_of(w);
        const int32_t v = gidv[i];
        if (ko + ki) {
            for (int d = 0; d < 2; d++) {
                const int32_t *row = d == 0 ? t.row : t.rowT, *col = d == 0 ? t.col : t.colT;
                uint16_t *a = d == 0 ? adj0 : adj1;
                uint32_t p = d == 0 ? po : pi;
                for (int32_t x = row[v]; x < row[v + 1]; x++) {
                    const int32_t u = col[x];
                    if (tail_arc(t, v, u, c)) a[p++] = (uint16_t)map[u];
                }
            }
        }
        po += ko;
        pi += ki;
    }
    if (tid == 0) {
        off0[L] = to;
        off1[L] = ti;
    }
    for (uint32_t k = tid; k < npiv; k += nt) {
        const uint32_t j = map[r.piv[k]];
        q0[k] = q1[k] = (uint16_t)j;
        if (j >= L || gidv[j] != r.piv[k]) sh.flag = 1;
    }
    if (ls.H) ls.clear_cur(t, pt ^ 1, tid);  // the LDS table of the last election: its winners are the pivots
    tail_sync();
    if (sh.flag) return -1;
    if (ls.H)
        for (uint32_t k = tid; k < npiv; k += nt)
            ls.insert(pt, slot_of(colour_of(lw[q0[k]]), t.slot_bits), (uint32_t)orig_id(t.oid, gidv[q0[k]]));
    tail_sync();
    return 1;
}

// Words, the live list (lsel) and the pivot list back to global memory, and,
// when the election tables were in LDS, the keys of the last election's
// winners into its global table (the host's next election reads them there).
using LocalSpace = LocalSpaceT<GlobalSlots>;

__device__ void tail_writeback(const TailArgs &t, const TailRun &r, const LocalSpace &ls, TailShared &sh, int lsel,
                               uint32_t npiv, int pt, uint32_t epoch) {
    const uint32_t tid = threadIdx.x, nt = TAIL_BLOCK;
    if (tid == 0) sh.lcnt = 0;
    tail_sync();
    int32_t *out = r.list(lsel);
    const uint32_t *lw = ls.w();
    const int32_t *gidv = ls.gidv();
    const uint16_t *q0 = ls.q(0);
    for (uint32_t i = tid; i < ls.L; i += nt) {
        const uint32_t w = lw[i];
        const int32_t v = gidv[i];
        t.word[v] = w;
        if (!(w & ST_E)) out[atomicAdd(&sh.lcnt, 1u)] = v;
    }
    for (uint32_t k = tid; k < npiv; k += nt) {
        const int32_t g = gidv[q0[k]];
        r.piv[k] = g;
        if (ls.H)
            t.slots[pt][slot_of(colour_of(lw[q0[k]]), t.slot_bits)] = elect_tag(epoch) | (uint32_t)orig_id(t.oid, g);
    }
    tail_sync();
    if (tid == 0) t.ctl->lcnt[lsel] = sh.lcnt;
}

__global__ __launch_bounds__(TAIL_BLOCK) void k_fb_tail(TailArgs t) {
    __shared__ TailShared sh;
    Ctl *ctl = t.ctl;
    const uint32_t live0 = ctl->live;  // block-uniform: no work-item writes ctl before a barrier
    if (t.gate_live && live0 > t.gate_live) {  // launched without a host wait: too many live vertices
        if (threadIdx.x == 0) {
            ctl->tail[0] = 0;
            ctl->tail[1] = 3;
            ctl->tail[2] = ctl->tail[3] = 0;
            ctl->tail_pivots = 0;
        }
        return;
    }
    if (t.instr && threadIdx.x == 0) {  // this launch's phase clocks (only thread 0 touches them)
        for (int q = 0; q < 4; q++) ctl->st_tailclk[q] = 0;
        ctl->tail_L = ctl->tail_lstep = 0;
    }
    TailRun r{};
    r.lsel = t.lsel;
    r.pt = t.prev_tab;
    r.epoch = t.epoch;
    r.npiv = ctl->npiv;
    r.q0 = t.q[0];
    r.q1 = t.q[1];
    r.q2 = t.q[2];
    r.q3 = t.q[3];
    r.l0 = t.list[0];
    r.l1 = t.list[1];
    r.piv = t.piv;
    r.vbase = t.vbase;
    r.max_steps = t.max_steps ? t.max_steps : live0 + 2;  // each step eliminates its pivots
    uint32_t steps = 0, status = 0, try_below = 0xFFFFFFFFu;
    unsigned long long pivots = 0;
    bool local = false;
    LocalSpace ls{};
    // first_step: a step from no pivots is trim-1 (every vertex checked, none holds F
    // or B) and an election whose colour split 3c + 0 keeps one colour per graph
    bool first = t.first_step != 0;
    while (r.npiv > 0 || first) {
        if (steps >= r.max_steps) {
            status = 1;
            break;
        }
        if (((r.epoch + 1u) & 0xFFFFu) < 2u) {  // the next election's tag wraps
            if (!t.self_refill) {                // the host takes over
                status = 1;
                break;
            }
            // clear the table this step's election writes: no key of an election 2^16
            // earlier may outrank the new tags
            unsigned long long *cur = r.pt ? t.slots[0] : t.slots[1];
            for (uint32_t i = threadIdx.x; i < (1u << t.slot_bits); i += TAIL_BLOCK) cur[i] = ~0ull;
            tail_sync();
        }
        r.keep_colour = first ? 1 : 0;
        first = false;
        if (!local) {  // the LDS copy, once the live list is short enough and its arcs fit
            const uint32_t L = ctl->lcnt[r.lsel];
            if (L < try_below && L < 65536u && tail_fixed_bytes(L) <= TAIL_ARENA) {
                const int b = tail_build(t, r, sh, ls, L, r.list(r.lsel), r.npiv, r.pt);
                if (b < 0) {
                    status = 2;
                    break;
                }
                local = b > 0;
                try_below = L / 2;
                if (local && t.instr && threadIdx.x == 0) {
                    ctl->tail_L = L;
                    ctl->tail_lstep = steps;
                }
            }
        }
        steps++;
        pivots += r.npiv;  // the pivots this step reaches from
        bool ok;
        if (local) {
            ok = tail_step(t, ls, sh, r);
        } else {
            const GlobalSpace gs{{}, &t, &r, r.list(r.lsel), ctl->lcnt[r.lsel], r.list(r.lsel ^ 1)};
            ok = tail_step(t, gs, sh, r);
        }
        if (!ok) {
            status = 2;
            break;
        }
    }
    if (local) tail_writeback(t, r, ls, sh, r.lsel, r.npiv, r.pt, r.epoch);
    if (threadIdx.x == 0) {
        ctl->tail_live0 = live0;
        ctl->tail[0] = steps;
        ctl->tail[1] = status;
        ctl->tail[2] = (uint32_t)r.depths;
        ctl->tail[3] = (uint32_t)r.trims;
        ctl->tail_pivots = pivots;
    }
}

void launch_fb_tail(hipStream_t s, const TailArgs &t) {
    hipLaunchKernelGGL(k_fb_tail, dim3(1), dim3(TAIL_BLOCK), 0, s, t);
}

// ---------------------------------------------------------------------------
// A batch of graphs in one persistent launch.  Each workgroup takes graphs from
// a work counter, largest first, and solves each one alone with the steps of
// the tail: set-up (colour g + 1, roots = self, the live list = the graph),
// a first step from no pivots (trim-1 and the first election, lowest id),
// then reach, trim, update and election until no pivot is left, in LDS once
// the graph's live part fits there.  Every per-vertex buffer is used through
// the graph's slice [o0, o1) (a graph's queues and lists never hold more than
// its vertices), and its elections go through tables of its own, so graphs
// never meet: each ends exactly as it would alone (same colours, states,
// roots and steps).  Statistics: steps, levels and trim rounds are maxima over
// the graphs, pivots a sum.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t batch_table_size(int32_t ng) {
    uint32_t H = 64;
    while (H < (uint32_t)ng + (uint32_t)ng / 2u + 1u) H <<= 1;
    return H;
}

__global__ __launch_bounds__(TAIL_BLOCK) void k_fb_batch(TailArgs base, BatchArgs b) {
    __shared__ TailShared sh;
    __shared__ uint32_t s_next;
    const uint32_t tid = threadIdx.x;
    Ctl *ctl = base.ctl;
    uint32_t max_steps = 0, max_depth = 0, max_trims = 0, status = 0;
    unsigned long long pivots = 0;
    for (;;) {
        if (tid == 0) s_next = atomicAdd(&ctl->wnext, 1u);
        tail_sync();
        const uint32_t k = s_next;
        tail_sync();  // s_next is written again for the next graph
        if (k >= (uint32_t)b.G) break;  // block-uniform
        const int32_t g = b.order[k];
        const int32_t o0 = b.off[g], ng = b.off[g + 1] - o0;
        if (ng <= 0) continue;
        const TailArgs &t = base;
        TailRun r{};
        r.q0 = base.q[0] + o0;
        r.q1 = base.q[1] + o0;
        r.q2 = base.q[2] + o0;
        r.q3 = base.q[3] + o0;
        r.l0 = base.list[0] + o0;
        r.l1 = base.list[1] + o0;
        r.piv = base.piv + o0;
        r.vbase = o0;
        r.max_steps = (uint32_t)ng + 2u;
        r.lc0 = (uint32_t)ng;
        const uint32_t colour = (uint32_t)(g + 1) & COLOUR_MASK;
        for (int32_t i = (int32_t)tid; i < ng; i += TAIL_BLOCK) {
            const int32_t v = o0 + i;
            t.word[v] = make_word(colour, 0);
            t.root[v] = orig_id(t.oid, v);
            r.l0[i] = v;
        }
        GraphSlots gs;
        gs.H = batch_table_size(ng);
        gs.tab[0] = b.tab + b.tab_off[g];
        gs.tab[1] = gs.tab[0] + 2u * gs.H;
        LocalSpaceT<GraphSlots> ls{};
        static_cast<GraphSlots &>(ls) = gs;
        bool local = false, first = true;
        uint32_t steps = 0, try_below = 0xFFFFFFFFu;
        tail_sync();
        while (r.npiv > 0 || first) {  // block-uniform
            if (steps >= r.max_steps) {
                status = 2;
                break;
            }
            if (!local) {  // the LDS copy, once the live list is short enough and its arcs fit
                const uint32_t L = r.lcnt(r.lsel);
                if (L < try_below && L < 65536u && tail_fixed_bytes(L) <= TAIL_ARENA) {
                    const int bld = tail_build(t, r, sh, ls, L, r.list(r.lsel), r.npiv, r.pt);
                    if (bld < 0) {
                        status = 2;
                        break;
                    }
                    local = bld > 0;
                    try_below = L / 2;
                }
            }
            r.keep_colour = first ? 1 : 0;
            first = false;
            steps++;
            pivots += r.npiv;  // the pivots this step reaches from
            bool ok;
            if (local) {
                ok = tail_step(t, ls, sh, r, BatchBook());
            } else {
                const GlobalSpaceT<GraphSlots> sp{gs, &t, &r, r.list(r.lsel), r.lcnt(r.lsel), r.list(r.lsel ^ 1)};
                ok = tail_step(t, sp, sh, r, BatchBook());
            }
            if (!ok) {
                status = 2;
                break;
            }
            if (steps == 1 && r.npiv && tid == 0) atomicMin(&ctl->first_piv, r.first);
        }
        if (local) {  // the words back to global memory
            const uint32_t *lw = ls.w();
            const int32_t *gidv = ls.gidv();
            for (uint32_t i = tid; i < ls.L; i += TAIL_BLOCK) t.word[gidv[i]] = lw[i];
        }
        tail_sync();  // LDS is rebuilt for the next graph
        // FB steps: the first step only trims and elects
        max_steps = steps - 1 > max_steps ? steps - 1 : max_steps;
        max_depth = (uint32_t)r.depths > max_depth ? (uint32_t)r.depths : max_depth;
        max_trims = (uint32_t)r.trims > max_trims ? (uint32_t)r.trims : max_trims;
        if (status) break;
    }
    if (tid == 0) {
        atomicMax(&ctl->tail[0], max_steps);
        if (status) atomicMax(&ctl->tail[1], status);
        atomicMax(&ctl->tail[2], max_depth);
        atomicMax(&ctl->tail[3], max_trims);
        atomicAdd(&ctl->tail_pivots, pivots);
    }
}

uint32_t batch_table_words(int32_t n_g) { return 4u * batch_table_size(n_g); }

void launch_fb_batch(hipStream_t s, uint32_t grid, const TailArgs &t, const BatchArgs &b) {
    hipLaunchKernelGGL(k_fb_batch, dim3(grid), dim3(TAIL_BLOCK), 0, s, t, b);
}

hipError_t preload_tail() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(k_fb_tail));
}

}  // namespace gpuscc
