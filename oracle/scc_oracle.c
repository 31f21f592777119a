/*
 * scc_oracle.c -- CPU restatement of the reference's Forward-Backward SCC path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * `scc` CLI) links, loads or calls this file.  It is used by tests/ as the
 * checker, by __graft_entry__.smoke() to check one small GPU run, and by
 * bench.py's cpu_baseline leg (single-threaded Tarjan timing).
 *
 * Every function cites the reference code it restates (file:line in
 * chenxuhao/gpuscc_code).  Vertex ids here are 0-based; the reference uses
 * 1-based ids with index 0 unused (load.cpp:128-139), which changes nothing
 * but an offset.
 *
 * State byte (scc.h:48-93): B=1 F=2 E=4 T=8 P=16 BExt=32 FExt=64.
 * Colours ("ranges", scc_routines.cu:207-209): c <- 3c + [F] + 2[B], wrapped
 * to `colour_mask` (0xFFFFFFFF restates the reference exactly; 0x00FFFFFF is
 * the packed 24-bit colour the HIP path keeps in one word with the state).
 *
 * Pivot rule: the reference elects, per hash slot, whichever vertex wins an
 * atomicCAS race (scc_routines.cu:214-221); the first pivot is the lowest
 * live id after trim-1 (scc.cu:235-245).  The race's outcome is not fixed,
 * so this restatement resolves every election to the lowest live id in the
 * slot, which is one of the outcomes the reference allows and the one the
 * HIP path computes deterministically (atomicMin).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

enum { ST_B = 1, ST_F = 2, ST_E = 4, ST_T = 8, ST_P = 16, ST_BX = 32, ST_FX = 64 };

typedef struct {
    int64_t steps;         /* FB loop iterations, scc.cu:266,279 */
    int64_t num_scc;       /* popcount of the P bit, scc.cu:914-919 */
    int64_t pivots_total;  /* pivots elected over the run (incl. the first) */
    int64_t wcc_step;      /* step at which the one WCC pass ran (0 = never) */
    int64_t trimmed;       /* vertices removed by trim-1 (state 28) */
    int64_t trim2_pairs;   /* pairs removed by the one trim-2 pass */
} orc_fb_stats;

/* ------------------------------------------------------------------ */
/* hash of a colour to an election slot                                */
/* kind 0: the reference's `colour & PIVOT_HASH_CONST` (scc.h:4,        */
/*         scc_routines.cu:214) generalised to `slot_bits` low bits.    */
/* kind 1: multiplicative (Fibonacci) hash, top `slot_bits` bits of     */
/*         colour * 0x9E3779B1 mod 2^32 -- what the HIP path uses.      */
/* ------------------------------------------------------------------ */
static inline uint32_t orc_slot(uint32_t c, int slot_bits, int kind) {
    if (kind == 0) return c & ((slot_bits >= 32) ? 0xFFFFFFFFu : ((1u << slot_bits) - 1u));
    if (slot_bits == 0) return 0;
    return (uint32_t)(c * 0x9E3779B1u) >> (32 - slot_bits);
}

uint32_t orc_hash_slot(uint32_t c, int slot_bits, int kind) { return orc_slot(c, slot_bits, kind); }

/* ------------------------------------------------------------------ */
/* Tarjan: the role of TR_Decomposition (tarjan.cpp:72-240).  Textbook */
/* iterative Tarjan (1972) with an explicit (vertex, next-edge) frame   */
/* stack.  root_out[v] = the DFS root of v's SCC (as scc_root in        */
/* tarjan.cpp:127); stats follow tarjan.cpp:137-152,237.  Returns the   */
/* number of SCCs; *seconds = DFS time only (tarjan.cpp:108->192).      */
/* ------------------------------------------------------------------ */
int64_t orc_tarjan(int32_t n, const int32_t *row, const int32_t *col, int32_t *root_out,
                   int64_t *trivial, int64_t *nontrivial, int64_t *biggest, double *seconds) {
    int32_t *index = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *low = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *sstack = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *fv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *fe = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    uint8_t *onst = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    int64_t count = 0, triv = 0, nontriv = 0, big = 0;
    int32_t counter = 0, sp = 0, fp = 0;
    struct timespec t0, t1;
    for (int32_t v = 0; v < n; v++) index[v] = -1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int32_t s = 0; s < n; s++) {
        if (index[s] != -1) continue;
        index[s] = low[s] = counter++;
        sstack[sp++] = s; onst[s] = 1;
        fv[fp] = s; fe[fp] = row[s]; fp++;
        while (fp > 0) {
            int32_t v = fv[fp - 1];
            int32_t e = fe[fp - 1];
            if (e < row[v + 1]) {
                fe[fp - 1] = e + 1;
                int32_t w = col[e];
                if (index[w] == -1) {
                    index[w] = low[w] = counter++;
                    sstack[sp++] = w; onst[w] = 1;
                    fv[fp] = w; fe[fp] = row[w]; fp++;
                } else if (onst[w] && index[w] < low[v]) {
                    low[v] = index[w];
                }
            } else {
                fp--;
                if (low[v] == index[v]) {
                    int64_t size = 0;
                    int32_t x;
                    do {
                        x = sstack[--sp];
                        onst[x] = 0;
                        root_out[x] = v;
                        size++;
                    } while (x != v);
                    count++;
                    if (size > big) big = size;
                    if (size == 1) triv++; else nontriv++;
                }
                if (fp > 0) {
                    int32_t u = fv[fp - 1];
                    if (low[v] < low[u]) low[u] = low[v];
                }
            }
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    if (trivial) *trivial = triv;
    if (nontrivial) *nontrivial = nontriv;
    if (biggest) *biggest = big;
    free(index); free(low); free(sstack); free(fv); free(fe); free(onst);
    return count;
}

/* Canonical labelling: label[v] = min vertex id of the class root[v]. */
void orc_canonicalize(int32_t n, const int32_t *root, int32_t *label) {
    int32_t *mn = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t v = 0; v < n; v++) mn[v] = INT32_MAX;
    for (int32_t v = 0; v < n; v++) if (v < mn[root[v]]) mn[root[v]] = v;
    for (int32_t v = 0; v < n; v++) label[v] = mn[root[v]];
    free(mn);
}

/* Brute force for tiny graphs: v and w share an SCC iff each reaches the */
/* other.  O(n*(n+m)) BFS closure; label = min id of the class.           */
void orc_brute_scc(int32_t n, const int32_t *row, const int32_t *col, int32_t *label) {
    uint8_t *reach = (uint8_t *)calloc((size_t)n * (size_t)n + 1, 1);
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t s = 0; s < n; s++) {
        uint8_t *r = reach + (size_t)s * (size_t)n;
        int32_t h = 0, t = 0;
        r[s] = 1; q[t++] = s;
        while (h < t) {
            int32_t v = q[h++];
            for (int32_t e = row[v]; e < row[v + 1]; e++) {
                int32_t w = col[e];
                if (!r[w]) { r[w] = 1; q[t++] = w; }
            }
        }
    }
    for (int32_t v = 0; v < n; v++) {
        label[v] = v;
        for (int32_t w = 0; w < v; w++) {
            if (reach[(size_t)v * n + w] && reach[(size_t)w * n + v]) { label[v] = w; break; }
        }
    }
    free(reach); free(q);
}

/* ------------------------------------------------------------------ */
/* Phase restatements.  All act in place on colour[] / state[] / root[] */
/* ------------------------------------------------------------------ */

/* Trim-1 within colours, run to its fixpoint: trim1_kernel / trim_kernel */
/* (scc_routines.cu:276-394) repeated by trim()/trim1() (:584-643).  A    */
/* live vertex with no live same-colour out-neighbour other than itself, */
/* or no such in-neighbour, becomes state 28 (P|T|E).  The fixpoint is   */
/* unique, so it is computed here with degree counters and a queue.      */
/* colour_blind != 0 drops the colour test (trim1, step 1).  Returns the */
/* number of vertices trimmed.                                           */
int64_t orc_trim(int32_t n, const int32_t *row, const int32_t *col, const int32_t *rowT,
                 const int32_t *colT, const uint32_t *colour, uint8_t *state, int32_t *root,
                 int colour_blind) {
    int32_t *od = (int32_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    int32_t *id = (int32_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    uint8_t *inq = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    int32_t h = 0, t = 0;
    int64_t trimmed = 0;
#define SAMEC(a, b) (colour_blind || colour[a] == colour[b])
    for (int32_t v = 0; v < n; v++) {
        if (state[v] & ST_E) continue;
        for (int32_t e = row[v]; e < row[v + 1]; e++) {
            int32_t w = col[e];
            if (w != v && !(state[w] & ST_E) && SAMEC(v, w)) od[v]++;
        }
        for (int32_t e = rowT[v]; e < rowT[v + 1]; e++) {
            int32_t w = colT[e];
            if (w != v && !(state[w] & ST_E) && SAMEC(v, w)) id[v]++;
        }
        if (od[v] == 0 || id[v] == 0) { q[t++] = v; inq[v] = 1; }
    }
    while (h < t) {
        int32_t v = q[h++];
        state[v] = 28;
        root[v] = v;
        trimmed++;
        for (int32_t e = row[v]; e < row[v + 1]; e++) {
            int32_t w = col[e];
            if (w == v || (state[w] & ST_E) || !SAMEC(v, w) || inq[w]) continue;
            if (--id[w] == 0) { q[t++] = w; inq[w] = 1; }
        }
        for (int32_t e = rowT[v]; e < rowT[v + 1]; e++) {
            int32_t w = colT[e];
            if (w == v || (state[w] & ST_E) || !SAMEC(v, w) || inq[w]) continue;
            if (--od[w] == 0) { q[t++] = w; inq[w] = 1; }
        }
    }
#undef SAMEC
    free(od); free(id); free(q); free(inq);
    return trimmed;
}

/* Reachability closure: fwd_step/bwd_step (reach_q.cu:12-79) repeated by */
/* fwd_reach_q/bwd_reach_q (:81-157).  Every live vertex holding `bit`    */
/* (F=2 over the CSR, B=1 over CSR^T) spreads it to live same-colour      */
/* neighbours.  The closure is schedule-independent, so a BFS computes it.*/
/* Returns the number of vertices that newly received the bit.           */
int64_t orc_reach(int32_t n, const int32_t *row, const int32_t *col, const uint32_t *colour,
                  uint8_t *state, int bit) {
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t h = 0, t = 0;
    int64_t added = 0;
    for (int32_t v = 0; v < n; v++)
        if ((state[v] & bit) && !(state[v] & ST_E)) q[t++] = v;
    while (h < t) {
        int32_t v = q[h++];
        for (int32_t e = row[v]; e < row[v + 1]; e++) {
            int32_t w = col[e];
            if ((state[w] & (bit | ST_E)) == 0 && colour[w] == colour[v]) {
                state[w] |= (uint8_t)bit;
                q[t++] = w;
                added++;
            }
        }
    }
    free(q);
    return added;
}

/* Colour update and pivot election: update_kernel (scc_routines.cu:187-224). */
/* Live v with F and B joins the pivot's SCC (E set; root = the pivot elected  */
/* for its colour in the previous election, prev_winner[slot]).  Other live v  */
/* get colour 3c+[F]+2[B] (masked), state 0, and compete for slot(colour):     */
/* lowest id wins and becomes state 19 (P|F|B).  recolour == 0 skips the split */
/* (the first election after trim-1, scc.cu:235-245).  winner[] (size         */
/* 2^slot_bits) receives the winners, -1 for empty slots.  Returns #pivots.   */
int64_t orc_update_elect(int32_t n, uint32_t *colour, uint8_t *state, int32_t *root,
                         uint32_t colour_mask, int slot_bits, int hash_kind,
                         const int32_t *prev_winner, int32_t *winner, int recolour) {
    size_t nslots = (size_t)1 << slot_bits;
    int64_t npiv = 0;
    for (size_t s = 0; s < nslots; s++) winner[s] = -1;
    for (int32_t v = 0; v < n; v++) {
        uint8_t s = state[v];
        if (s & ST_E) continue;
        if (recolour) {
            uint32_t k = ((s & ST_F) ? 1u : 0u) + ((s & ST_B) ? 2u : 0u);
            if (k == 3) {
                state[v] = (uint8_t)(s | ST_E);
                root[v] = prev_winner ? prev_winner[orc_slot(colour[v], slot_bits, hash_kind)] : v;
                continue;
            }
            colour[v] = (3u * colour[v] + k) & colour_mask;
            state[v] = 0;
        }
        uint32_t sl = orc_slot(colour[v], slot_bits, hash_kind);
        if (winner[sl] < 0) winner[sl] = v; /* ascending v: first seen is the lowest id */
    }
    for (size_t s = 0; s < nslots; s++) {
        if (winner[s] >= 0) { state[winner[s]] = 19; npiv++; }
    }
    return npiv;
}

/* Count of FB-eliminated, not trimmed vertices: countElimmed (scc.cu:42-51). */
int64_t orc_count_elimmed(int32_t n, const uint8_t *state) {
    int64_t c = 0;
    for (int32_t v = 0; v < n; v++) if ((state[v] & 12) == 4) c++;
    return c;
}

/* One trim-2 pass: trim2_kernel (scc_routines.cu:397-523).  A live v whose  */
/* only live same-colour out-neighbour (other than itself) is u, while u's   */
/* only such out-neighbour is v, forms a size-2 SCC with u; likewise over    */
/* in-edges.  The lower id gets 20 (P|E), the other 4 (E).  Every test reads */
/* the states as they were before the pass (the reference reads them while  */
/* other threads write, which makes its result timing-dependent).  Returns   */
/* the number of pairs.                                                      */
static int32_t only_nb(int32_t v, const int32_t *row, const int32_t *col, const uint32_t *colour,
                       const uint8_t *st) {
    int32_t found = -1;
    for (int32_t e = row[v]; e < row[v + 1]; e++) {
        int32_t w = col[e];
        if (w == v || (st[w] & ST_E) || colour[w] != colour[v]) continue;
        if (found >= 0) return -2;
        found = w;
    }
    return found;
}

int64_t orc_trim2(int32_t n, const int32_t *row, const int32_t *col, const int32_t *rowT,
                  const int32_t *colT, const uint32_t *colour, uint8_t *state, int32_t *root) {
    uint8_t *snap = (uint8_t *)malloc((size_t)(n > 0 ? n : 1));
    int64_t pairs = 0;
    memcpy(snap, state, (size_t)n);
    for (int32_t v = 0; v < n; v++) {
        if (snap[v] & ST_E) continue;
        int32_t u = -1;
        int32_t a = only_nb(v, row, col, colour, snap);
        if (a >= 0 && only_nb(a, row, col, colour, snap) == v) u = a;
        if (u < 0) {
            a = only_nb(v, rowT, colT, colour, snap);
            if (a >= 0 && only_nb(a, rowT, colT, colour, snap) == v) u = a;
        }
        if (u < 0) continue;
        int32_t lo = v < u ? v : u;
        if (v == lo) pairs++;
        state[v] = (uint8_t)(v == lo ? 20 : 4);
        root[v] = lo;
    }
    free(snap);
    return pairs;
}

/* Labels of the one WCC pass: computeWCC (wcc.cu:7-52) iterated to its      */
/* fixpoint gives each live vertex the minimum id among live vertices it      */
/* reaches over same-colour out-edges.  Computed here by backward BFS from    */
/* each vertex in increasing id order, stopping at already-labelled vertices. */
/* Eliminated vertices keep label = own id (thrust::sequence, wcc.cu:164).    */
void orc_wcc_labels(int32_t n, const int32_t *rowT, const int32_t *colT, const uint32_t *colour,
                    const uint8_t *state, int32_t *label) {
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    uint8_t *done = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    for (int32_t v = 0; v < n; v++) label[v] = v;
    for (int32_t x = 0; x < n; x++) {
        if ((state[x] & ST_E) || done[x]) continue;
        int32_t h = 0, t = 0;
        done[x] = 1; label[x] = x; q[t++] = x;
        while (h < t) {
            int32_t v = q[h++];
            for (int32_t e = rowT[v]; e < rowT[v + 1]; e++) {
                int32_t w = colT[e];
                if (done[w] || (state[w] & ST_E) || colour[w] != colour[v]) continue;
                done[w] = 1; label[w] = x; q[t++] = w;
            }
        }
    }
    free(q); free(done);
}

/* Recolouring after the WCC labels: update_ranges2/3 (wcc.cu:79-144).  Each */
/* live class gets a fresh colour taken from its root (label == id) and all  */
/* members copy it.  The reference draws fresh colours with atomicAdd on    */
/* max(colour) (wcc.cu:92), whose order is not fixed; here the fresh colour */
/* is (maxc + 1 + root id) & mask.  The reference then makes every root a   */
/* pivot (state 19, wcc.cu:98); here the caller runs the ordinary election  */
/* on the new colours instead, which elects the root of each class (it is  */
/* the class's lowest id) unless two classes share an election slot, and   */
/* never puts two pivots in one colour even when masked colours collide.    */
/* Returns the number of classes.                                           */
int64_t orc_wcc_recolour(int32_t n, uint32_t *colour, uint8_t *state, const int32_t *label,
                         uint32_t maxc, uint32_t colour_mask) {
    int64_t roots = 0;
    for (int32_t v = 0; v < n; v++) {
        if (state[v] & ST_E) continue;
        if (label[v] == v) {
            colour[v] = (maxc + 1u + (uint32_t)v) & colour_mask;
            roots++;
        }
    }
    for (int32_t v = 0; v < n; v++) {
        if (state[v] & ST_E) continue;
        if (label[v] != v) colour[v] = colour[label[v]];
        state[v] = 0;
    }
    return roots;
}

/* ------------------------------------------------------------------ */
/* Whole FB loop: computeSCC (scc.cu:150-1022) in the order trim-1,     */
/* first pivot, then per step {reach F, reach B, trim, update+elect}.   */
/* variant 0 = scc, 1 = scc_w (one WCC pass once FB-eliminated >= n/100, */
/* scc.cu:745-823), 2 = scc_2_w (plus one trim-2 pass, scc.cu:797-799). */
/* init_colour NULL means colour 1 everywhere (thrust::fill, scc.cu:214).*/
/* Outputs: final colour/state, root (class representative), stats.     */
/* ------------------------------------------------------------------ */
int64_t orc_fb(int32_t n, const int32_t *row, const int32_t *col, const int32_t *rowT,
               const int32_t *colT, const uint32_t *init_colour, uint32_t colour_mask,
               int slot_bits, int hash_kind, int variant, uint32_t *colour, uint8_t *state,
               int32_t *root, orc_fb_stats *st) {
    size_t nslots = (size_t)1 << slot_bits;
    int32_t *win_a = (int32_t *)malloc(sizeof(int32_t) * nslots);
    int32_t *win_b = (int32_t *)malloc(sizeof(int32_t) * nslots);
    int32_t *lab = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *prev = win_a, *cur = win_b, *tmp;
    orc_fb_stats s;
    int wcc_done = 0;
    int64_t threshold = n / 100;
    memset(&s, 0, sizeof(s));
    for (int32_t v = 0; v < n; v++) {
        colour[v] = init_colour ? (init_colour[v] & colour_mask) : 1u;
        state[v] = 0;
        root[v] = v;
    }
    s.trimmed += orc_trim(n, row, col, rowT, colT, colour, state, root, 0);
    int64_t np = orc_update_elect(n, colour, state, root, colour_mask, slot_bits, hash_kind, NULL,
                                  prev, 0);
    s.pivots_total += np;
    int has_pivot = np > 0;
    while (has_pivot) {
        s.steps++;
        orc_reach(n, row, col, colour, state, ST_F);
        orc_reach(n, rowT, colT, colour, state, ST_B);
        s.trimmed += orc_trim(n, row, col, rowT, colT, colour, state, root, 0);
        if (variant != 0 && !wcc_done) {
            np = orc_update_elect(n, colour, state, root, colour_mask, slot_bits, hash_kind, prev,
                                  cur, 1);
            tmp = prev; prev = cur; cur = tmp;
            int64_t ne = orc_count_elimmed(n, state);
            if (ne >= threshold) {
                if (ne == n) { s.pivots_total += np; break; }
                wcc_done = 1;
                s.wcc_step = s.steps;
                for (int32_t v = 0; v < n; v++) if (state[v] == 19) state[v] = 0; /* unsetPivots */
                if (variant == 2) s.trim2_pairs += orc_trim2(n, row, col, rowT, colT, colour, state, root);
                uint32_t maxc = 0;
                for (int32_t v = 0; v < n; v++) if (colour[v] > maxc) maxc = colour[v];
                orc_wcc_labels(n, rowT, colT, colour, state, lab);
                orc_wcc_recolour(n, colour, state, lab, maxc, colour_mask);
                np = orc_update_elect(n, colour, state, root, colour_mask, slot_bits, hash_kind,
                                      NULL, prev, 0);
            }
        } else {
            np = orc_update_elect(n, colour, state, root, colour_mask, slot_bits, hash_kind, prev,
                                  cur, 1);
            tmp = prev; prev = cur; cur = tmp;
        }
        s.pivots_total += np;
        has_pivot = np > 0;
    }
    s.num_scc = 0;
    for (int32_t v = 0; v < n; v++) if (state[v] & ST_P) s.num_scc++;
    if (st) *st = s;
    free(win_a); free(win_b); free(lab);
    return s.num_scc;
}
