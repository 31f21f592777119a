// Launchers of the HIP kernels (kernels_fb.hip, kernels_reach.hip), called by
// the host engine (engine.hip).  Host-compilable declarations only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpuscc {

struct Ctl;

struct Launch {
    hipStream_t stream;
    uint32_t grid;   // workgroups (persistent grid-stride)
    uint32_t block;  // work-items per workgroup, multiple of 64
};

struct Graph {  // device CSR + CSR^T
    int32_t n;
    int64_t m;
    const int32_t *row, *col, *rowT, *colT;
};

// kernels_fb.hip
void launch_init(const Launch &l, uint32_t *word, int32_t *root, int32_t n, const uint32_t *ic);
void launch_fill_u64(const Launch &l, unsigned long long *p, int64_t count, unsigned long long v);
void launch_fill_i32(const Launch &l, int32_t *p, int64_t count, int32_t v);
void launch_ctl_reset(hipStream_t s, Ctl *ctl, int what);
void launch_trim_sweep(const Launch &l, const Graph &g, uint32_t *word, int32_t *root, Ctl *ctl,
                       int iter);
void launch_update(const Launch &l, uint32_t *word, int32_t *root, int32_t n,
                   const unsigned long long *prev, unsigned long long *cur, int slot_bits,
                   unsigned long long tag, int recolour, const uint16_t *score);
void launch_confirm(const Launch &l, uint32_t *word, int32_t n, const unsigned long long *cur,
                    int slot_bits, int32_t *piv, Ctl *ctl);
void launch_count_state(const Launch &l, const uint32_t *word, int32_t n, int what, Ctl *ctl);
void launch_max_colour(const Launch &l, const uint32_t *word, int32_t n, Ctl *ctl);
void launch_unset_pivots(const Launch &l, uint32_t *word, int32_t n);
void launch_trim2(const Launch &l, const Graph &g, uint32_t *word, int32_t *root, Ctl *ctl);
void launch_wcc_init(const Launch &l, int32_t *lab, int32_t n);
void launch_wcc_iter(const Launch &l, const Graph &g, const uint32_t *word, int32_t *lab, Ctl *ctl,
                     int iter);
void launch_wcc_recolour(const Launch &l, uint32_t *word, const int32_t *lab, int32_t n,
                         const Ctl *ctl);
void launch_labels(const Launch &l, const int32_t *root, int32_t *mn, int32_t *label, int32_t n);
void launch_collect(const Launch &l, const uint32_t *word, int32_t n, uint32_t bit, int32_t *q,
                    Ctl *ctl);

// kernels_reach.hip -- one level of a reachability chain.  dir 0: forward over
// (row, col) with bit F / FExt; dir 1: backward over (rowT, colT) with B / BExt.
// Queue levels read `inq` (length ctl->qsize[iter&3]) and append to `outq`.
void launch_reach_sweep(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                        int32_t n, int dir, Ctl *ctl, int iter, int instr);
void launch_reach_topo(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                       int32_t n, int dir, Ctl *ctl, int iter, int instr);
void launch_reach_queue(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                        const int32_t *inq, int32_t *outq, int dir, Ctl *ctl, int iter, int instr);
void launch_reach_lb(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                     const int32_t *inq, int32_t *outq, int dir, Ctl *ctl, int iter, int instr);
void launch_reach_hybrid(const Launch &l, const int32_t *row, const int32_t *col, uint32_t *word,
                         int32_t n, const int32_t *inq, int32_t *outq, int dir, Ctl *ctl, int iter,
                         uint32_t threshold, int instr);
// clears the expanded bit (FExt or BExt) of every vertex: queue levels of the
// hybrid mark expanded vertices so a later sweep level skips them
void launch_clear_bits(const Launch &l, uint32_t *word, int32_t n, uint32_t bits);

}  // namespace gpuscc
