// Learned pivot scorer for gfx950: per-vertex features, fused feature +
// MLP inference (12 -> 32 -> 32 -> 1, BF16 weights and activations, FP32
// accumulation), and the training kernels (forward + backward per sample,
// gradient reduction, Adam).  No reference counterpart: the reference's pivot
// is whichever vertex wins an atomicCAS race (scc_routines.cu:214-221).
//
// Parameter vector (1505 floats): W1[32][12], b1[32], W2[32][32], b2[32],
// w3[32], b3.  Feature vector (12 floats per vertex), all normalised:
//   0 log1p(out-degree within the colour, live)   1 log1p(in-degree within ...)
//   2 log1p(out-degree)                          3 log1p(in-degree)
//   4 log1p(sum of out-neighbours' out-degrees)  5 log1p(sum of in-neighbours' in-degrees)
//   6 log1p(FB step)                             7 log1p(colour-set size)
//   8 out_c / (out + 1)                          9 in_c / (in + 1)
//  10 min(in_c, out_c) / (max(in_c, out_c) + 1) 11 log1p(in_c * out_c)
#include "fb_device.hpp"
#include "kernels.hpp"

namespace gpuscc {

constexpr int NF = MLP_IN, NH = MLP_HID;
constexpr int OFF_W1 = 0, OFF_B1 = NH * NF, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH;
static_assert(OFF_B3 + 1 == MLP_PARAMS, "parameter layout");

constexpr float LOG_SCALE = 1.0f / 16.0f;  // log1p of degrees up to ~2^23 maps below ~1
constexpr int32_t DEG_SAMPLE = 256;        // rows longer than this are sampled for within-colour degrees
// The per-vertex network keeps its feature vector and hidden layer in
// registers: these kernels run 256-wide workgroups (one wave per SIMD), which
// lets the compiler use up to 512 VGPRs per wave instead of spilling them.
constexpr int MLP_BLOCK = 256;
constexpr int FEAT_ILP = 8;                // arcs per step of a feature row scan: their loads in flight together

__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (finite inputs): gfx950's v_cvt_pk_bf16_f32
__device__ inline uint16_t f_to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ inline float bf16_round(float f) { return bf16_to_f(f_to_bf16(f)); }

// ---------------------------------------------------------------------------
// static features (computed once per graph): log1p of out/in degree and of
// the 2-hop degree sums
// ---------------------------------------------------------------------------
__global__ void k_static_features(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                  const int32_t *__restrict__ rowT, const int32_t *__restrict__ colT,
                                  int32_t n, float4 *out) {
    int64_t beg, end;
    my_range(n, beg, end);
    for (int64_t vv = beg + threadIdx.x; vv < end; vv += blockDim.x) {
        int32_t v = (int32_t)vv;
        int32_t o0 = row[v], o1 = row[v + 1], i0 = rowT[v], i1 = rowT[v + 1];
        float h_out = 0.f, h_in = 0.f;
        // neighbour degrees FEAT_ILP at a time, summed in row order
        int32_t oe = o1 - o0 > DEG_SAMPLE ? o0 + DEG_SAMPLE : o1;
        for (int32_t e = o0; e < oe; e += FEAT_ILP) {
            int32_t u[FEAT_ILP], d[FEAT_ILP];
#pragma unroll
            for (int j = 0; j < FEAT_ILP; j++) u[j] = e + j < oe ? col[e + j] : -1;
#pragma unroll
            for (int j = 0; j < FEAT_ILP; j++) d[j] = u[j] >= 0 ? row[u[j] + 1] - row[u[j]] : 0;
#pragma unroll
            for (int j = 0; j < FEAT_ILP; j++)
                if (u[j] >= 0) h_out += (float)d[j];
        }
        if (o1 - o0 > DEG_SAMPLE) h_out *= (float)(o1 - o0) / (float)DEG_SAMPLE;
        int32_t ie = i1 - i0 > DEG_SAMPLE ? i0 + DEG_SAMPLE : i1;
        for (int32_t e = i0; e < ie; e += FEAT_ILP) {
            int32_t u[FEAT_ILP], d[FEAT_ILP];
#pragma unroll
            for (int j = 0; j < FEAT_ILP; j++) u[j] = e + j < ie ? colT[e + j] : -1;
#pragma unroll
            for (int j = 0; j < FEAT_ILP; j++) d[j] = u[j] >= 0 ? rowT[u[j] + 1] - rowT[u[j]] : 0;
#pragma unroll
            for (int j = 0; j < FEAT_ILP; j++)
                if (u[j] >= 0) h_in += (float)d[j];
        }
        if (i1 - i0 > DEG_SAMPLE) h_in *= (float)(i1 - i0) / (float)DEG_SAMPLE;
        out[v] = make_float4(log1pf((float)(o1 - o0)), log1pf((float)(i1 - i0)), log1pf(h_out),
                             log1pf(h_in));
    }
}

// live vertices per colour slot (the colour-set size, slots shared by
// colliding colours are summed), over the vertices lin[0 .. *lin_cnt) or, with
// lin null, all of [0, n)
__global__ void k_colour_count(const uint32_t *word, int32_t n, uint32_t *cnt, int slot_bits, const int32_t *lin,
                               const uint32_t *lin_cnt, const Ctl *ctl) {
    if (ctl && __atomic_load_n(&ctl->abort, __ATOMIC_RELAXED)) return;
    int64_t beg, end;
    my_range(lin ? (int64_t)*lin_cnt : (int64_t)n, beg, end);
    int lane = threadIdx.x & (WAVE - 1);
    // a wave-uniform pending count of one slot, carried across rounds: a colour set
    // spread over many waves (the first election's one colour per graph) costs one
    // atomic per wave, not one per round
    uint32_t pslot = 0xFFFFFFFFu, pcnt = 0;
    for (int64_t base = beg; base < end; base += blockDim.x) {
        int64_t i = base + threadIdx.x;
        bool live = false;
        uint32_t sl = 0;
        if (i < end) {
            uint32_t w = word[lin ? lin[i] : (int32_t)i];
            live = !(w & ST_E);
            sl = slot_of(colour_of(w), slot_bits);
        }
        unsigned long long mask = __ballot(live);
        if (mask == 0) continue;
        int first = __ffsll((long long)mask) - 1;
        uint32_t sl0 = __shfl(sl, first, WAVE);
        if (__all(!live || sl == sl0)) {
            if (sl0 != pslot) {
                if (pcnt && lane == 0) atomicAdd(&cnt[pslot], pcnt);
                pslot = sl0;
                pcnt = 0;
            }
            pcnt += (uint32_t)__popcll(mask);
        } else if (live) {
            atomicAdd(&cnt[sl], 1u);
        }
    }
    if (pcnt && lane == 0) atomicAdd(&cnt[pslot], pcnt);
}

__device__ inline uint32_t live_same_colour(const int32_t *__restrict__ row,
                                            const int32_t *__restrict__ col, const uint32_t *word,
                                            int32_t v, uint32_t c) {
    int32_t e0 = row[v], e1 = row[v + 1];
    int32_t ee = e1 - e0 > DEG_SAMPLE ? e0 + DEG_SAMPLE : e1;
    uint32_t k = 0;
    for (int32_t e = e0; e < ee; e += FEAT_ILP) {  // FEAT_ILP arcs: column ids, then words, then tests
        int32_t u[FEAT_ILP];
        uint32_t wu[FEAT_ILP];
#pragma unroll
        for (int j = 0; j < FEAT_ILP; j++) u[j] = e + j < ee ? col[e + j] : -1;
#pragma unroll
        for (int j = 0; j < FEAT_ILP; j++) wu[j] = u[j] >= 0 && u[j] != v ? word[u[j]] : ST_E;  // no self-loop
#pragma unroll
        for (int j = 0; j < FEAT_ILP; j++) k += (!(wu[j] & ST_E) && colour_of(wu[j]) == c) ? 1u : 0u;
    }
    if (e1 - e0 > DEG_SAMPLE) k = (uint32_t)((uint64_t)k * (uint64_t)(e1 - e0) / DEG_SAMPLE);
    return k;
}

__device__ inline void load_params(const uint16_t *pb, float *sp) {
    for (int i = threadIdx.x; i < MLP_PARAMS; i += blockDim.x) sp[i] = bf16_to_f(pb[i]);
    __syncthreads();
}

// ---------------------------------------------------------------------------
// The network on MFMA (v_mfma_f32_32x32x8_bf16): one wave scores 32 vertices.
// The layers are computed transposed, D1 = W1' X^T (hidden x vertices) and
// D2 = W2 H1^T, so that the layer-1 result is already in the B-operand layout
// of layer 2.  Lane l = 32h + j holds A[j][8c + 4h .. +3] and B[8c + 4h .. +3][j]
// of K chunk c, and D[8q + 4h + s][j] in accumulator register 4q + s.  W1' is
// W1 with the bias as feature 12 (input 1) and features 13-15 zero.  FP32
// features enter as two BF16 parts (x = hi + lo), so layer 1 sees 16 bits of
// each feature, not 8; weights and hidden activations are BF16, accumulation
// FP32, as in the scalar restatement (tests compare them).
// ---------------------------------------------------------------------------
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

struct MlpRegs {
    bf16x4_t a1[2];  // W1' chunks
    bf16x4_t a2[4];  // W2 chunks
    float b2[16];    // b2 of the accumulator rows this lane holds
    float w3[16];    // w3 of the same rows
    float b3;
};

__device__ inline void mlp_regs_load(const uint16_t *__restrict__ pb, MlpRegs &R) {
    const int l = threadIdx.x & (WAVE - 1), j = l & 31, h = l >> 5;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = 8 * c + 4 * h + t;
            R.a1[c][t] = (short)(k < NF ? pb[OFF_W1 + j * NF + k] : (k == NF ? pb[OFF_B1 + j] : 0));
        }
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int t = 0; t < 4; t++) R.a2[c][t] = (short)pb[OFF_W2 + j * NH + 8 * c + 4 * h + t];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = 8 * (r >> 2) + 4 * h + (r & 3);
        R.b2[r] = bf16_to_f(pb[OFF_B2 + i]);
        R.w3[r] = bf16_to_f(pb[OFF_W3 + i]);
    }
    R.b3 = bf16_to_f(pb[OFF_B3]);
}

// Logit of vertex j of the tile, returned in lanes j and j + 32.  x[0..3] are
// this lane's features 4h .. 4h + 3 and x[4..7] its features 8 + 4h .. 11 + 4h
// (for h = 1: 12 .. 15 = 1, 0, 0, 0).  Every lane of the wave must call it.
__device__ inline float mlp_logit_mfma(const MlpRegs &R, const float (&x)[8]) {
    bf16x4_t hi[2], lo[2];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const float f = x[4 * c + t];
            const uint16_t a = f_to_bf16(f);
            hi[c][t] = (short)a;
            lo[c][t] = (short)f_to_bf16(f - bf16_to_f(a));
        }
    f32x16_t d1;
#pragma unroll
    for (int r = 0; r < 16; r++) d1[r] = 0.f;
#pragma unroll
    for (int c = 0; c < 2; c++) {
        d1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(R.a1[c], hi[c], d1, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(R.a1[c], lo[c], d1, 0, 0, 0);
    }
    bf16x4_t h1[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const float v = d1[4 * c + t];
            h1[c][t] = (short)f_to_bf16(v > 0.f ? v : 0.f);
        }
    f32x16_t d2;
#pragma unroll
    for (int r = 0; r < 16; r++) d2[r] = R.b2[r];
#pragma unroll
    for (int c = 0; c < 4; c++) d2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(R.a2[c], h1[c], d2, 0, 0, 0);
    float part = 0.f;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const float v = d2[r];
        part += R.w3[r] * bf16_round(v > 0.f ? v : 0.f);
    }
    part += __shfl_xor(part, 32, WAVE);
    return part + R.b3;
}

// Features, scores and the learned election, one wave per 64 vertices of the
// list lin[0 .. *lin_cnt) (all of [0, n) with lin null): every lane loads and
// computes the 12 features of its own vertex (the within-colour degrees by row
// scans, or, in the first election of a run, where every arc joins two
// vertices of one colour, the degrees themselves); the network then runs as
// two MFMA tiles of 32 vertices, their operands gathered with shuffles, and
// each tile's logit comes back to the lane of its vertex.  Outputs, each
// optional: feat_out[v][12] (zeros for eliminated vertices), score_out[v] =
// round(sigmoid(logit) * 65535) (0 when eliminated), and the election key
// [tag][65535 - score][v] offered to cur[slot of the colour] (lowest key wins:
// the highest score, then the lowest id).  The 16-bit probability ties the
// most confident vertices, and the lowest id among them wins; ranking by the
// logit itself instead (resolution 1/2048) elects the largest hub of R-MAT
// scale 24, whose first reach level pushes its whole neighbourhood: one level
// fewer, but 1.18 ms of reach instead of 0.87.
struct ScoreArgs {
    const int32_t *row, *col, *rowT, *colT;
    const uint32_t *word;
    const float4 *stat;
    const uint32_t *cnt;
    const uint32_t *one_set;  // not null: every live vertex is in one colour set of this size
    int slot_bits;
    float step_f;
    int static_deg;
    int32_t n;
    const int32_t *lin;
    const uint32_t *lin_cnt;
    const uint16_t *params;
    uint16_t *score_out;
    float *feat_out;
    unsigned long long *cur;
    unsigned long long tag;
    const Ctl *ctl;
    const int32_t *oid;  // the caller's ids (the keys' last field), null: the same
};

__global__ __launch_bounds__(MLP_BLOCK) void k_score(ScoreArgs a) {
    if (a.ctl && __atomic_load_n(&a.ctl->abort, __ATOMIC_RELAXED)) return;
    MlpRegs R;
    if (a.params) mlp_regs_load(a.params, R);
    const int l = threadIdx.x & (WAVE - 1), j = l & 31, h = l >> 5;
    const int64_t total = a.lin ? (int64_t)*a.lin_cnt : (int64_t)a.n;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / WAVE;
    // the wave's pending minimum key of one slot, carried across rounds (flushed when
    // the slot changes and at the end)
    uint32_t pslot = 0xFFFFFFFFu;
    unsigned long long pmin = ~0ull;
    // feature 7 of the last slot whose count the wave read (every lane asks for one
    // slot in the first election: one read per wave instead of one per round)
    uint32_t cslot = 0xFFFFFFFFu;
    float cfeat = 0.f;
    if (a.one_set) {
        cslot = 0xFFFFFFFEu;
        cfeat = log1pf((float)*a.one_set) * LOG_SCALE;
    }
    for (int64_t base = wave * WAVE; base < total; base += nwaves * WAVE) {  // wave-uniform
        const int64_t i = base + l;
        const int32_t v = i < total ? (a.lin ? a.lin[i] : (int32_t)i) : -1;
        const uint32_t w = v >= 0 ? a.word[v] : ST_E;
        const bool live = !(w & ST_E);
        float f[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) f[k] = 0.f;
        const uint32_t sl = live ? slot_of(colour_of(w), a.slot_bits) : 0u;
        if (!a.one_set) {  // the colour-set sizes: one read per wave while its slot repeats
            const unsigned long long lm = __ballot(live);
            const uint32_t s0 = lm ? __shfl(sl, __ffsll((long long)lm) - 1, WAVE) : 0u;
            if (lm && __all(!live || sl == s0)) {
                if (s0 != cslot) {
                    cslot = s0;
                    cfeat = log1pf((float)a.cnt[s0]) * LOG_SCALE;
                }
            } else if (lm) {
                cslot = 0xFFFFFFFFu;
                cfeat = live ? log1pf((float)a.cnt[sl]) * LOG_SCALE : 0.f;
            }
        }
        if (live) {
            // the hardware log, exp and reciprocal (a few ulp): the kernel is bound by its
            // VALU issue, and training reads its features from this same code
            const uint32_t c = colour_of(w);
            const float4 st = a.stat[v];
            const float dout = (float)(a.row[v + 1] - a.row[v]), din = (float)(a.rowT[v + 1] - a.rowT[v]);
            float oc = dout, ic = din;
            if (!a.static_deg) {
                oc = (float)live_same_colour(a.row, a.col, a.word, v, c);
                ic = (float)live_same_colour(a.rowT, a.colT, a.word, v, c);
            }
            f[0] = __logf(oc + 1.f) * LOG_SCALE;
            f[1] = __logf(ic + 1.f) * LOG_SCALE;
            f[2] = st.x * LOG_SCALE;
            f[3] = st.y * LOG_SCALE;
            f[4] = st.z * LOG_SCALE * 0.5f;
            f[5] = st.w * LOG_SCALE * 0.5f;
            f[6] = a.step_f;
            f[7] = cfeat;
            f[8] = __fdividef(oc, dout + 1.f);
            f[9] = __fdividef(ic, din + 1.f);
            const float mx = oc > ic ? oc : ic, mn = oc > ic ? ic : oc;
            f[10] = __fdividef(mn, mx + 1.f);
            f[11] = __logf(oc * ic + 1.f) * LOG_SCALE * 0.5f;
        }
        if (a.feat_out && v >= 0) {
            float *o = a.feat_out + (int64_t)v * NF;
#pragma unroll
            for (int k = 0; k < NF; k++) o[k] = f[k];
        }
        if (!a.params) continue;  // kernel-uniform
        float logit = 0.f;
#pragma unroll
        for (int t = 0; t < 2; t++) {
            // B operands of tile t: lane (h, j) takes features 4h.. and 8 + 4h.. of vertex 32t + j
            const int src = 32 * t + j;
            float g[NF];
#pragma unroll
            for (int k = 0; k < NF; k++) g[k] = __shfl(f[k], src, WAVE);
            float x[8];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                x[k] = h == 0 ? g[k] : g[4 + k];
                x[4 + k] = h == 0 ? g[8 + k] : (k == 0 ? 1.f : 0.f);  // feature 12: the bias input
            }
            const float s = mlp_logit_mfma(R, x);  // vertex 32t + j, in lanes j and j + 32
            if (h == t) logit = s;                  // lane 32t + j owns vertex 32t + j
        }
        uint16_t sc = 0;
        if (live) sc = (uint16_t)__float2uint_rn(__fdividef(65535.f, 1.f + __expf(-logit)));
        if (v >= 0 && a.score_out) a.score_out[v] = sc;
        if (!a.cur) continue;  // kernel-uniform
        // the election: one pending minimum per wave while its candidates share a slot
        const unsigned long long key =
            live ? (a.tag | ((unsigned long long)(0xFFFFu - sc) << 32) | (uint32_t)orig_id(a.oid, v)) : ~0ull;
        const unsigned long long bm = __ballot(live);
        if (bm == 0) continue;  // wave-uniform
        const uint32_t sl0 = __shfl(sl, __ffsll((long long)bm) - 1, WAVE);
        if (__all(!live || sl == sl0)) {
            unsigned long long m = key;
            for (int o = WAVE / 2; o > 0; o >>= 1) {
                const unsigned long long y = __shfl_xor(m, o, WAVE);
                m = y < m ? y : m;
            }
            if (sl0 != pslot) {
                if (pslot != 0xFFFFFFFFu && l == 0 && pmin < a.cur[pslot]) atomicMin(&a.cur[pslot], pmin);
                pslot = sl0;
                pmin = m;
            } else {
                pmin = m < pmin ? m : pmin;
            }
        } else if (live && key < a.cur[sl]) {
            atomicMin(&a.cur[sl], key);
        }
    }
    if (a.cur && pslot != 0xFFFFFFFFu && l == 0 && pmin < a.cur[pslot]) atomicMin(&a.cur[pslot], pmin);
}

// inference on given feature rows (tests): logits out
__global__ __launch_bounds__(MLP_BLOCK) void k_mlp_infer(const float *X, int64_t N, const uint16_t *params, float *logit) {
    MlpRegs R;
    mlp_regs_load(params, R);
    const int l = threadIdx.x & (WAVE - 1), j = l & 31, h = l >> 5;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / WAVE;
    for (int64_t base = wave * 32; base < N; base += nwaves * 32) {  // wave-uniform
        const int64_t i = base + j;
        float x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = 0.f;
        if (i < N) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                x[t] = X[i * NF + 4 * h + t];
                if (h == 0) x[4 + t] = X[i * NF + 8 + t];
            }
        }
        if (h == 1) x[4] = 1.f;
        const float s = mlp_logit_mfma(R, x);
        if (h == 0 && i < N) logit[i] = s;
    }
}

// ---------------------------------------------------------------------------
// training: per-sample forward + backward of the binary cross-entropy of
// sigmoid(logit) against y, mean over the batch
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(MLP_BLOCK) void k_mlp_fwdbwd(const float *X, const float *y, const int32_t *idx, int64_t N,
                             const uint16_t *params, uint16_t *H1, uint16_t *H2, float *dH1,
                             float *dH2, float *ds, float *Xb, float *loss_sum) {
    __shared__ float sp[MLP_PARAMS];
    __shared__ float sh[16];
    load_params(params, sp);
    float lsum = 0.f;
    const float invN = 1.f / (float)N;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < N;
         b += (int64_t)gridDim.x * blockDim.x) {
        int64_t src = idx ? idx[b] : b;
        float f[NF], h1[NH], h2[NH];
        for (int k = 0; k < NF; k++) {
            f[k] = X[src * NF + k];
            Xb[b * NF + k] = f[k];
        }
        for (int i = 0; i < NH; i++) {
            float a = sp[OFF_B1 + i];
            for (int k = 0; k < NF; k++) a += sp[OFF_W1 + i * NF + k] * f[k];
            h1[i] = bf16_round(a > 0.f ? a : 0.f);
        }
        float s = sp[OFF_B3];
        for (int i = 0; i < NH; i++) {
            float a = sp[OFF_B2 + i];
            for (int j = 0; j < NH; j++) a += sp[OFF_W2 + i * NH + j] * h1[j];
            h2[i] = bf16_round(a > 0.f ? a : 0.f);
            s += sp[OFF_W3 + i] * h2[i];
        }
        float t = y[src];
        float p = 1.f / (1.f + __expf(-s));
        lsum += fmaxf(s, 0.f) - s * t + log1pf(__expf(-fabsf(s)));
        float g = (p - t) * invN;
        ds[b] = g;
        float d2[NH];
        for (int i = 0; i < NH; i++) {
            d2[i] = h2[i] > 0.f ? g * sp[OFF_W3 + i] : 0.f;
            dH2[b * NH + i] = d2[i];
            H1[b * NH + i] = f_to_bf16(h1[i]);
            H2[b * NH + i] = f_to_bf16(h2[i]);
        }
        for (int j = 0; j < NH; j++) {
            float a = 0.f;
            for (int i = 0; i < NH; i++) a += sp[OFF_W2 + i * NH + j] * d2[i];
            dH1[b * NH + j] = h1[j] > 0.f ? a : 0.f;
        }
    }
    lsum = block_sum(lsum, sh);
    if (threadIdx.x == 0) atomicAdd(loss_sum, lsum);
}

// gradient of parameter p summed over the samples [blockIdx.y share]; grad
// must be zeroed before.  One work-item per parameter.
__global__ void k_mlp_grad(const float *Xb, const uint16_t *H1, const uint16_t *H2, const float *dH1,
                           const float *dH2, const float *ds, int64_t N, float *grad) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    int64_t per = (N + gridDim.y - 1) / gridDim.y;
    int64_t b0 = (int64_t)blockIdx.y * per, b1 = b0 + per < N ? b0 + per : N;
    float acc = 0.f;
    if (p < OFF_B1) {
        int i = p / NF, k = p % NF;
        for (int64_t b = b0; b < b1; b++) acc += dH1[b * NH + i] * Xb[b * NF + k];
    } else if (p < OFF_W2) {
        int i = p - OFF_B1;
        for (int64_t b = b0; b < b1; b++) acc += dH1[b * NH + i];
    } else if (p < OFF_B2) {
        int i = (p - OFF_W2) / NH, j = (p - OFF_W2) % NH;
        for (int64_t b = b0; b < b1; b++) acc += dH2[b * NH + i] * bf16_to_f(H1[b * NH + j]);
    } else if (p < OFF_W3) {
        int i = p - OFF_B2;
        for (int64_t b = b0; b < b1; b++) acc += dH2[b * NH + i];
    } else if (p < OFF_B3) {
        int j = p - OFF_W3;
        for (int64_t b = b0; b < b1; b++) acc += ds[b] * bf16_to_f(H2[b * NH + j]);
    } else {
        for (int64_t b = b0; b < b1; b++) acc += ds[b];
    }
    if (b1 > b0) atomicAdd(&grad[p], acc);
}

// Adam (Kingma & Ba 2015) on FP32 master weights; writes the BF16 copy
__global__ void k_adam(float *w, const float *g, float *m, float *v, uint16_t *wb, float lr,
                       float beta1, float beta2, float eps, float bc1, float bc2) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    float gp = g[p];
    float mp = beta1 * m[p] + (1.f - beta1) * gp;
    float vp = beta2 * v[p] + (1.f - beta2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    float wp = w[p] - lr * (mp / bc1) / (sqrtf(vp / bc2) + eps);
    w[p] = wp;
    wb[p] = f_to_bf16(wp);
}

__global__ void k_to_bf16(const float *w, uint16_t *wb, int count) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) wb[p] = f_to_bf16(w[p]);
}

// rows idx[i] of a [n][12] feature matrix -> out[i]
__global__ void k_gather_rows(const float *feat, const int32_t *idx, int64_t count, float *out) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count * NF;
         t += (int64_t)gridDim.x * blockDim.x)
        out[t] = feat[(int64_t)idx[t / NF] * NF + t % NF];
}

__global__ void k_fill_f32(float *p, int64_t count, float val) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

// ---------------------------------------------------------------------------
#define GRID(l) dim3((l).grid), dim3((l).block), 0, (l).stream
// the same number of work-items in MLP_BLOCK-wide workgroups
#define GRID_MLP(l) dim3(((l).grid * (l).block + MLP_BLOCK - 1) / MLP_BLOCK), dim3(MLP_BLOCK), 0, (l).stream

void launch_static_features(const Launch &l, const Graph &g, float4 *out) {
    hipLaunchKernelGGL(k_static_features, GRID(l), g.row, g.col, g.rowT, g.colT, g.n, out);
}
void launch_colour_count(const Launch &l, const uint32_t *word, int32_t n, uint32_t *cnt, int slot_bits,
                         const int32_t *lin, const uint32_t *lin_cnt, const Ctl *ctl) {
    hipLaunchKernelGGL(k_colour_count, GRID(l), word, n, cnt, slot_bits, lin, lin_cnt, ctl);
}
void launch_score(const Launch &l, const Graph &g, const uint32_t *word, const float4 *stat, const uint32_t *cnt,
                  const uint32_t *one_set, int slot_bits, float step_f, int static_deg, const int32_t *lin, const uint32_t *lin_cnt,
                  const uint16_t *params, uint16_t *score_out, float *feat_out, unsigned long long *cur,
                  unsigned long long tag, const Ctl *ctl, const int32_t *oid) {
    ScoreArgs a{g.row, g.col, g.rowT, g.colT, word, stat, cnt, one_set, slot_bits, step_f, static_deg, g.n,
                lin, lin_cnt, params, score_out, feat_out, cur, tag, ctl, oid};
    hipLaunchKernelGGL(k_score, GRID_MLP(l), a);
}
void launch_mlp_infer(const Launch &l, const float *X, int64_t N, const uint16_t *params, float *logit) {
    hipLaunchKernelGGL(k_mlp_infer, GRID_MLP(l), X, N, params, logit);
}
void launch_mlp_fwdbwd(const Launch &l, const float *X, const float *y, const int32_t *idx, int64_t N,
                       const uint16_t *params, MlpWork &w) {
    hipLaunchKernelGGL(k_mlp_fwdbwd, GRID_MLP(l), X, y, idx, N, params, w.H1, w.H2, w.dH1, w.dH2, w.ds,
                       w.Xb, w.loss);
}
void launch_mlp_grad(hipStream_t s, int64_t N, MlpWork &w) {
    hipLaunchKernelGGL(k_fill_f32, dim3(8), dim3(256), 0, s, w.grad, (int64_t)MLP_PARAMS, 0.f);
    uint32_t split = (uint32_t)((N + 1023) / 1024);
    split = split < 1 ? 1 : (split > 256 ? 256 : split);
    hipLaunchKernelGGL(k_mlp_grad, dim3((MLP_PARAMS + 255) / 256, split), dim3(256), 0, s, w.Xb, w.H1,
                       w.H2, w.dH1, w.dH2, w.ds, N, w.grad);
}
void launch_adam(hipStream_t s, float *wf, const float *g, float *m, float *v, uint16_t *wb, float lr,
                 float beta1, float beta2, float eps, int t) {
    float bc1 = 1.f - powf(beta1, (float)t), bc2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s, wf, g, m, v, wb, lr,
                       beta1, beta2, eps, bc1, bc2);
}
void launch_to_bf16(hipStream_t s, const float *w, uint16_t *wb) {
    hipLaunchKernelGGL(k_to_bf16, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s, w, wb, MLP_PARAMS);
}
void launch_gather_rows(const Launch &l, const float *feat, const int32_t *idx, int64_t count, float *out) {
    hipLaunchKernelGGL(k_gather_rows, GRID(l), feat, idx, count, out);
}
void launch_fill_f32(const Launch &l, float *p, int64_t count, float v) {
    hipLaunchKernelGGL(k_fill_f32, GRID(l), p, count, v);
}

hipError_t preload_mlp() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(k_score));
}

}  // namespace gpuscc
