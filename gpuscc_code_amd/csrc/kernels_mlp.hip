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

// round to nearest even on the upper 16 bits (finite inputs)
__device__ inline uint16_t f_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

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
// colliding colours are summed)
__global__ void k_colour_count(const uint32_t *word, int32_t n, uint32_t *cnt, int slot_bits) {
    int64_t beg, end;
    my_range(n, beg, end);
    int lane = threadIdx.x & (WAVE - 1);
    for (int64_t base = beg; base < end; base += blockDim.x) {
        int64_t v = base + threadIdx.x;
        bool live = false;
        uint32_t sl = 0;
        if (v < end) {
            uint32_t w = word[v];
            live = !(w & ST_E);
            sl = slot_of(colour_of(w), slot_bits);
        }
        unsigned long long mask = __ballot(live);
        if (mask == 0) continue;
        int first = __ffsll((long long)mask) - 1;
        uint32_t sl0 = __shfl(sl, first, WAVE);
        if (__all(!live || sl == sl0)) {
            if (lane == first) atomicAdd(&cnt[sl0], (uint32_t)__popcll(mask));
        } else if (live) {
            atomicAdd(&cnt[sl], 1u);
        }
    }
}

__device__ inline uint32_t live_same_colour(const int32_t *__restrict__ row,
                                            const int32_t *__restrict__ col, const uint32_t *word,
                                            int32_t v, uint32_t c, uint32_t &deg) {
    int32_t e0 = row[v], e1 = row[v + 1];
    deg = (uint32_t)(e1 - e0);
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

__device__ inline void vertex_features(const int32_t *row, const int32_t *col, const int32_t *rowT,
                                       const int32_t *colT, const uint32_t *word, const float4 *stat,
                                       const uint32_t *cnt, int slot_bits, float step_f, int32_t v,
                                       uint32_t w, float *f) {
    uint32_t c = colour_of(w), dout, din;
    uint32_t oc = live_same_colour(row, col, word, v, c, dout);
    uint32_t ic = live_same_colour(rowT, colT, word, v, c, din);
    float4 s = stat[v];
    float foc = (float)oc, fic = (float)ic;
    f[0] = log1pf(foc) * LOG_SCALE;
    f[1] = log1pf(fic) * LOG_SCALE;
    f[2] = s.x * LOG_SCALE;
    f[3] = s.y * LOG_SCALE;
    f[4] = s.z * LOG_SCALE * 0.5f;
    f[5] = s.w * LOG_SCALE * 0.5f;
    f[6] = step_f;
    f[7] = log1pf((float)cnt[slot_of(c, slot_bits)]) * LOG_SCALE;
    f[8] = foc / ((float)dout + 1.f);
    f[9] = fic / ((float)din + 1.f);
    float mx = foc > fic ? foc : fic, mn = foc > fic ? fic : foc;
    f[10] = mn / (mx + 1.f);
    f[11] = log1pf(foc * fic) * LOG_SCALE * 0.5f;
}

// the network on one feature vector; weights in LDS as floats, activations
// rounded to BF16, FP32 accumulation
__device__ inline float mlp_logit(const float *p, const float *f, float *h1) {
    for (int i = 0; i < NH; i++) {
        float a = p[OFF_B1 + i];
        for (int k = 0; k < NF; k++) a += p[OFF_W1 + i * NF + k] * f[k];
        h1[i] = bf16_round(a > 0.f ? a : 0.f);
    }
    float s = p[OFF_B3];
    for (int i = 0; i < NH; i++) {
        float a = p[OFF_B2 + i];
        for (int j = 0; j < NH; j++) a += p[OFF_W2 + i * NH + j] * h1[j];
        float h2 = bf16_round(a > 0.f ? a : 0.f);
        s += p[OFF_W3 + i] * h2;
    }
    return s;
}

__device__ inline void load_params(const uint16_t *pb, float *sp) {
    for (int i = threadIdx.x; i < MLP_PARAMS; i += blockDim.x) sp[i] = bf16_to_f(pb[i]);
    __syncthreads();
}

// Features (optional output) and, with parameters, the fused inference: the
// score of a live vertex is round(sigmoid(logit) * 65535); eliminated vertices
// score 0 and get zero features.
__global__ __launch_bounds__(MLP_BLOCK) void k_score(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                        const int32_t *__restrict__ rowT, const int32_t *__restrict__ colT,
                        const uint32_t *word, const float4 *stat, const uint32_t *cnt, int slot_bits,
                        float step_f, int32_t n, const uint16_t *params, uint16_t *score,
                        float *feat_out) {
    __shared__ float sp[MLP_PARAMS];
    if (params) load_params(params, sp);
    int64_t beg, end;
    my_range(n, beg, end);
    for (int64_t vv = beg + threadIdx.x; vv < end; vv += blockDim.x) {
        int32_t v = (int32_t)vv;
        uint32_t w = word[v];
        float f[NF];
        bool live = !(w & ST_E);
        if (live) {
            vertex_features(row, col, rowT, colT, word, stat, cnt, slot_bits, step_f, v, w, f);
        } else {
            for (int k = 0; k < NF; k++) f[k] = 0.f;
        }
        if (feat_out)
            for (int k = 0; k < NF; k++) feat_out[(int64_t)v * NF + k] = f[k];
        if (params) {
            uint16_t sc = 0;
            if (live) {
                float h1[NH];
                float s = mlp_logit(sp, f, h1);
                float p = 1.f / (1.f + __expf(-s));
                sc = (uint16_t)__float2uint_rn(p * 65535.f);
            }
            score[v] = sc;
        }
    }
}

// inference on given feature rows (tests): logits out
__global__ __launch_bounds__(MLP_BLOCK) void k_mlp_infer(const float *X, int64_t N, const uint16_t *params, float *logit) {
    __shared__ float sp[MLP_PARAMS];
    load_params(params, sp);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        float f[NF], h1[NH];
        for (int k = 0; k < NF; k++) f[k] = X[i * NF + k];
        logit[i] = mlp_logit(sp, f, h1);
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
void launch_colour_count(const Launch &l, const uint32_t *word, int32_t n, uint32_t *cnt, int slot_bits) {
    hipLaunchKernelGGL(k_colour_count, GRID(l), word, n, cnt, slot_bits);
}
void launch_score(const Launch &l, const Graph &g, const uint32_t *word, const float4 *stat,
                  const uint32_t *cnt, int slot_bits, float step_f, const uint16_t *params,
                  uint16_t *score, float *feat_out) {
    hipLaunchKernelGGL(k_score, GRID_MLP(l), g.row, g.col, g.rowT, g.colT, word, stat, cnt, slot_bits,
                       step_f, g.n, params, score, feat_out);
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

}  // namespace gpuscc
