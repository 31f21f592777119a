// scc_run_noalloc: a whole FB decomposition of a small graph that allocates
// nothing.  The caller's CSR and CSR^T and the output labels are device
// memory; every other buffer (vertex words, roots, frontier queues, live and
// pivot lists, the two election tables, the control block) is carved out of
// one file-scope device array of fixed size, so the call makes no hipMalloc,
// hipFree or any other allocating call and suits callers that cannot allocate
// (code under stream capture, tight serving loops).
//
// The FB steps are those of the one-workgroup tail (kernels_tail.hip): k_fb_tail
// starts from no pivots, so its first step is trim-1 plus the first election
// (the lowest live id, scc.cu:232-245), and every later step is reach F and B,
// trim, update and election in one launch of one workgroup; it clears an
// election table itself when the 16-bit election tag would wrap.  With the
// lowest-id rule every pivot is the lowest id of its colour set, so the roots
// are the canonical labels (minimum vertex id of each SCC) and are copied out.
//
// Every call on a device shares that device's array and finds it as the last
// call left it; the set-up kernel clears what the steps read.  Calls that
// overlap on one device (two host threads, or two streams) race on the array.
#include <algorithm>
#include <mutex>

#include "fb_device.hpp"
#include "gpuscc.h"
#include "kernels.hpp"

namespace gpuscc {
namespace {

constexpr int64_t NA_MAX_N = SCC_NOALLOC_MAX_VERTICES;
constexpr int NA_MAX_SLOT_BITS = 18;  // the election table size for NA_MAX_N vertices (2^18 slots)
static_assert(((int64_t)1 << NA_MAX_SLOT_BITS) >= NA_MAX_N, "a slot per vertex");

// Layout (bytes): words, roots, four queues, two live lists, the pivot list
// (9 int32 arrays of NA_MAX_N), two tables of 2^18 keys, the control block.
constexpr size_t NA_VEC = sizeof(int32_t) * (size_t)NA_MAX_N;
constexpr size_t NA_TAB = sizeof(unsigned long long) << NA_MAX_SLOT_BITS;
constexpr size_t NA_CTL_OFF = 9 * NA_VEC + 2 * NA_TAB;
constexpr size_t NA_BYTES = NA_CTL_OFF + ((sizeof(Ctl) + 255) & ~(size_t)255);

static __device__ __align__(256) uint8_t scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
uint8_t *na_ptr[NA_MAX_DEVICES] = {};  // device address of `scratch` on each device
// Election epoch of each device's calls: the tags go on from call to call (both
// tables are cleared at the start of a call, so no key outranks a new one; a
// call that crosses a 16-bit tag wrap clears the next table in the kernel).
uint32_t na_epoch[NA_MAX_DEVICES] = {};

// Device address of the array on the current device: hipGetSymbolAddress once
// per device, with that device current.
hipError_t scratch_of(int dev, uint8_t **out) {
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        na_ptr[dev] = (uint8_t *)p;
    }
    *out = na_ptr[dev];
    return hipSuccess;
}

// Set-up: one colour for every vertex, roots = self, the live list = all
// vertices.  Also checks the CSR pair the caller gave (offsets from 0 to m and
// never decreasing, ids in range); a violation sets ctl->count.
__global__ void k_noalloc_setup(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                const int32_t *__restrict__ rowT, const int32_t *__restrict__ colT, int32_t n,
                                int64_t m, uint32_t *word, int32_t *root, int32_t *list0, Ctl *ctl) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    for (int64_t v = i0; v < n; v += stride) {
        word[v] = make_word(1u, 0u);
        root[v] = (int32_t)v;
        list0[v] = (int32_t)v;
        const int32_t r0 = row[v], r1 = row[v + 1], t0 = rowT[v], t1 = rowT[v + 1];
        bad = bad || r1 < r0 || t1 < t0 || r0 < 0 || t0 < 0 || r1 > m || t1 > m;
        if (v == 0) bad = bad || r0 != 0 || t0 != 0;
        if (v == n - 1) bad = bad || r1 != m || t1 != m;
    }
    for (int64_t i = i0; i < m; i += stride) {
        const int32_t a = col[i], b = colT[i];
        bad = bad || a < 0 || a >= n || b < 0 || b >= n;
    }
    if (bad) atomicAdd(&ctl->count, 1ull);
}

// the control block as k_fb_tail expects it before a first step
__global__ void k_noalloc_ctl(Ctl *ctl, int32_t n) {
    uint32_t *w = (uint32_t *)ctl;
    for (uint32_t i = threadIdx.x; i < sizeof(Ctl) / 4; i += blockDim.x) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl->live = (uint32_t)n;
        ctl->lcnt[0] = (uint32_t)n;
        ctl->first_piv = 0xFFFFFFFFu;
        ctl->first_piv_prev = 0xFFFFFFFFu;
    }
}

}  // namespace
hipError_t preload_noalloc() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(k_noalloc_setup));
}

}  // namespace gpuscc

using namespace gpuscc;

extern "C" scc_status scc_run_noalloc(const int32_t *d_row, const int32_t *d_col, const int32_t *d_rowT,
                                      const int32_t *d_colT, int32_t n, int64_t m, int32_t *d_labels,
                                      int64_t *num_scc, void *stream) {
    if (n < 0 || m < 0) return SCC_E_INPUT;
    if (n > NA_MAX_N || m >= (int64_t)INT32_MAX) return SCC_E_OVERFLOW;
    if (num_scc) *num_scc = 0;
    if (n == 0) return SCC_OK;
    if (!d_row || !d_rowT || !d_labels || (m > 0 && (!d_col || !d_colT))) return SCC_E_INPUT;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return SCC_E_HIP;
    uint8_t *base = nullptr;
    if (scratch_of(dev, &base) != hipSuccess) return SCC_E_HIP;
    hipStream_t s = (hipStream_t)stream;
    // carve the array (device addresses from hipGetSymbolAddress, never the symbol)
    int32_t *vec[9];
    for (int k = 0; k < 9; k++) vec[k] = (int32_t *)(base + (size_t)k * NA_VEC);
    unsigned long long *tab0 = (unsigned long long *)(base + 9 * NA_VEC);
    unsigned long long *tab1 = (unsigned long long *)(base + 9 * NA_VEC + NA_TAB);
    Ctl *ctl = (Ctl *)(base + NA_CTL_OFF);
    int sb = 10;  // the election table size of the multi-kernel path for n vertices
    while (sb < 22 && ((int64_t)1 << sb) < (int64_t)n) sb++;
    const int64_t nslots = (int64_t)1 << sb;  // sb <= NA_MAX_SLOT_BITS: n <= 2^18
    // set-up with kernels only (no memset nodes for a caller that captures the stream)
    hipLaunchKernelGGL(k_noalloc_ctl, dim3(1), dim3(256), 0, s, ctl, n);
    const uint32_t grid = (uint32_t)std::min<int64_t>(1024, (std::max<int64_t>(n, m) + 255) / 256 + 1);
    hipLaunchKernelGGL(k_noalloc_setup, dim3(grid), dim3(256), 0, s, d_row, d_col, d_rowT, d_colT, n, m,
                       (uint32_t *)vec[0], vec[1], vec[6], ctl);
    const Launch lf{s, (uint32_t)std::min<int64_t>(256, nslots / 256), 256};
    launch_fill_u64(lf, tab0, nslots, ~0ull);
    launch_fill_u64(lf, tab1, nslots, ~0ull);
    Ctl h;
    if (hipMemcpyAsync(&h, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s) != hipSuccess) return SCC_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return SCC_E_HIP;
    if (h.count != 0) return SCC_E_INPUT;
    TailArgs t{};
    t.row = d_row;
    t.col = d_col;
    t.rowT = d_rowT;
    t.colT = d_colT;
    t.word = (uint32_t *)vec[0];
    t.root = vec[1];
    t.n = n;
    t.slots[0] = tab0;
    t.slots[1] = tab1;
    t.slot_bits = sb;
    t.prev_tab = 0;
    {
        std::lock_guard<std::mutex> lock(na_mutex);
        t.epoch = na_epoch[dev];
    }
    for (int k = 0; k < 4; k++) t.q[k] = vec[2 + k];
    t.list[0] = vec[6];
    t.list[1] = vec[7];
    t.lsel = 0;
    t.piv = vec[8];
    t.ctl = ctl;
    t.max_steps = (uint32_t)n + 2;
    t.xbits = 0;
    t.first_step = 1;
    t.self_refill = 1;
    launch_fb_tail(s, t);
    if (hipGetLastError() != hipSuccess) return SCC_E_HIP;
    if (hipMemcpyAsync(d_labels, t.root, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return SCC_E_HIP;
    if (hipMemcpyAsync(&h, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s) != hipSuccess) return SCC_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return SCC_E_HIP;
    {
        std::lock_guard<std::mutex> lock(na_mutex);
        na_epoch[dev] = t.epoch + h.tail[0];  // one election per step
    }
    if (h.tail[1] != 0) return SCC_E_STATE;  // a bound every correct run stays within was passed
    if (num_scc) *num_scc = (int64_t)h.pcount;
    return SCC_OK;
}
