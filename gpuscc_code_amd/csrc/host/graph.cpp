// Host input path: the load.cpp-compatible parser, the CSR/CSR^T builder and
// the seeded generators for the benchmark graphs.  Plain C++17, no GPU.
//
// The reference builds CSR with one std::set per vertex (load.cpp:66-164),
// O(m log d) time and ~40 bytes per arc; here arcs are bucketed with a
// counting sort and each row is sorted and deduplicated in place, in parallel,
// so scale-24 R-MAT (268M arcs) loads in seconds.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gpuscc.h"
#include "host/graph.hpp"

namespace gpuscc {

int host_threads() {
    unsigned hc = std::thread::hardware_concurrency();
    const char *env = std::getenv("GPUSCC_HOST_THREADS");
    if (env && std::atoi(env) > 0) return std::atoi(env);
    if (hc == 0) hc = 1;
    return (int)std::min(hc, 16u);
}

template <class F>
static void parallel_for(int64_t total, F fn) {
    int nt = host_threads();
    if (total < 1 << 16 || nt <= 1) {
        fn(0, total, 0);
        return;
    }
    std::vector<std::thread> th;
    int64_t chunk = (total + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        int64_t b = t * chunk, e = std::min(total, b + chunk);
        if (b >= e) break;
        th.emplace_back([=] { fn(b, e, t); });
    }
    for (auto &x : th) x.join();
}

// -------------------------------------------------------------------------
// seeded counter-based random numbers (SplitMix64 finaliser): value j of
// stream s under seed is mix(mix(seed*G + s) + (j+1)*G).  Any draw can be
// computed independently, so generation is parallel and thread-count free.
// -------------------------------------------------------------------------
static constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t stream_key(uint64_t seed, uint64_t s) { return mix64(seed * kGolden + s); }
static inline uint64_t draw(uint64_t key, uint64_t j) { return mix64(key + (j + 1) * kGolden); }
static inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

uint64_t rng_draw(uint64_t seed, uint64_t stream, uint64_t j) { return draw(stream_key(seed, stream), j); }

// -------------------------------------------------------------------------
// CSR construction
// -------------------------------------------------------------------------
static void exclusive_scan(const std::vector<int64_t> &cnt, std::vector<int32_t> &row) {
    int64_t acc = 0;
    for (size_t i = 0; i < cnt.size(); i++) {
        row[i] = (int32_t)acc;
        acc += cnt[i];
    }
    row[cnt.size()] = (int32_t)acc;
}

// bucket arcs by source, then sort + dedup each row (optionally dropping self-loops)
static scc_status bucket_rows(int32_t n, int64_t m, const int32_t *src, const int32_t *dst,
                              int drop_self, std::vector<int32_t> &row, std::vector<int32_t> &col) {
    std::vector<std::atomic<int32_t>> cnt(n);
    for (int32_t v = 0; v < n; v++) cnt[v].store(0, std::memory_order_relaxed);
    std::atomic<int> bad{0};
    parallel_for(m, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; i++) {
            int32_t s = src[i], d = dst[i];
            if (s < 0 || s >= n || d < 0 || d >= n) { bad.store(1); continue; }
            cnt[s].fetch_add(1, std::memory_order_relaxed);
        }
    });
    if (bad.load()) return SCC_E_INPUT;
    std::vector<int32_t> start(n + 1);
    {
        int64_t acc = 0;
        for (int32_t v = 0; v < n; v++) {
            start[v] = (int32_t)acc;
            acc += cnt[v].load(std::memory_order_relaxed);
            cnt[v].store(start[v], std::memory_order_relaxed);
        }
        start[n] = (int32_t)acc;
    }
    std::vector<int32_t> tmp((size_t)std::max<int64_t>(m, 1));
    parallel_for(m, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; i++) tmp[cnt[src[i]].fetch_add(1, std::memory_order_relaxed)] = dst[i];
    });
    std::vector<int64_t> deg(n);
    parallel_for(n, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; v++) {
            int32_t *p = tmp.data() + start[v], *q = tmp.data() + start[v + 1];
            std::sort(p, q);
            int32_t *o = p;
            for (int32_t *x = p; x < q; x++) {
                if (drop_self && *x == (int32_t)v) continue;
                if (o > p && o[-1] == *x) continue;
                *o++ = *x;
            }
            deg[v] = o - p;
        }
    });
    row.assign((size_t)n + 1, 0);
    exclusive_scan(deg, row);
    col.resize((size_t)row[n]);
    parallel_for(n, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; v++)
            std::copy(tmp.data() + start[v], tmp.data() + start[v] + deg[v], col.data() + row[v]);
    });
    return SCC_OK;
}

static void transpose(int32_t n, const std::vector<int32_t> &row, const std::vector<int32_t> &col,
                      std::vector<int32_t> &rowT, std::vector<int32_t> &colT) {
    int64_t m = row[n];
    std::vector<std::atomic<int32_t>> cnt(n);
    for (int32_t v = 0; v < n; v++) cnt[v].store(0, std::memory_order_relaxed);
    parallel_for(m, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; i++) cnt[col[i]].fetch_add(1, std::memory_order_relaxed);
    });
    rowT.assign((size_t)n + 1, 0);
    int64_t acc = 0;
    for (int32_t v = 0; v < n; v++) {
        rowT[v] = (int32_t)acc;
        acc += cnt[v].load(std::memory_order_relaxed);
        cnt[v].store(rowT[v], std::memory_order_relaxed);
    }
    rowT[n] = (int32_t)acc;
    colT.resize((size_t)m);
    parallel_for(n, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; v++)
            for (int32_t i = row[v]; i < row[v + 1]; i++)
                colT[cnt[col[i]].fetch_add(1, std::memory_order_relaxed)] = (int32_t)v;
    });
    parallel_for(n, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e; v++) std::sort(colT.data() + rowT[v], colT.data() + rowT[v + 1]);
    });
}

static int32_t *to_malloc(const std::vector<int32_t> &v) {
    int32_t *p = (int32_t *)std::malloc(sizeof(int32_t) * std::max<size_t>(v.size(), 1));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(int32_t) * v.size());
    return p;
}

static scc_status finish_graph(int32_t n, std::vector<int32_t> &row, std::vector<int32_t> &col,
                               scc_host_graph *g) {
    std::vector<int32_t> rowT, colT;
    transpose(n, row, col, rowT, colT);
    g->n = n;
    g->m = row[n];
    g->row = to_malloc(row);
    g->col = to_malloc(col);
    g->rowT = to_malloc(rowT);
    g->colT = to_malloc(colT);
    if (!g->row || !g->col || !g->rowT || !g->colT) return SCC_E_OOM;
    int32_t mxo = 0, mno = n, mxi = 0, mni = n;
    for (int32_t v = 0; v < n; v++) {
        int32_t o = row[v + 1] - row[v], i = rowT[v + 1] - rowT[v];
        mxo = std::max(mxo, o); mno = std::min(mno, o);
        mxi = std::max(mxi, i); mni = std::min(mni, i);
    }
    g->max_out = mxo; g->min_out = mno; g->max_in = mxi; g->min_in = mni;
    if (g->m_reported == 0) g->m_reported = g->m;
    if (g->average == 0.0 && n > 0) g->average = (double)g->m_reported / n;
    return SCC_OK;
}

static void zero_graph(scc_host_graph *g) { std::memset(g, 0, sizeof(*g)); }

scc_status csr_from_arcs(int32_t n, int64_t m, const int32_t *src, const int32_t *dst,
                         int drop_self, scc_host_graph *g) {
    if (!g || n < 0 || m < 0) return SCC_E_INPUT;
    if (m >= (int64_t)INT32_MAX) return SCC_E_OVERFLOW;  // beyond int32 CSR offsets
    if (m > 0 && (!src || !dst)) return SCC_E_INPUT;
    zero_graph(g);
    std::vector<int32_t> row, col;
    scc_status s = bucket_rows(n, m, src, dst, drop_self, row, col);
    if (s != SCC_OK) return s;
    return finish_graph(n, row, col, g);
}

// -------------------------------------------------------------------------
// binary CSR files: a loaded graph saved once (save_graph) loads again without
// parsing text.  Little-endian layout: the 8-byte magic, int64 n, m,
// m_reported, int32 flags (1 pattern, 2 symmetric, 4 mtx), num_graphs, max_in,
// min_in, max_out, min_out, double average, then int32 row[n + 1], col[m] and,
// for a batch, graph_offsets[num_graphs + 1].  CSR^T is rebuilt on load.
// -------------------------------------------------------------------------
static constexpr char kBinMagic[8] = {'G', 'P', 'U', 'S', 'C', 'S', 'R', '1'};
static constexpr size_t kBinHeader = 64;

scc_status save_graph(const scc_host_graph *g, const char *path, std::string *err) {
    if (!g || !path || g->n < 0 || g->m < 0 || (g->n > 0 && !g->row) || (g->m > 0 && !g->col))
        return SCC_E_INPUT;
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        if (err) *err = std::string("cannot create ") + path;
        return SCC_E_INPUT;
    }
    unsigned char h[kBinHeader];
    std::memset(h, 0, sizeof(h));
    std::memcpy(h, kBinMagic, 8);
    const int64_t n64 = g->n, m64 = g->m, mr = g->m_reported;
    const int32_t fl = (g->pattern ? 1 : 0) | (g->symmetric ? 2 : 0) | (g->mtx ? 4 : 0);
    const int32_t ng = g->graph_offsets ? g->num_graphs : 0;
    const int32_t st4[4] = {g->max_in, g->min_in, g->max_out, g->min_out};
    std::memcpy(h + 8, &n64, 8);
    std::memcpy(h + 16, &m64, 8);
    std::memcpy(h + 24, &mr, 8);
    std::memcpy(h + 32, &fl, 4);
    std::memcpy(h + 36, &ng, 4);
    std::memcpy(h + 40, st4, 16);
    std::memcpy(h + 56, &g->average, 8);
    bool ok = std::fwrite(h, 1, sizeof(h), f) == sizeof(h);
    const int32_t zero = 0;
    const int32_t *rowp = g->n > 0 ? g->row : &zero;
    ok = ok && std::fwrite(rowp, sizeof(int32_t), (size_t)g->n + 1, f) == (size_t)g->n + 1;
    if (g->m > 0) ok = ok && std::fwrite(g->col, sizeof(int32_t), (size_t)g->m, f) == (size_t)g->m;
    if (ng > 0) ok = ok && std::fwrite(g->graph_offsets, sizeof(int32_t), (size_t)ng + 1, f) == (size_t)ng + 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok && err) *err = std::string("write failed: ") + path;
    return ok ? SCC_OK : SCC_E_INPUT;
}

// Checks everything the library assumes of a CSR (offsets from 0 to m, never
// decreasing; ids in range; rows strictly increasing), then rebuilds CSR^T.
static scc_status load_binary(const std::string &buf, scc_host_graph *g, std::string *err) {
    auto bad = [&](const char *msg) {
        if (err) *err = msg;
        return SCC_E_INPUT;
    };
    if (buf.size() < kBinHeader) return bad("binary graph: truncated header");
    const unsigned char *h = (const unsigned char *)buf.data();
    int64_t n64, m64, mr;
    int32_t fl, ng, st4[4];
    double avg;
    std::memcpy(&n64, h + 8, 8);
    std::memcpy(&m64, h + 16, 8);
    std::memcpy(&mr, h + 24, 8);
    std::memcpy(&fl, h + 32, 4);
    std::memcpy(&ng, h + 36, 4);
    std::memcpy(st4, h + 40, 16);
    std::memcpy(&avg, h + 56, 8);
    if (n64 < 0 || m64 < 0 || ng < 0) return bad("binary graph: negative sizes");
    if (n64 >= INT32_MAX - 1 || m64 >= INT32_MAX) {
        if (err) *err = "binary graph: sizes beyond int32 CSR";
        return SCC_E_OVERFLOW;
    }
    if (ng > n64 + 1) return bad("binary graph: more graphs than vertices");
    const int32_t n = (int32_t)n64;
    const size_t words = (size_t)n + 1 + (size_t)m64 + (ng > 0 ? (size_t)ng + 1 : 0);
    if (buf.size() != kBinHeader + 4 * words) return bad("binary graph: file size does not match its header");
    const int32_t *w = (const int32_t *)(buf.data() + kBinHeader);
    std::vector<int32_t> row(w, w + n + 1), col(w + n + 1, w + n + 1 + m64);
    if (row[0] != 0 || row[n] != m64) return bad("binary graph: row offsets do not span the arcs");
    // every offset in [0, m] and never decreasing, checked in full before any
    // column is read through an offset
    for (int32_t v = 0; v < n; v++)
        if (row[v + 1] < row[v] || row[v + 1] > m64) return bad("binary graph: row offsets out of range or decreasing");
    std::atomic<bool> fine{true};
    parallel_for(n, [&](int64_t b, int64_t e, int) {
        for (int64_t v = b; v < e && fine.load(std::memory_order_relaxed); v++) {
            for (int32_t i = row[v]; i < row[v + 1]; i++)
                if (col[i] < 0 || col[i] >= n || (i > row[v] && col[i] <= col[i - 1])) { fine = false; break; }
        }
    });
    if (!fine) return bad("binary graph: rows out of order, duplicated or out of range");
    std::vector<int32_t> off;
    if (ng > 0) {
        off.assign(w + n + 1 + m64, w + n + 1 + m64 + ng + 1);
        if (off[0] != 0 || off[ng] != n) return bad("binary graph: bad graph offsets");
        for (int32_t i = 0; i < ng; i++)
            if (off[i + 1] < off[i]) return bad("binary graph: bad graph offsets");
    }
    zero_graph(g);
    g->pattern = fl & 1;
    g->symmetric = (fl >> 1) & 1;
    g->mtx = (fl >> 2) & 1;
    g->m_reported = mr;
    g->average = avg;
    scc_status s = finish_graph(n, row, col, g);
    if (s != SCC_OK) return s;
    g->max_in = st4[0];
    g->min_in = st4[1];
    g->max_out = st4[2];
    g->min_out = st4[3];
    if (ng > 0) {
        g->num_graphs = ng;
        g->graph_offsets = to_malloc(off);
        if (!g->graph_offsets) return SCC_E_OOM;
    }
    return SCC_OK;
}

// -------------------------------------------------------------------------
// load.cpp-compatible parser (load.cpp:18-207)
// -------------------------------------------------------------------------
static inline bool is_space(char c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// One "%lld" of sscanf on [p, e): white space, an optional sign, at least one
// digit (values beyond the long long range clamp, as strtoll does).  Returns
// the position after the number, or null when there is none.
static const char *scan_ll(const char *p, const char *e, long long &out) {
    while (p < e && is_space(*p)) p++;
    bool neg = false;
    if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
    if (p >= e || *p < '0' || *p > '9') return nullptr;
    unsigned long long v = 0;
    bool big = false;
    for (; p < e && *p >= '0' && *p <= '9'; p++) {
        if (v > (ULLONG_MAX - 9) / 10) big = true;
        else v = v * 10 + (unsigned long long)(*p - '0');
    }
    if (big || v > (unsigned long long)LLONG_MAX) out = neg ? LLONG_MIN : LLONG_MAX;
    else out = neg ? -(long long)v : (long long)v;
    return p;
}

// The nnz entry lines that follow the header (load.cpp:70-92): line "a b" is
// the arc b -> a (1-based ids in [1, n]); with symmetric, a -> b as well unless
// a == b (counted into *reported, load.cpp:83-89).  The text from pos on is cut
// into chunks at line ends and parsed in parallel: a first pass counts each
// chunk's lines, so every line has its global index, and the first failing line
// (the one a sequential reader meets first) gives the error.
static scc_status parse_entries(const std::string &buf, size_t pos, int32_t n, int64_t nnz, bool symmetric,
                                std::vector<int32_t> &src, std::vector<int32_t> &dst, int64_t &reported,
                                std::string &msg) {
    const char *base = buf.data();
    const size_t beg = std::min(pos, buf.size()), end = buf.size();
    const size_t bytes = end - beg;
    const int nc = (int)std::max<size_t>(1, std::min<size_t>((size_t)host_threads(), bytes / 4096));
    std::vector<size_t> cb(nc + 1);
    cb[0] = beg;
    cb[nc] = end;
    for (int k = 1; k < nc; k++) {  // each cut just after a line end
        size_t x = std::max(beg + bytes * (size_t)k / (size_t)nc, cb[k - 1]);
        const void *nl = x < end ? std::memchr(base + x, '\n', end - x) : nullptr;
        cb[k] = nl ? (size_t)((const char *)nl - base) + 1 : end;
    }
    auto run = [&](auto &&fn) {  // fn(chunk) for every chunk, in parallel
        std::vector<std::thread> th;
        for (int k = 1; k < nc; k++) th.emplace_back([&, k] { fn(k); });
        fn(0);
        for (auto &t : th) t.join();
    };
    // pass 1: lines per chunk (a last line without its line end counts when non-empty)
    std::vector<int64_t> lines(nc + 1, 0);
    run([&](int k) {
        int64_t c = 0;
        for (const char *p = base + cb[k], *e = base + cb[k + 1]; p < e;) {
            const void *nl = std::memchr(p, '\n', (size_t)(e - p));
            c++;
            p = nl ? (const char *)nl + 1 : e;
        }
        lines[k + 1] = c;
    });
    for (int k = 0; k < nc; k++) lines[k + 1] += lines[k];  // lines[k]: index of chunk k's first line
    // pass 2: the entries; chunk k keeps its arcs and its first failing line
    std::vector<std::vector<int32_t>> cs(nc), cd(nc);
    std::vector<int64_t> err_line(nc, INT64_MAX), mirrored(nc, 0);
    std::vector<int> err_kind(nc, 0);
    run([&](int k) {
        int64_t li = lines[k];
        if (li >= nnz) return;
        const int64_t cap = std::min<int64_t>(nnz, lines[k + 1]) - li;
        cs[k].reserve((size_t)cap * (symmetric ? 2 : 1));
        cd[k].reserve((size_t)cap * (symmetric ? 2 : 1));
        for (const char *p = base + cb[k], *e = base + cb[k + 1]; p < e && li < nnz; li++) {
            const void *nl = std::memchr(p, '\n', (size_t)(e - p));
            const char *le = nl ? (const char *)nl : e;
            long long a = 0, b = 0;
            const char *q = scan_ll(p, le, a);
            if (q) q = scan_ll(q, le, b);
            if (!q) {
                err_line[k] = li;
                err_kind[k] = 1;
                return;
            }
            if (a < 1 || a > n || b < 1 || b > n) {
                err_line[k] = li;
                err_kind[k] = 2;
                return;
            }
            const int32_t d = (int32_t)(a - 1), s = (int32_t)(b - 1);
            cs[k].push_back(s);
            cd[k].push_back(d);
            if (symmetric && d != s) {
                cs[k].push_back(d);
                cd[k].push_back(s);
                mirrored[k]++;
            }
            p = nl ? le + 1 : e;
        }
    });
    int64_t first = INT64_MAX;
    int kind = 0;
    for (int k = 0; k < nc; k++)
        if (err_line[k] < first) {
            first = err_line[k];
            kind = err_kind[k];
        }
    if (first == INT64_MAX && lines[nc] < nnz) {
        msg = "file ends before nnz entries";
        return SCC_E_INPUT;
    }
    if (kind == 1) {
        msg = "bad entry line " + std::to_string(first + 1);
        return SCC_E_INPUT;
    }
    if (kind == 2) {
        msg = "vertex id out of range on entry line " + std::to_string(first + 1);
        return SCC_E_INPUT;
    }
    std::vector<size_t> off(nc + 1, 0);
    for (int k = 0; k < nc; k++) {
        off[k + 1] = off[k] + cs[k].size();
        reported += mirrored[k];
    }
    src.resize(off[nc]);
    dst.resize(off[nc]);
    run([&](int k) {
        if (!cs[k].empty()) {
            std::memcpy(src.data() + off[k], cs[k].data(), sizeof(int32_t) * cs[k].size());
            std::memcpy(dst.data() + off[k], cd[k].data(), sizeof(int32_t) * cd[k].size());
        }
    });
    return SCC_OK;
}
scc_status load_file(const char *path, scc_host_graph *g, std::string *err) {
    if (!g || !path) return SCC_E_INPUT;
    zero_graph(g);
    auto bad = [&](const std::string &msg, scc_status code) {
        if (err) *err = msg;
        return code;
    };
    FILE *f = std::fopen(path, "rb");
    if (!f) { return bad(std::string("cannot open ") + path, SCC_E_INPUT); }
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::string buf((size_t)std::max(sz, 0L), '\0');
    size_t rd = sz > 0 ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    buf.resize(rd);
    size_t pos = 0;
    auto next_line = [&](std::string &line) -> bool {
        if (pos >= buf.size()) return false;
        size_t e = buf.find('\n', pos);
        if (e == std::string::npos) e = buf.size();
        line.assign(buf, pos, e - pos);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        pos = e + 1;
        return true;
    };
    if (buf.size() >= sizeof(kBinMagic) && std::memcmp(buf.data(), kBinMagic, sizeof(kBinMagic)) == 0)
        return load_binary(buf, g, err);
    std::string line;
    if (!next_line(line)) { return bad("empty file", SCC_E_INPUT); }
    // load.cpp:30-33: the first line carries the MatrixMarket qualifiers
    bool pattern = line.find("pattern") != std::string::npos;
    bool symmetric = line.find("symmetric") != std::string::npos;
    // load.cpp:38-44: skip lines whose first character is '%'
    while (!line.empty() && line[0] == '%') {
        if (!next_line(line)) { return bad("no header line", SCC_E_INPUT); }
    }
    bool mtx = std::strstr(path, ".mtx") != nullptr;
    long long m_rows = 0, n_cols = 0, nnz = 0;
    if (mtx) {
        if (std::sscanf(line.c_str(), "%lld %lld %lld", &m_rows, &n_cols, &nnz) != 3) {
            return bad("bad MatrixMarket size line", SCC_E_INPUT);
        }
        if (m_rows != n_cols) { return bad("matrix is not square (m != n)", SCC_E_INPUT); }
    } else if (std::sscanf(line.c_str(), "%lld %lld", &m_rows, &nnz) != 2) {
        return bad("bad header line", SCC_E_INPUT);
    }
    if (m_rows < 0 || nnz < 0) { return bad("negative sizes", SCC_E_INPUT); }
    if (m_rows >= INT32_MAX || nnz >= INT32_MAX / 2) {
        return bad("sizes beyond int32 CSR", SCC_E_OVERFLOW);
    }
    int32_t n = (int32_t)m_rows;
    std::vector<int32_t> src, dst;
    int64_t reported = nnz;
    {
        std::string msg;
        scc_status es = parse_entries(buf, pos, n, (int64_t)nnz, symmetric, src, dst, reported, msg);
        if (es != SCC_OK) return bad(msg, es);
    }
    std::vector<int32_t> row, col;
    scc_status st = bucket_rows(n, (int64_t)src.size(), src.data(), dst.data(), 0, row, col);
    if (st != SCC_OK) return st;
    g->pattern = pattern;
    g->symmetric = symmetric;
    g->mtx = mtx;
    g->m_reported = reported;
    g->average = n > 0 ? (double)nnz / (double)n : 0.0;
    st = finish_graph(n, row, col, g);
    if (st == SCC_OK && symmetric) {  // load.cpp:102,146-164: in-degree stats only for non-symmetric
        g->max_in = 0;
        g->min_in = n;
    }
    return st;
}

// -------------------------------------------------------------------------
// generators
// -------------------------------------------------------------------------
// bijection on [0, 2^bits): odd multiplies and xor-shifts, all invertible mod 2^bits
static inline uint64_t permute_id(uint64_t x, int bits, uint64_t k1, uint64_t k2, uint64_t k3) {
    uint64_t mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
    int sh = bits / 2 + 1;
    x = (x ^ k3) & mask;
    x = (x * k1) & mask;
    x ^= x >> sh;
    x = (x * k2) & mask;
    x ^= x >> sh;
    x = (x * k1) & mask;
    return x;
}

scc_status gen_rmat(int scale, int ef, double a, double b, double c, uint64_t seed, int permute,
                    scc_host_graph *g) {
    if (!g || scale < 1 || scale > 30 || ef < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return SCC_E_INPUT;
    int64_t n = 1ll << scale;
    int64_t m = (int64_t)ef << scale;
    if (m >= INT32_MAX) return SCC_E_OVERFLOW;
    std::vector<int32_t> src((size_t)m), dst((size_t)m);
    const double ab = a + b, abc = a + b + c;
    uint64_t pk = stream_key(seed, ~0ull);
    uint64_t k1 = draw(pk, 0) | 1ull, k2 = draw(pk, 1) | 1ull, k3 = draw(pk, 2);
    parallel_for(m, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; i++) {
            uint64_t key = stream_key(seed, (uint64_t)i);
            uint64_t s = 0, d = 0;
            for (int l = 0; l < scale; l++) {
                double u = unit(draw(key, (uint64_t)l));
                uint64_t sb = 0, db = 0;
                if (u < a) {
                } else if (u < ab) {
                    db = 1;
                } else if (u < abc) {
                    sb = 1;
                } else {
                    sb = 1; db = 1;
                }
                s = (s << 1) | sb;
                d = (d << 1) | db;
            }
            if (permute) {
                s = permute_id(s, scale, k1, k2, k3);
                d = permute_id(d, scale, k1, k2, k3);
            }
            src[i] = (int32_t)s;
            dst[i] = (int32_t)d;
        }
    });
    return csr_from_arcs((int32_t)n, m, src.data(), dst.data(), 1, g);
}

scc_status gen_mesh(int side, double p, uint64_t seed, scc_host_graph *g) {
    if (!g || side < 1 || p < 0 || p > 1) return SCC_E_INPUT;
    if ((int64_t)side * side >= INT32_MAX) return SCC_E_OVERFLOW;
    int64_t n = (int64_t)side * side;
    int64_t horiz = (int64_t)side * (side - 1);
    int64_t L = 2 * horiz;
    std::vector<int32_t> src((size_t)(2 * L)), dst((size_t)(2 * L));
    std::vector<uint8_t> kind((size_t)std::max<int64_t>(L, 1));
    parallel_for(L, [&](int64_t lo, int64_t hi, int) {
        for (int64_t j = lo; j < hi; j++) {
            int64_t u, v;
            if (j < horiz) {
                int64_t r = j / (side - 1), cc = j % (side - 1);
                u = r * side + cc;
                v = u + 1;
            } else {
                int64_t jj = j - horiz;
                u = jj;  // r * side + c with r = jj / side, c = jj % side
                v = u + side;
            }
            uint64_t key = stream_key(seed, (uint64_t)j);
            uint8_t k;
            if (unit(draw(key, 0)) < p) k = 3;                 // both directions
            else k = (draw(key, 1) & 1) ? 1 : 2;               // u->v or v->u
            kind[j] = k;
            src[2 * j] = (int32_t)u; dst[2 * j] = (int32_t)v;
            src[2 * j + 1] = (int32_t)v; dst[2 * j + 1] = (int32_t)u;
        }
    });
    // compact the arcs that exist
    int64_t o = 0;
    for (int64_t j = 0; j < L; j++) {
        if (kind[j] & 1) { src[o] = src[2 * j]; dst[o] = dst[2 * j]; o++; }
        if (kind[j] & 2) { src[o] = src[2 * j + 1]; dst[o] = dst[2 * j + 1]; o++; }
    }
    return csr_from_arcs((int32_t)n, o, src.data(), dst.data(), 1, g);
}

scc_status gen_batch(int G, int nmin, int nmax, double dlo, double dhi, double cyc_p, double cover,
                     uint64_t seed, scc_host_graph *g) {
    if (!g || G < 1 || nmin < 2 || nmax < nmin || dlo < 0 || dhi < dlo || cyc_p <= 0 || cyc_p > 1 ||
        cover < 0 || cover > 1)
        return SCC_E_INPUT;
    std::vector<int64_t> nv(G), off(G + 1, 0);
    std::vector<double> deg(G);
    for (int i = 0; i < G; i++) {
        uint64_t key = stream_key(seed, (uint64_t)i);
        double u = unit(draw(key, 0));
        double x = std::exp(std::log((double)nmin) + u * (std::log((double)nmax) - std::log((double)nmin)));
        int64_t ni = (int64_t)std::llround(x);
        nv[i] = std::min<int64_t>(std::max<int64_t>(ni, nmin), nmax);
        deg[i] = dlo + unit(draw(key, 1)) * (dhi - dlo);
        off[i + 1] = off[i] + nv[i];
    }
    if (off[G] >= INT32_MAX) return SCC_E_OVERFLOW;
    // arcs per graph: ER arcs then cycle arcs (at most cover*n_i)
    std::vector<int64_t> aoff(G + 1, 0);
    for (int i = 0; i < G; i++) {
        int64_t mi = (int64_t)std::llround((double)nv[i] * deg[i]);
        aoff[i + 1] = aoff[i] + mi + (int64_t)(cover * (double)nv[i]) + 1;
    }
    if (aoff[G] >= INT32_MAX) return SCC_E_OVERFLOW;
    std::vector<int32_t> src((size_t)aoff[G], -1), dst((size_t)aoff[G], -1);
    std::atomic<int> next{0};
    int nt = host_threads();
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) {
        th.emplace_back([&] {
            std::vector<int32_t> perm;
            for (;;) {
                int i = next.fetch_add(1);
                if (i >= G) break;
                uint64_t ek = stream_key(seed ^ 0x5EEDull, (uint64_t)i);
                int64_t ni = nv[i], base = off[i], o = aoff[i];
                int64_t mi = (int64_t)std::llround((double)ni * deg[i]);
                for (int64_t j = 0; j < mi; j++) {
                    int64_t u = (int64_t)(unit(draw(ek, 2 * j)) * ni);
                    int64_t w = (u + 1 + (int64_t)(unit(draw(ek, 2 * j + 1)) * (ni - 1))) % ni;
                    src[o] = (int32_t)(base + u);
                    dst[o] = (int32_t)(base + w);
                    o++;
                }
                // planted cycles over a random order of the graph's vertices
                perm.resize(ni);
                for (int64_t v = 0; v < ni; v++) perm[v] = (int32_t)v;
                uint64_t pk = stream_key(seed ^ 0xC7C1Eull, (uint64_t)i);
                for (int64_t v = ni - 1; v > 0; v--) {
                    int64_t r = (int64_t)(draw(pk, (uint64_t)v) % (uint64_t)(v + 1));
                    std::swap(perm[v], perm[r]);
                }
                int64_t covered = (int64_t)(cover * (double)ni), at = 0;
                uint64_t ck = stream_key(seed ^ 0xCC11ull, (uint64_t)i);
                uint64_t cj = 0;
                while (at + 2 <= covered) {
                    double u = unit(draw(ck, cj++));
                    int64_t s = 1 + (int64_t)std::floor(std::log(1.0 - u) / std::log(1.0 - cyc_p));
                    if (s < 2) s = 2;
                    if (at + s > covered) s = covered - at;
                    if (s < 2) break;
                    for (int64_t k = 0; k < s; k++) {
                        src[o] = (int32_t)(base + perm[at + k]);
                        dst[o] = (int32_t)(base + perm[at + (k + 1) % s]);
                        o++;
                    }
                    at += s;
                }
            }
        });
    }
    for (auto &x : th) x.join();
    // compact unused slots
    int64_t o = 0;
    for (int64_t j = 0; j < aoff[G]; j++)
        if (src[j] >= 0) { src[o] = src[j]; dst[o] = dst[j]; o++; }
    scc_status s = csr_from_arcs((int32_t)off[G], o, src.data(), dst.data(), 1, g);
    if (s != SCC_OK) return s;
    g->num_graphs = G;
    g->graph_offsets = (int32_t *)std::malloc(sizeof(int32_t) * (G + 1));
    if (!g->graph_offsets) return SCC_E_OOM;
    for (int i = 0; i <= G; i++) g->graph_offsets[i] = (int32_t)off[i];
    return SCC_OK;
}

void graph_free(scc_host_graph *g) {
    if (!g) return;
    std::free(g->row); std::free(g->col); std::free(g->rowT); std::free(g->colT);
    std::free(g->graph_offsets);
    zero_graph(g);
}

}  // namespace gpuscc
