// The drop-in process interface: ./scc | ./scc_w | ./scc_2_w  [-t] [-a q|t|d|l|h|a] [-b N] -f graph
// (README:1-9, scc.cu:57-148).  Same flags and the same stdout lines as the
// reference (load.cpp:35,48,121,123,166,204; scc.cu:111-147,237,913,998-1005;
// tarjan.cpp:193,237), with the FB run on the GPU through the C ABI.
// Extras: -j prints one JSON line per run; -v variant override (0/1/2); -V
// writes the reference's VERIFY outputs (scc.cu:920-995, a compile-time
// option there) from the canonical labels.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>

#include "gpuscc.h"
#include "host/graph.hpp"

#ifndef SCC_BIN_VARIANT
#define SCC_BIN_VARIANT 0
#endif

static const char *kStrategyName[] = {"QUADRATIC", "TOPO", "DATA", "LINEAR", "HYBRID", "COMBINE", "BEST"};
static const char kStrategyChar[] = "qtdlhcb";

static void usage() {
    std::printf("usage: ./scc [-t] [-a q/t/d/l/h/c/b/a] [-b block] -f graph\n"
                "\t-t tarjan's algorithm, optional.\n"
                "\t-a q for topology-driven, t for topology-driven with workgroup load balance,\n"
                "\t   d for data-driven, l for load-balance, h for hybrid, c for combine,\n"
                "\t   b for best (direction-optimising), a for all of them.\n"
                "\t   topology-driven with load balance (t) in default if not specific.\n"
                "\t-f input directed graph (edge list, MatrixMarket, or a binary graph written by -o)\n"
                "\t-o also write the loaded graph as a binary graph file (loads without parsing text)\n"
                "\t-j also print one JSON line per run\n"
                "\t-H 1: -a h switches as the reference's HYBRID (t until n/100 vertices are FB-eliminated,\n"
                "\t   then q); default 0: each level pushes or pulls by its frontier's size\n"
                "\t-V write trivial.txt, scc_dist.txt, parallel_scc.txt, state_range_root.txt\n");
}

// The VERIFY files of scc.cu:920-995 (1-based ids).  A class is represented by
// its minimum id (the canonical label), so the SCC lines come in that order;
// the reference orders them by pivot id.
static void write_verify(scc_ctx *ctx, const scc_host_graph &g, const std::vector<int32_t> &labels) {
    const int32_t n = g.n;
    std::vector<uint32_t> words((size_t)std::max<int32_t>(n, 1));
    std::vector<int32_t> roots((size_t)std::max<int32_t>(n, 1));
    if (scc_state_get(ctx, words.data(), roots.data()) != SCC_OK) {
        std::fprintf(stderr, "scc_state_get failed: %s\n", scc_last_error(ctx));
        return;
    }
    FILE *fp_trivial = std::fopen("trivial.txt", "w"), *fp_dist = std::fopen("scc_dist.txt", "w");
    FILE *fp_scc = std::fopen("parallel_scc.txt", "w"), *fp_detail = std::fopen("state_range_root.txt", "w");
    if (!fp_trivial || !fp_dist || !fp_scc || !fp_detail) {
        std::fprintf(stderr, "cannot write the VERIFY files\n");
        for (FILE *f : {fp_trivial, fp_dist, fp_scc, fp_detail})
            if (f) std::fclose(f);
        return;
    }
    std::printf("vector init done!\n");
    std::fprintf(fp_detail, "ID\t\tSTATE\tP  T  ROOT\t\tRANGE\tRANGE\n");
    std::vector<int32_t> size((size_t)std::max<int32_t>(n, 1), 0), first((size_t)n + 1, 0), next((size_t)n, -1);
    std::vector<int32_t> tail((size_t)std::max<int32_t>(n, 1), -1);
    int64_t number = 0;
    for (int32_t v = 0; v < n; v++) {
        uint32_t st = words[v] & 0xFFu, colour = words[v] >> 8;
        std::fprintf(fp_detail, "%d\t\t%u\t%u  %u  %d\t\t%u\t%u\n", v + 1, st, st & 16u, st & 8u, roots[v] + 1,
                     colour & 0xFFFFFu, colour);
        int32_t r = labels[v];
        if (tail[r] < 0) first[r] = v; else next[tail[r]] = v;  // members in increasing id
        tail[r] = v;
        size[r]++;
        if (r == v) number++;
    }
    std::fclose(fp_detail);
    std::printf("number=%lld\n", (long long)number);
    std::printf("sum=%d\n", n);
    std::vector<int64_t> dist((size_t)n + 1, 0);
    for (int32_t r = 0; r < n; r++) {
        if (labels[r] != r) continue;
        if (size[r] == 1) std::fprintf(fp_trivial, "%d\n", r + 1);
        dist[size[r]]++;
        for (int32_t v = first[r]; v >= 0; v = next[v]) std::fprintf(fp_scc, "%d ", v + 1);
        std::fprintf(fp_scc, "\n");
    }
    for (int32_t k = 1; k <= n; k++)
        if (dist[k]) std::fprintf(fp_dist, "%d %lld\n", k, (long long)dist[k]);
    std::fclose(fp_scc);
    std::fclose(fp_dist);
    std::fclose(fp_trivial);
    std::printf("sum=%d\n", n);
}

static int run_one(scc_ctx *ctx, int alg, int variant, const scc_host_graph &g, bool json,
                   const char *file, bool verify) {
    std::printf("computing SCC in parallel - alg: %d...\n", alg);
    scc_stats st;
    std::vector<int32_t> labels((size_t)std::max<int32_t>(g.n, 1));
    scc_status s = scc_run(ctx, alg, variant, SCC_PIVOT_PLAIN, labels.data(), &st);
    if (s != SCC_OK) {
        std::fprintf(stderr, "scc_run failed: %s (%s)\n", scc_status_string(s), scc_last_error(ctx));
        return 1;
    }
    if (st.first_pivot >= 0) std::printf("%lld not elimmed, set as first pivot\n", (long long)st.first_pivot + 1);
    std::printf("running %s strategy\n", kStrategyName[alg]);
    if (st.wcc_step > 0) std::printf("wcc done in step %lld\n", (long long)st.wcc_step);
    std::printf("FB done!\n");
    if (verify) write_verify(ctx, g, labels);
    double steps = st.steps > 0 ? (double)st.steps : 1.0;
    std::printf("\n");
    std::printf("    ->steps: %lld\n", (long long)st.steps);
    std::printf("    ->sum of fwdDepths: %lld and avg fwd depth: %.1f\n", (long long)st.fwd_depths,
                st.steps ? st.fwd_depths / steps : 0.0);
    std::printf("    ->sum of bwdDepths: %lld and avg bwd depth: %.1f\n", (long long)st.bwd_depths,
                st.steps ? st.bwd_depths / steps : 0.0);
    std::printf("    ->sum of trimIterations: %lld and avg of trimIterations: %.1f\n",
                (long long)st.trim_iters, st.steps ? st.trim_iters / steps : 0.0);
    std::printf("    ->time spent by fwd_reach: %f\n", st.ms_fwd / 1000.0);
    std::printf("    ->time spent by bwd_reach: %f\n", st.ms_bwd / 1000.0);
    std::printf("    ->time spent by trimming:  %f\n", st.ms_trim / 1000.0);
    std::printf("    ->time spent by updating:  %f\n", st.ms_update / 1000.0);
    std::printf("    ->count of SCC: %lld\n", (long long)st.num_scc);
    std::printf("    ->time: %f\n", st.ms_total / 1000.0);
    if (json) {
        double mteps = st.ms_total > 0 ? (double)g.m / (st.ms_total * 1e3) : 0.0;
        std::printf("{\"file\": \"%s\", \"variant\": %d, \"alg\": \"%c\", \"n\": %d, \"m\": %lld, "
                    "\"num_scc\": %lld, \"trivial\": %lld, \"biggest\": %lld, \"steps\": %lld, "
                    "\"fwd_depths\": %lld, \"bwd_depths\": %lld, \"trim_iters\": %lld, "
                    "\"wcc_step\": %lld, \"trim2_pairs\": %lld, \"ms\": %.4f, \"mteps\": %.2f}\n",
                    file, variant, kStrategyChar[alg], g.n, (long long)g.m, (long long)st.num_scc,
                    (long long)st.trivial, (long long)st.biggest, (long long)st.steps,
                    (long long)st.fwd_depths, (long long)st.bwd_depths, (long long)st.trim_iters,
                    (long long)st.wcc_step, (long long)st.trim2_pairs, st.ms_total, mteps);
    }
    return 0;
}

int main(int argc, char **argv) {
    bool tarjan = false, json = false, all = false, verify = false;
    const char *file = nullptr, *save_to = nullptr;
    int alg = SCC_SCHED_T;  // the reference's default (scc.cu:80)
    int variant = SCC_BIN_VARIANT;
    scc_opts opts;
    std::memset(&opts, 0, sizeof(opts));
    int c;
    while ((c = getopt(argc, argv, "htjVf:a:w:b:v:o:H:")) != -1) {
        switch (c) {
        case 'h': break;  // header flag: accepted, unused (as in load.cpp)
        case 't': tarjan = true; break;
        case 'j': json = true; break;
        case 'V': verify = true; break;
        case 'f': file = optarg; break;
        case 'a':
            switch (optarg[0]) {
            case 'q': alg = SCC_SCHED_Q; break;
            case 't': alg = SCC_SCHED_T; break;
            case 'd': alg = SCC_SCHED_D; break;
            case 'l': alg = SCC_SCHED_L; break;
            case 'h': alg = SCC_SCHED_H; break;
            case 'c': alg = SCC_SCHED_C; break;
            case 'b': alg = SCC_SCHED_B; break;
            case 'a': all = true; break;
            default:
                std::fprintf(stderr, "strategy '%c' is not provided (q, t, d, l, h, c, b, a)\n", optarg[0]);
                return 1;
            }
            break;
        case 'w': break;  // warp size: accepted, unused (never read by the reference either)
        case 'b': opts.block_size = (int32_t)std::strtol(optarg, nullptr, 0); break;
        case 'v': variant = std::atoi(optarg); break;
        case 'o': save_to = optarg; break;  // binary CSR file for later runs (no reference counterpart)
        case 'H': opts.hybrid_policy = (int32_t)std::atoi(optarg); break;  // 1: the reference's HYBRID rule
        default: usage(); return 1;
        }
    }
    if (!file) {
        usage();
        return 1;
    }
    std::printf("loading graph...\n");
    scc_host_graph g;
    std::string err;
    scc_status s = gpuscc::load_file(file, &g, &err);
    if (s != SCC_OK) {
        std::fprintf(stderr, "error: %s\n", err.c_str());
        return 1;
    }
    std::printf("%d %d\n", g.pattern, g.symmetric);
    if (g.mtx) std::printf("m=%d\tn=%d\tnnz=%lld\n", g.n, g.n, (long long)std::llround(g.average * g.n));
    std::printf("nodes=%d\tedges=%lld\n", g.n, (long long)g.m_reported);
    std::printf("OK\n");
    std::printf("OK\n");
    std::printf("max_in=%d min_in=%d max_out=%d min_out=%d average=%f\n", g.max_in, g.min_in, g.max_out,
                g.min_out, g.average);
    std::printf("    ->nodes: %u\n", (unsigned)g.n);
    std::printf("    ->edges: %u\n", (unsigned)g.m_reported);
    if (save_to) {
        s = gpuscc::save_graph(&g, save_to, &err);
        if (s != SCC_OK) {
            std::fprintf(stderr, "error: %s\n", err.c_str());
            gpuscc::graph_free(&g);
            return 1;
        }
        std::printf("saved binary graph to %s\n", save_to);
    }
    if (tarjan) {
        std::printf("tarjan decomposition...\n");
        gpuscc::TarjanResult r = gpuscc::tarjan_count(g.n, g.row, g.col);
        std::printf("Tarjan done!\n");
        std::printf("trivial = %lld\tnon_trivial=%lld\tnSCC = %lld\tbiggestSCC = %lld\n", (long long)r.trivial,
                    (long long)r.nontrivial, (long long)r.num_scc, (long long)r.biggest);
        std::printf("    ->count of SCC: %lld\n", (long long)r.num_scc);
        std::printf("    ->time: %f\n", r.seconds);
    }
    scc_ctx *ctx = nullptr;
    s = scc_create(0, &opts, &ctx);
    if (s != SCC_OK) {
        std::fprintf(stderr, "scc_create failed: %s\n", scc_status_string(s));
        gpuscc::graph_free(&g);
        return 1;
    }
    s = scc_upload(ctx, g.n, g.m, g.row, g.col, g.rowT, g.colT);
    int rc = 0;
    if (s != SCC_OK) {
        std::fprintf(stderr, "scc_upload failed: %s (%s)\n", scc_status_string(s), scc_last_error(ctx));
        rc = 1;
    } else if (all) {
        for (int a = SCC_SCHED_Q; a <= SCC_SCHED_B && rc == 0; a++)  // QUADRATIC..BEST (scc.cu:128-139)
            rc = run_one(ctx, a, variant, g, json, file, verify);
    } else {
        rc = run_one(ctx, alg, variant, g, json, file, verify);
    }
    scc_destroy(ctx);
    gpuscc::graph_free(&g);
    return rc;
}
