// The drop-in process interface: ./scc | ./scc_w | ./scc_2_w  [-t] [-a q|t|d|l|h|a] [-b N] -f graph
// (README:1-9, scc.cu:57-148).  Same flags and the same stdout lines as the
// reference (load.cpp:35,48,121,123,166,204; scc.cu:111-147,237,913,998-1005;
// tarjan.cpp:193,237), with the FB run on the GPU through the C ABI.
// Extras: -j prints one JSON line per run; -v variant override (0/1/2).
// The reference's COMBINE (-a c) and BEST (-a b) strategies are not provided.
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

static const char *kStrategyName[] = {"QUADRATIC", "TOPO", "DATA", "LINEAR", "HYBRID"};

static void usage() {
    std::printf("usage: ./scc [-t] [-a q/t/d/l/h/a] [-b block] -f graph\n"
                "\t-t tarjan's algorithm, optional.\n"
                "\t-a q for topology-driven, t for topology-driven with workgroup load balance,\n"
                "\t   d for data-driven, l for load-balance, h for hybrid, a for all of them.\n"
                "\t   topology-driven with load balance (t) in default if not specific.\n"
                "\t-f input directed graph\n"
                "\t-j also print one JSON line per run\n");
}

static int run_one(scc_ctx *ctx, int alg, int variant, const scc_host_graph &g, bool json,
                   const char *file) {
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
                    file, variant, "qtdlh"[alg], g.n, (long long)g.m, (long long)st.num_scc,
                    (long long)st.trivial, (long long)st.biggest, (long long)st.steps,
                    (long long)st.fwd_depths, (long long)st.bwd_depths, (long long)st.trim_iters,
                    (long long)st.wcc_step, (long long)st.trim2_pairs, st.ms_total, mteps);
    }
    return 0;
}

int main(int argc, char **argv) {
    bool tarjan = false, json = false, all = false;
    const char *file = nullptr;
    int alg = SCC_SCHED_T;  // the reference's default (scc.cu:80)
    int variant = SCC_BIN_VARIANT;
    scc_opts opts;
    std::memset(&opts, 0, sizeof(opts));
    int c;
    while ((c = getopt(argc, argv, "htjf:a:w:b:v:")) != -1) {
        switch (c) {
        case 'h': break;  // header flag: accepted, unused (as in load.cpp)
        case 't': tarjan = true; break;
        case 'j': json = true; break;
        case 'f': file = optarg; break;
        case 'a':
            switch (optarg[0]) {
            case 'q': alg = SCC_SCHED_Q; break;
            case 't': alg = SCC_SCHED_T; break;
            case 'd': alg = SCC_SCHED_D; break;
            case 'l': alg = SCC_SCHED_L; break;
            case 'h': alg = SCC_SCHED_H; break;
            case 'a': all = true; break;
            default:
                std::fprintf(stderr, "strategy '%c' is not provided (q, t, d, l, h, a)\n", optarg[0]);
                return 1;
            }
            break;
        case 'w': break;  // warp size: accepted, unused (never read by the reference either)
        case 'b': opts.block_size = (int32_t)std::strtol(optarg, nullptr, 0); break;
        case 'v': variant = std::atoi(optarg); break;
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
        for (int a = SCC_SCHED_Q; a <= SCC_SCHED_H && rc == 0; a++) rc = run_one(ctx, a, variant, g, json, file);
    } else {
        rc = run_one(ctx, alg, variant, g, json, file);
    }
    scc_destroy(ctx);
    gpuscc::graph_free(&g);
    return rc;
}
