// Internal C++ interface of the host input path (graph.cpp) and the CLI's
// Tarjan (tarjan.cpp).  The C ABI wrappers live in abi_host.cpp.
#pragma once
#include <cstdint>
#include <string>

#include "gpuscc.h"

namespace gpuscc {

int host_threads();
uint64_t rng_draw(uint64_t seed, uint64_t stream, uint64_t j);
scc_status csr_from_arcs(int32_t n, int64_t m, const int32_t *src, const int32_t *dst,
                         int drop_self, scc_host_graph *g);
scc_status load_file(const char *path, scc_host_graph *g, std::string *err);
scc_status save_graph(const scc_host_graph *g, const char *path, std::string *err);
scc_status gen_rmat(int scale, int ef, double a, double b, double c, uint64_t seed, int permute,
                    scc_host_graph *g);
scc_status gen_mesh(int side, double p, uint64_t seed, scc_host_graph *g);
scc_status gen_batch(int G, int nmin, int nmax, double dlo, double dhi, double cyc_p,
                     double cover, uint64_t seed, scc_host_graph *g);
void graph_free(scc_host_graph *g);

struct TarjanResult {
    int64_t num_scc, trivial, nontrivial, biggest;
    double seconds;  // DFS only
};
// The CLI's -t option (tarjan.cpp:72-240): sequential SCC count on the host CSR.
TarjanResult tarjan_count(int32_t n, const int32_t *row, const int32_t *col);

}  // namespace gpuscc
