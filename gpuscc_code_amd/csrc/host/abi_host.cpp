// C ABI wrappers of the host input path (declared in include/gpuscc.h).
#include "gpuscc.h"
#include "host/graph.hpp"

extern "C" {

scc_status scc_load_file(const char *path, scc_host_graph *g) {
    return gpuscc::load_file(path, g, nullptr);
}

scc_status scc_save_graph(const scc_host_graph *g, const char *path) {
    return gpuscc::save_graph(g, path, nullptr);
}

scc_status scc_csr_from_arcs(int32_t n, int64_t m, const int32_t *src, const int32_t *dst,
                             int drop_self_loops, scc_host_graph *g) {
    return gpuscc::csr_from_arcs(n, m, src, dst, drop_self_loops, g);
}

scc_status scc_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                        int permute, scc_host_graph *g) {
    return gpuscc::gen_rmat(scale, edge_factor, a, b, c, seed, permute, g);
}

scc_status scc_gen_mesh(int side, double p_bidir, uint64_t seed, scc_host_graph *g) {
    return gpuscc::gen_mesh(side, p_bidir, seed, g);
}

scc_status scc_gen_batch(int num_graphs, int n_min, int n_max, double deg_lo, double deg_hi,
                         double cycle_p, double cycle_cover, uint64_t seed, scc_host_graph *g) {
    return gpuscc::gen_batch(num_graphs, n_min, n_max, deg_lo, deg_hi, cycle_p, cycle_cover, seed, g);
}

void scc_graph_free(scc_host_graph *g) { gpuscc::graph_free(g); }

}  // extern "C"
