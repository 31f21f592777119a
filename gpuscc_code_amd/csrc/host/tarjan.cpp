// The CLI's -t option: sequential Tarjan SCC count on the host CSR, printed
// beside the GPU count as the reference does (scc.cu:119-125, tarjan.cpp:72-240).
// It never feeds the GPU path.  Iterative, with an explicit frame stack so deep
// graphs (the 2048x2048 mesh) do not overflow the call stack.
#include <chrono>
#include <vector>

#include "host/graph.hpp"

namespace gpuscc {

TarjanResult tarjan_count(int32_t n, const int32_t *row, const int32_t *col) {
    struct Frame { int32_t v, next; };
    std::vector<int32_t> order(n, -1), low(n, 0), comp_stack;
    std::vector<uint8_t> on_stack(n, 0);
    std::vector<Frame> frames;
    comp_stack.reserve(n);
    TarjanResult r{0, 0, 0, 0, 0.0};
    int32_t clock = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int32_t root = 0; root < n; root++) {
        if (order[root] >= 0) continue;
        order[root] = low[root] = clock++;
        comp_stack.push_back(root);
        on_stack[root] = 1;
        frames.push_back({root, row[root]});
        while (!frames.empty()) {
            Frame &f = frames.back();
            if (f.next < row[f.v + 1]) {
                int32_t w = col[f.next++];
                if (order[w] < 0) {
                    order[w] = low[w] = clock++;
                    comp_stack.push_back(w);
                    on_stack[w] = 1;
                    frames.push_back({w, row[w]});
                } else if (on_stack[w] && order[w] < low[f.v]) {
                    low[f.v] = order[w];
                }
                continue;
            }
            int32_t v = f.v;
            frames.pop_back();
            if (!frames.empty() && low[v] < low[frames.back().v]) low[frames.back().v] = low[v];
            if (low[v] != order[v]) continue;
            int64_t size = 0;
            for (;;) {
                int32_t x = comp_stack.back();
                comp_stack.pop_back();
                on_stack[x] = 0;
                size++;
                if (x == v) break;
            }
            r.num_scc++;
            if (size == 1) r.trivial++; else r.nontrivial++;
            if (size > r.biggest) r.biggest = size;
        }
    }
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

}  // namespace gpuscc
