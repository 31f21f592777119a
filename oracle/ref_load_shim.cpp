// extern "C" entry to the reference's own loader (load.cpp:18-207, declared in
// load.h:5), so tests can compare this project's parser with it on the same file.
// Test infrastructure only; built into oracle/_ref/ from the reference's sources.
#include <cstdlib>
#include "load.h"

extern "C" int ref_load(const char *path, unsigned *n, unsigned *m, unsigned **nodes,
                        unsigned **edges, unsigned **nodesT, unsigned **edgesT) {
    return load(const_cast<char *>(path), false, n, m, nodes, edges, nodesT, edgesT, nullptr,
                nullptr) ? 1 : 0;
}

extern "C" void ref_free(void *p) { std::free(p); }
