// The log tabulator of the reference (process_rst.cpp:7-41): reads the log of
// a batch of runs (the runall script prints each command line, then the run's
// output) and writes one row per graph: the graph file (the sixth field of a
// line holding "./scc") followed by the run's "->time:" values in groups of
// seven, as the reference does.  Same arguments: input log, output table.
//
//   process_rst LOG TABLE
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: process_rst LOG TABLE\n");
        return 1;
    }
    std::ifstream in(argv[1]);
    if (!in) {
        std::fprintf(stderr, "error: cannot open %s\n", argv[1]);
        return 1;
    }
    std::ofstream out(argv[2]);
    if (!out) {
        std::fprintf(stderr, "error: cannot create %s\n", argv[2]);
        return 1;
    }
    constexpr int kGroup = 7;  // values per row (process_rst.cpp:17,34)
    float times[kGroup] = {0};
    float t = 0.f;  // a "time:" line without a number repeats the value before
    long count = 0;
    std::string line, name;
    while (std::getline(in, line)) {
        if (line.find("./scc") != std::string::npos) {
            // the sixth whitespace-separated field ("./scc -t -a a -f FILE");
            // with fewer fields, the last one
            std::istringstream f(line);
            std::string tok;
            for (int k = 0; k < 6 && (f >> tok); k++) name = tok;
            out << name << ' ';
        }
        if (line.find("time:") != std::string::npos) {
            std::istringstream f(line);
            std::string first;
            float v;
            if ((f >> first) && (f >> v)) t = v;
            times[count++ % kGroup] = t;
            if (count % kGroup == 0) {
                for (int k = 0; k < kGroup; k++) out << times[k] << ' ';
                out << std::endl;
            }
        }
    }
    return 0;
}
