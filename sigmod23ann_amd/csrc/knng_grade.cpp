// Recall grader with the reference's semantics and output (src/grade.cpp:40-98):
//   knng_grade <baseline.bin> <test.bin> <n>
// For every row, the number of baseline ids (duplicates counted each time) present in the test
// row; prints their mean over the n rows as "avg percent" (out of 100 for k = 100). Host-only.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unordered_set>
#include <vector>

#include "knng_io.hpp"

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: knng_grade <baseline.bin> <test.bin> <n>\n");
        return 2;
    }
    const std::string baseline = argv[1], test = argv[2];
    const long n = std::strtol(argv[3], nullptr, 10);
    if (n <= 0) {
        std::fprintf(stderr, "knng_grade: n must be positive\n");
        return 2;
    }
    std::printf("loaded baseline file: %s\n", baseline.c_str());
    std::printf("loaded test file: %s\n", test.c_str());
    std::printf("num points: %ld\n", n);
    const auto b = knng_io::read_graph(baseline, (uint32_t)n);
    const auto t = knng_io::read_graph(test, (uint32_t)n);
    uint64_t total = 0;
    for (long i = 0; i < n; ++i) {
        const std::unordered_set<uint32_t> ts(t[i].begin(), t[i].end());
        for (uint32_t id : b[i]) total += ts.count(id);
    }
    std::printf("avg percent: %g\n", (double)total / (double)n);  // iostream default: 6 digits
    return 0;
}
