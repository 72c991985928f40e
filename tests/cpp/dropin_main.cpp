// A caller written the way the reference's main.cpp calls its solution (src/main.cpp:51-63):
// stride-112 rows, a pre-sized vector<vector<uint32_t>>, one constructResult call, output.bin in
// SaveKNNG's format (src/io.h:21-33: n rows of k uint32, no header).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "knng_solution.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s points.bin output.bin [kd]\n", argv[0]);
        return 2;
    }
    const bool kd = argc > 3 && std::string(argv[3]) == "kd";  // the other line of main.cpp:58-59
    // ReadBinArray (src/io.h:66-94): uint32 N, then N x 100 float32 rows, into stride-112 rows
    std::ifstream in(argv[1], std::ios::binary);
    uint32_t n = 0;
    in.read(reinterpret_cast<char*>(&n), 4);
    auto* points = static_cast<float(*)[112]>(std::aligned_alloc(64, (size_t)n * 112 * sizeof(float)));
    for (uint32_t i = 0; i < n; ++i) in.read(reinterpret_cast<char*>(points[i]), 100 * sizeof(float));
    std::vector<std::vector<uint32_t>> knng(n);
    if (kd) {
        SolutionRandomKDMI355X::params().seed = 7;  // otherwise the reference's 5 rounds, leaves < 400
        SolutionRandomKDMI355X::constructResult(points, n, knng);
    } else {
        SolutionKmeansMI355X::params().k = 100;
        SolutionKmeansMI355X::params().leaf_max = 256;
        SolutionKmeansMI355X::params().rounds = 4;
        SolutionKmeansMI355X::params().non_passes = 1;
        SolutionKmeansMI355X::params().seed = 7;
        SolutionKmeansMI355X::constructResult(points, n, knng);
    }
    std::ofstream out(argv[2], std::ios::binary);
    for (auto& row : knng) out.write(reinterpret_cast<const char*>(row.data()), row.size() * 4);
    std::free(points);
    return 0;
}
