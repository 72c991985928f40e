// Contest file formats (src/io.h): points are a uint32 N followed by N rows of 100 float32;
// a graph is N rows of 100 uint32 ids, no header.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace knng_io {

// Points of `dims` floats per row into rows of `stride` floats (the reference uses 112,
// src/io.h:77). Rows are read until EOF as ReadBinArray does (src/io.h:83-88), but never more
// than the header's N, and a file with fewer rows than N is an error (the reference would leave
// the rest uninitialised). Returns false with `err` set on failure.
inline bool read_points(const std::string& path, uint32_t dims, uint32_t stride,
                        std::vector<float>& rows, uint32_t& n, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    uint32_t hdr = 0;
    if (std::fread(&hdr, 4, 1, f) != 1) {
        std::fclose(f);
        err = "missing point count in " + path;
        return false;
    }
    rows.assign((size_t)hdr * stride, 0.f);
    uint32_t got = 0;
    std::vector<float> buf(dims);
    while (got < hdr && std::fread(buf.data(), 4, dims, f) == dims) {
        std::copy(buf.begin(), buf.end(), rows.begin() + (size_t)got * stride);
        ++got;
    }
    std::fclose(f);
    if (got < hdr) {
        err = path + ": header says " + std::to_string(hdr) + " points, file holds " + std::to_string(got);
        return false;
    }
    n = hdr;
    return true;
}

// SaveKNNG (src/io.h:21-33): n rows of k ids.
inline bool save_graph(const std::string& path, const uint32_t* ids, uint32_t n, uint32_t k,
                       std::string& err) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    const size_t tot = (size_t)n * k;
    const bool ok = std::fwrite(ids, 4, tot, f) == tot;
    std::fclose(f);
    if (!ok) err = "short write to " + path;
    return ok;
}

// grade.cpp's read (src/grade.cpp:20-38): rows of 100 uint32 until EOF, at most n of them;
// missing rows stay empty.
inline std::vector<std::vector<uint32_t>> read_graph(const std::string& path, uint32_t n,
                                                     uint32_t k = 100) {
    std::vector<std::vector<uint32_t>> g(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return g;
    std::vector<uint32_t> buf(k);
    for (uint32_t i = 0; i < n && std::fread(buf.data(), 4, k, f) == k; ++i) g[i] = buf;
    std::fclose(f);
    return g;
}

}  // namespace knng_io
