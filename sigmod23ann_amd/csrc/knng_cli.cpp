// Command-line build with the reference's file formats and flow (src/main.cpp:39-71):
//   knng_cli [source.bin] [--out output.bin] [--k 100] [--rounds 6] [--non 2] [--leaf 400]
//            [--iterations 1] [--seed S] [--precision fp32|fp16|bf16] [--search-k W] [--truth]
// Reads the points (default dummy-data.bin) into rows of 112 floats as ReadBinArray does, builds
// the graph through the drop-in entry point (knng_build), and writes n rows of k ids
// (default output.bin, SaveKNNG). --truth writes the exact FP64 k-NN instead: the approximate
// graph's k-th distances bound an exact brute force on the GPU (a truth file for knng_grade).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/knng.h"
#include "knng_io.hpp"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t) {
    return std::chrono::duration<double, std::milli>(clk::now() - t).count();
}

int die(const char* what) {
    std::fprintf(stderr, "knng_cli: %s: %s\n", what, knng_last_error());
    return 1;
}

// Exact k-NN of every point, as ids: device brute force in FP64, bounded per query by the
// approximate graph's k-th distance.
int exact_graph(knng_handle* h, const std::vector<float>& rows, uint32_t n, uint32_t d, uint32_t stride,
                uint32_t k, const std::vector<uint32_t>& approx, std::vector<uint32_t>& out) {
    const uint32_t ld = (d + 3) & ~3u;
    std::vector<float> packed((size_t)n * ld, 0.f);
    for (uint32_t i = 0; i < n; ++i)
        std::memcpy(&packed[(size_t)i * ld], &rows[(size_t)i * stride], d * sizeof(float));
    void *dX = nullptr, *dq = nullptr, *dl = nullptr, *dub = nullptr, *dti = nullptr, *dtd = nullptr;
    const uint32_t batch = std::min<uint32_t>(n, 100000);
    int rc = KNNG_OK;
    auto chk = [&](int r) {
        if (rc == KNNG_OK) rc = r;
        return rc == KNNG_OK;
    };
    if (chk(knng_device_alloc(packed.size() * 4, &dX)) && chk(knng_memcpy_h2d(dX, packed.data(), packed.size() * 4)) &&
        chk(knng_device_alloc((size_t)batch * 4, &dq)) && chk(knng_device_alloc((size_t)batch * k * 4, &dl)) &&
        chk(knng_device_alloc((size_t)batch * 8, &dub)) && chk(knng_device_alloc((size_t)batch * k * 4, &dti)) &&
        chk(knng_device_alloc((size_t)batch * k * 8, &dtd))) {
        out.assign((size_t)n * k, 0);
        std::vector<uint32_t> q(batch);
        for (uint32_t q0 = 0; q0 < n && rc == KNNG_OK; q0 += batch) {
            const uint32_t nq = std::min(batch, n - q0);
            for (uint32_t i = 0; i < nq; ++i) q[i] = q0 + i;
            chk(knng_memcpy_h2d(dq, q.data(), (size_t)nq * 4)) &&
                chk(knng_memcpy_h2d(dl, &approx[(size_t)q0 * k], (size_t)nq * k * 4)) &&
                chk(knng_exact_upper_from_lists(h, (const float*)dX, n, d, ld, (const uint32_t*)dq, nq,
                                                (const uint32_t*)dl, k, (double*)dub)) &&
                chk(knng_exact_knn_device(h, (const float*)dX, n, d, ld, (const uint32_t*)dq, nq, k,
                                          (const double*)dub, 4 * k + 256, (uint32_t*)dti, (double*)dtd)) &&
                chk(knng_memcpy_d2h(&out[(size_t)q0 * k], dti, (size_t)nq * k * 4));
        }
    }
    for (void* p : {dX, dq, dl, dub, dti, dtd})
        if (p) knng_device_free(p);
    return rc;
}

}  // namespace

int main(int argc, char** argv) {
    const auto t_start = clk::now();
    std::string src = "dummy-data.bin", out = "output.bin";
    knng_params p;
    knng_default_params(&p);
    p.non_passes = 2;
    uint32_t dims = 100;
    bool truth = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "knng_cli: %s needs a value\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        auto num = [&]() { return (uint32_t)std::strtoul(next(), nullptr, 10); };
        if (a == "--out") out = next();
        else if (a == "--k") p.k = num();
        else if (a == "--rounds") p.rounds = num();
        else if (a == "--non") p.non_passes = num();
        else if (a == "--leaf") p.leaf_max = num();
        else if (a == "--iterations") p.iterations = num();
        else if (a == "--search-k") p.search_k = num();
        else if (a == "--dims") dims = num();
        else if (a == "--seed") p.seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--truth") truth = true;
        else if (a == "--precision") {
            const std::string v = next();
            p.precision = v == "bf16" ? KNNG_PREC_BF16 : v == "fp16" ? KNNG_PREC_FP16 : KNNG_PREC_FP32;
        } else if (!a.empty() && a[0] == '-') {
            std::fprintf(stderr, "knng_cli: unknown option %s\n", a.c_str());
            return 2;
        } else {
            src = a;
        }
    }
    const uint32_t stride = dims <= 112 ? 112 : (dims + 3) & ~3u;

    std::printf("Reading Data: %s\n", src.c_str());
    const auto t_read = clk::now();
    std::vector<float> rows;
    uint32_t n = 0;
    std::string err;
    if (!knng_io::read_points(src, dims, stride, rows, n, err)) {
        std::fprintf(stderr, "knng_cli: %s\n", err.c_str());
        return 1;
    }
    std::printf("# of points: %u\nFinish Reading Data\nread time: %.0f\n", n, ms_since(t_read));

    const auto t_build = clk::now();
    knng_handle* h = nullptr;
    if (knng_create(&p, &h) != KNNG_OK) return die("create");
    std::vector<uint32_t> ids((size_t)n * p.k);
    if (knng_build(h, rows.data(), n, dims, stride, ids.data(), nullptr) != KNNG_OK) return die("build");
    if (truth) {
        std::vector<uint32_t> exact;
        if (exact_graph(h, rows, n, dims, stride, p.k, ids, exact) != KNNG_OK) return die("exact k-NN");
        ids.swap(exact);
    }
    knng_destroy(h);
    std::printf("build time: %.0f\n", ms_since(t_build));

    const auto t_save = clk::now();
    std::printf("Saving KNN Graph (%u X %u) to %s\n", n, p.k, out.c_str());
    if (!knng_io::save_graph(out, ids.data(), n, p.k, err)) {
        std::fprintf(stderr, "knng_cli: %s\n", err.c_str());
        return 1;
    }
    std::printf("save time: %.0f\ntotal time (ms): %.0f\n", ms_since(t_save), ms_since(t_start));
    return 0;
}
