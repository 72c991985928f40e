// Drop-ins for the reference's solution structs (src/SolutionKmeans.hpp:46, :558;
// src/SolutionRandomKD.hpp:377): a caller written against
// `SolutionKmeans::constructResult(points, numPoints, result)` (src/main.cpp:59) compiles unchanged
// against `SolutionKmeansMI355X::constructResult` (or `SolutionRandomKDMI355X::constructResult`
// for the other line, :58) and gets the graph built on the GPU.
//
// Contract kept from the reference:
//  * points: numPoints rows of 112 floats (src/io.h:77), the first 100 used; read only, not freed.
//  * result: pre-sized to numPoints (src/main.cpp:57); each row receives exactly k ids, ascending by
//    distance, padded with id 1 when fewer than k neighbours were found (src/KnnSets.hpp:334-344).
// Differences, by design: the round count is a parameter instead of a wall-clock budget
// (src/SolutionKmeans.hpp:560-579), and a library error throws std::runtime_error instead of the
// reference's undefined behaviour on degenerate input.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "knng.h"

// Shared body of the drop-in solutions: build with P::params() and copy the ids out.
template <class P>
struct KnngSolutionMI355X {
    template <size_t Stride>
    static void constructResult(float points[][Stride], uint32_t numPoints,
                                std::vector<std::vector<uint32_t>>& result, uint32_t dims = 100) {
        const knng_params& p = P::params();
        knng_handle* h = nullptr;
        check(knng_create(&p, &h));
        std::vector<uint32_t> ids((size_t)numPoints * p.k);
        const int rc = knng_build(h, &points[0][0], numPoints, dims, Stride, ids.data(), nullptr);
        knng_destroy(h);
        check(rc);
        result.resize(numPoints);
        for (uint32_t i = 0; i < numPoints; ++i)
            result[i].assign(ids.begin() + (size_t)i * p.k, ids.begin() + (size_t)(i + 1) * p.k);
    }

   private:
    static void check(int rc) {
        if (rc != KNNG_OK) throw std::runtime_error(std::string("knng: ") + knng_last_error());
    }
};

struct SolutionKmeansMI355X : KnngSolutionMI355X<SolutionKmeansMI355X> {
    // Build parameters; defaults follow the reference (k = dims = 100, knnIterations = 1, the
    // 3000-point small branch) with leaf <= 400 and the rounds / NoN passes that reach
    // recall@100 >= 0.94 on 10M embedding-like points.
    static knng_params& params() {
        static knng_params p = [] {
            knng_params q;
            knng_default_params(&q);
            return q;
        }();
        return p;
    }
};

// Drop-in for SolutionRandomKD::constructResult (src/SolutionRandomKD.hpp:377, the other line of
// src/main.cpp:58-59): the random-rotation KD splitter, with the reference's 5 rounds, leaves of
// fewer than 400 points and no neighbours-of-neighbours pass (:385-417).
struct SolutionRandomKDMI355X : KnngSolutionMI355X<SolutionRandomKDMI355X> {
    static knng_params& params() {
        static knng_params p = [] {
            knng_params q;
            knng_default_params(&q);
            q.flags |= KNNG_FLAG_KD_SPLIT;
            q.rounds = 5;
            q.leaf_max = 399;
            q.leaf_max_odd = 0;
            q.non_passes = 0;
            return q;
        }();
        return p;
    }
};
