// Exact FP64 brute-force k nearest neighbours on the GPU (recall measurement).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace knng {

struct ExactScanArgs {
    const float* X;
    uint32_t ld, d, n;
    const double* norms;       // n, FP64 squared norms
    const uint32_t* queries;   // nq point ids
    uint32_t nq;
    const double* ub;          // nq upper bounds (inclusive), +inf allowed
    uint32_t cap;              // candidate slots per query
    uint32_t* count;           // nq, zeroed before the scan
    double* cand_d;            // nq * cap
    uint32_t* cand_id;         // nq * cap
};

void launch_norms64(const float* X, uint32_t ld, uint32_t d, uint32_t n, double* out,
                    hipStream_t st);
size_t exact_scan_lds(uint32_t d);
void launch_exact_scan(const ExactScanArgs& a, uint32_t slices, hipStream_t st);
void launch_exact_select(const ExactScanArgs& a, uint32_t k, uint32_t* out_ids, double* out_d,
                         uint32_t* overflow, hipStream_t st);
void launch_exact_scatter(const uint32_t* where, uint32_t nsub, uint32_t k, const uint32_t* sub_ids,
                          const double* sub_d, uint32_t* out_ids, double* out_d, hipStream_t st);
void launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* dst,
                       hipStream_t st);
void launch_gather_f64(const double* src, const uint32_t* idx, uint32_t n, double* dst,
                       hipStream_t st);
void launch_exact_ub(const ExactScanArgs& a, const uint32_t* lists, uint32_t k, double* ub,
                     hipStream_t st);

}  // namespace knng
