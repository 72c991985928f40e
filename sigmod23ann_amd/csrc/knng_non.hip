// Neighbours-of-neighbours pass (topUpSingle, SolutionKmeans.hpp:480-555) for gfx950.
//
// The snapshot copies every point's k id slots (unfilled slots read as id 0, as the reference's
// zero-initialised current_ids do, KnnSets.hpp:29). Each point p is then offered every id in the
// snapshot lists of its snapshot neighbours except itself; the reference dedupes candidates in a
// hash set first, here duplicates reach the list insert, whose id check rejects them (same pair,
// same exact distance, so the result is the same set).
#include "knng_device.hpp"
#include "knng_kernels.hpp"

namespace knng {

__global__ void k_non_snapshot(const uint32_t* __restrict__ ids, uint32_t* __restrict__ snap,
                               size_t total) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) {
        const uint32_t v = ids[i];
        snap[i] = v == kNoId ? 0u : v;
    }
}

void launch_non_snapshot(const uint32_t* ids, uint32_t* snap, uint32_t n, uint32_t k,
                         hipStream_t st) {
    const size_t tot = (size_t)n * k;
    if (tot) k_non_snapshot<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(ids, snap, tot);
}

// One wave per point: for each snapshot neighbour q, lanes take q's snapshot ids, compute the exact
// reference-order distance and offer the ones under the live bound.
constexpr uint32_t kNonWaves = 4;

__global__ __launch_bounds__(kNonWaves * 64) void k_non_direct(NonArgs a) {
    __shared__ __align__(16) float rowbuf[kNonWaves][kMaxDim];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t p = blockIdx.x * kNonWaves + wave;
    if (p >= a.n) return;
    const uint32_t k = a.k, ld = a.ld, d = a.d;
    float* xr = rowbuf[wave];
    for (uint32_t j = lane; j < ld; j += 64) xr[j] = a.X[(size_t)p * ld + j];
    uint32_t* gid = a.knn_ids + (size_t)p * k;
    float* gd = a.knn_d + (size_t)p * k;
    WaveList lst;
    lst.load(gid, gd, k, lane);
    bool changed = false;
    const uint32_t* sp = a.snap + (size_t)p * k;
    for (uint32_t a2 = 0; a2 < k; ++a2) {
        const uint32_t q = sp[a2];
        const uint32_t* sq = a.snap + (size_t)q * k;
        for (uint32_t c0 = 0; c0 < k; c0 += 64) {
            float dd = kFltMax;
            uint32_t cid = kNoId;
            if (c0 + lane < k) {
                cid = sq[c0 + lane];
                if (cid != p) dd = dist_ref(a.X + (size_t)cid * ld, xr, d);
            }
            uint64_t acc = __ballot(dd < lst.bound);
            while (acc) {
                const uint32_t src = __ffsll((unsigned long long)acc) - 1;
                acc &= acc - 1;
                const float nd = __shfl(dd, src);
                const uint32_t nid = __shfl(cid, src);
                changed |= lst.insert(nd, nid, k, lane);
            }
        }
    }
    if (changed) {
        lst.store(gid, gd, k, lane);
        if (lane == 0) a.bound[p] = lst.bound;
    }
}

void launch_non_direct(const NonArgs& a, hipStream_t st) {
    if (a.n) k_non_direct<<<(a.n + kNonWaves - 1) / kNonWaves, kNonWaves * 64, 0, st>>>(a);
}

}  // namespace knng
