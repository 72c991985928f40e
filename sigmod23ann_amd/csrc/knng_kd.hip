// Random-rotation KD splitter (SolutionRandomKD, src/SolutionRandomKD.hpp:258-428): the device
// parts of one round. The host draws the rotation, picks each node's dimension from the variances
// and drives the levels; these kernels project the points, reduce per-node statistics and write
// the side flags that the shared stable partition (knng_split.hip) consumes.
//
// Orders fixed here (the reference leaves them to TBB's combine order, SolutionRandomKD.hpp:
// 104-151): a node's FP64 sums run sequentially over each 1024-position tile of its range, from
// 0.0, and the tile partials are added in tile order from 0.0.
#include "knng_device.hpp"
#include "knng_kernels.hpp"

namespace knng {

constexpr uint32_t kKdRows = 8;  // rows projected per workgroup (32 threads per row)

// rot[i][c] = dot(u_c, x_i) for c < m (rotateRandom, SolutionRandomKD.hpp:268-273: the
// reference-order dot), rows of kKdStride floats, zero beyond m.
__global__ __launch_bounds__(256) void k_kd_rotate(const float* __restrict__ X, uint32_t ld,
                                                   uint32_t d, const float* __restrict__ R,
                                                   uint32_t m, uint32_t n, float* __restrict__ rot) {
    extern __shared__ __align__(16) float kd_lds[];
    float* sR = kd_lds;                    // kKdStride rows of ld floats
    float* sX = kd_lds + kKdStride * ld;   // kKdRows rows of ld floats
    const uint32_t r0 = blockIdx.x * kKdRows;
    for (uint32_t t = threadIdx.x; t < m * ld; t += blockDim.x) sR[t] = R[t];
    for (uint32_t t = threadIdx.x; t < kKdRows * ld; t += blockDim.x) {
        const uint32_t r = t / ld, j = t - r * ld;
        sX[t] = r0 + r < n ? X[(size_t)(r0 + r) * ld + j] : 0.f;
    }
    __syncthreads();
    const uint32_t r = threadIdx.x >> 5, c = threadIdx.x & 31;
    if (r0 + r >= n) return;
    rot[(size_t)(r0 + r) * kKdStride + c] = c < m ? dot_ref(sR + c * ld, sX + r * ld, d) : 0.f;
}

// part[tile][c] = sum over the tile's positions, in order, of x (mode 0) or (x - mean[node][c])^2
// (mode 1), x = rot[perm[pos]][c]; getMean / getVariance, SolutionRandomKD.hpp:104-151.
__global__ __launch_bounds__(64) void k_kd_tile_sums(uint32_t mode, const float* __restrict__ rot,
                                                     const uint32_t* __restrict__ perm,
                                                     const NodeDesc* __restrict__ nodes,
                                                     const TileDesc* __restrict__ tiles,
                                                     const double* __restrict__ mean, uint32_t m,
                                                     double* __restrict__ part) {
    const TileDesc tl = tiles[blockIdx.x];
    const uint32_t c = threadIdx.x;
    if (c >= m) return;
    const uint32_t base = nodes[tl.node].start + tl.begin;
    const double mu = mode ? mean[(size_t)tl.node * kKdStride + c] : 0.0;
    double s = 0.0;
    auto add = [&](float v) {
        const double x = (double)v;
        if (mode) {
            const double t = x - mu;
            s = s + t * t;
        } else {
            s = s + x;
        }
    };
    // the loads of 16 positions are issued before their adds, which stay in position order
    constexpr uint32_t kAhead = 16;
    uint32_t e = 0;
    for (; e + kAhead <= tl.count; e += kAhead) {
        float v[kAhead];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u) v[u] = rot[(size_t)perm[base + e + u] * kKdStride + c];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u) add(v[u]);
    }
    for (; e < tl.count; ++e) add(rot[(size_t)perm[base + e] * kKdStride + c]);
    part[(size_t)blockIdx.x * kKdStride + c] = s;
}

// out[node][c] = (tile partials added in order) / n (mode 0: the mean) or / (n - 1) (mode 1: the
// sample variance, SolutionRandomKD.hpp:146-148).
__global__ __launch_bounds__(64) void k_kd_node_finish(uint32_t mode, const NodeDesc* __restrict__ nodes,
                                                       const double* __restrict__ part, uint32_t m,
                                                       double* __restrict__ out) {
    const NodeDesc nd = nodes[blockIdx.x];
    const uint32_t c = threadIdx.x;
    if (c >= m) return;
    double s = 0.0;
    for (uint32_t t = 0; t < nd.ntiles; ++t) s = s + part[(size_t)(nd.tile0 + t) * kKdStride + c];
    out[(size_t)blockIdx.x * kKdStride + c] = s / (double)(mode ? nd.size - 1 : nd.size);
}

// flags[pos] = rot[perm[pos]][dim] < split (the float promoted to double, SolutionRandomKD.hpp:
// 338-339: true goes first); tile_lo = the tile's count of such positions.
__global__ __launch_bounds__(256) void k_kd_side(const float* __restrict__ rot,
                                                 const uint32_t* __restrict__ perm,
                                                 const NodeDesc* __restrict__ nodes,
                                                 const TileDesc* __restrict__ tiles,
                                                 const uint32_t* __restrict__ kdim,
                                                 const double* __restrict__ ksplit,
                                                 uint8_t* __restrict__ flags,
                                                 uint32_t* __restrict__ tile_lo) {
    const TileDesc tl = tiles[blockIdx.x];
    const NodeDesc nd = nodes[tl.node];
    __shared__ uint32_t nlo;
    if (threadIdx.x == 0) nlo = 0;
    __syncthreads();
    const uint32_t dim = kdim[tl.node];
    const double split = ksplit[tl.node];
    const uint32_t base = nd.start + tl.begin;
    uint32_t mylo = 0;
    for (uint32_t e = threadIdx.x; e < tl.count; e += blockDim.x) {
        const bool lo = (double)rot[(size_t)perm[base + e] * kKdStride + dim] < split;
        flags[base + e] = lo;
        mylo += lo;
    }
    atomicAdd(&nlo, mylo);
    __syncthreads();
    if (threadIdx.x == 0) tile_lo[blockIdx.x] = nlo;
}

size_t kd_rotate_lds(uint32_t ld) { return (size_t)(kKdStride + kKdRows) * ld * sizeof(float); }

void launch_kd_rotate(const float* X, uint32_t ld, uint32_t d, const float* R, uint32_t m,
                      uint32_t n, float* rot, hipStream_t st) {
    if (!n) return;
    const size_t lds = kd_rotate_lds(ld);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_kd_rotate),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_kd_rotate<<<(n + kKdRows - 1) / kKdRows, 256, lds, st>>>(X, ld, d, R, m, n, rot);
}

void launch_kd_stats(const float* rot, const uint32_t* perm, const NodeDesc* nodes, uint32_t nnodes,
                     const TileDesc* tiles, uint32_t ntiles, uint32_t m, double* part,
                     double* mean, double* var, hipStream_t st) {
    if (!nnodes) return;
    k_kd_tile_sums<<<ntiles, 64, 0, st>>>(0, rot, perm, nodes, tiles, nullptr, m, part);
    k_kd_node_finish<<<nnodes, 64, 0, st>>>(0, nodes, part, m, mean);
    k_kd_tile_sums<<<ntiles, 64, 0, st>>>(1, rot, perm, nodes, tiles, mean, m, part);
    k_kd_node_finish<<<nnodes, 64, 0, st>>>(1, nodes, part, m, var);
}

void launch_kd_side(const float* rot, const uint32_t* perm, const NodeDesc* nodes,
                    const TileDesc* tiles, uint32_t ntiles, const uint32_t* kdim,
                    const double* ksplit, uint8_t* flags, uint32_t* tile_lo, hipStream_t st) {
    if (ntiles) k_kd_side<<<ntiles, 256, 0, st>>>(rot, perm, nodes, tiles, kdim, ksplit, flags, tile_lo);
}

}  // namespace knng
