// Seeded synthetic point sets, generated on the device (one wave per point).
//
//  low rank  x = z W + eps, z_i ~ N(0, 1/(1+i)) for i < latent, eps ~ N(0, noise^2), then scaled to
//            unit norm (the 10M x 100 embedding-like workload).
//  mixture   x = centre[c] + sigma[c] * N(0, I), c drawn from a cumulative weight table (uniform
//            weights for the 100k x 100 mixture, Zipf weights for the uneven-bucket workload).
// Normal draws are Box-Muller on counter-based uniforms keyed by (seed, point, draw index).
#include "knng_device.hpp"
#include "knng_kernels.hpp"

namespace knng {

__device__ inline float unit_open(uint64_t h) {  // (0, 1]
    return ((float)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);
}

__device__ inline float normal_draw(uint64_t seed, uint64_t point, uint32_t idx) {
    const uint64_t base = mix64(seed ^ mix64(point * 0x9E3779B97F4A7C15ull + idx));
    const float u1 = unit_open(mix64(base + 1));
    const float u2 = unit_open(mix64(base + 2));
    return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}

constexpr uint32_t kGenWaves = 4;

__global__ __launch_bounds__(kGenWaves * 64) void k_gen_lowrank(float* X, uint32_t n, uint32_t d,
                                                                 uint32_t ld, uint32_t latent,
                                                                 float noise, uint64_t seed,
                                                                 const float* W) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t p = (uint64_t)blockIdx.x * kGenWaves + (threadIdx.x >> 6);
    if (p >= n) return;
    const float z = lane < latent ? normal_draw(seed, p, lane) * rsqrtf(1.0f + (float)lane) : 0.f;
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t j = lane + 64 * c;
        float s = 0.f;
        if (j < d) {
            for (uint32_t i = 0; i < latent; ++i) s += __shfl(z, i) * W[(size_t)i * d + j];
            s += noise * normal_draw(seed, p, 64 + j);
        } else {
            for (uint32_t i = 0; i < latent; ++i) (void)__shfl(z, i);
        }
        v[c] = s;
    }
    float ss = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) ss += v[c] * v[c];
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
    const float inv = ss > 0.f ? 1.0f / sqrtf(ss) : 1.0f;
    float* x = X + p * ld;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t j = lane + 64 * c;
        if (j < ld) x[j] = j < d ? v[c] * inv : 0.f;
    }
}

__global__ __launch_bounds__(kGenWaves * 64) void k_gen_mixture(float* X, uint32_t n, uint32_t d,
                                                                 uint32_t ld, const float* centres,
                                                                 uint32_t nc, const float* cdf,
                                                                 const float* sigma, uint64_t seed) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t p = (uint64_t)blockIdx.x * kGenWaves + (threadIdx.x >> 6);
    if (p >= n) return;
    // cluster: first c with cdf[c] >= u (binary search; cdf[nc-1] == 1)
    const float u = unit_open(mix64(seed ^ mix64(p + 0x5851F42D4C957F2Dull)));
    uint32_t lo = 0, hi = nc - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] >= u) hi = mid;
        else lo = mid + 1;
    }
    const float sg = sigma[lo];
    const float* c = centres + (size_t)lo * d;
    float* x = X + p * ld;
    for (uint32_t j = lane; j < ld; j += 64)
        x[j] = j < d ? c[j] + sg * normal_draw(seed, p, j) : 0.f;
}

void launch_gen_lowrank(float* X, uint32_t n, uint32_t d, uint32_t ld, uint32_t latent,
                        float noise, uint64_t seed, const float* W, hipStream_t st) {
    if (n) k_gen_lowrank<<<(n + kGenWaves - 1) / kGenWaves, kGenWaves * 64, 0, st>>>(
        X, n, d, ld, latent, noise, seed, W);
}

void launch_gen_mixture(float* X, uint32_t n, uint32_t d, uint32_t ld, const float* centres,
                        uint32_t ncentres, const float* cdf, const float* sigma, uint64_t seed,
                        hipStream_t st) {
    if (n) k_gen_mixture<<<(n + kGenWaves - 1) / kGenWaves, kGenWaves * 64, 0, st>>>(
        X, n, d, ld, centres, ncentres, cdf, sigma, seed);
}

}  // namespace knng
