// Device-side building blocks shared by the gfx950 kernels of the KNN-graph builder.
//
// Arithmetic order matches the reference's AVX2 kernels (LinearAlgebra.hpp:53-77 distance,
// :179-200 dot): 8 lane accumulators, lane j takes dims j, j+8, ... below floor(d/8)*8 with
// fused multiply-add, the 8 lanes are added in order into 0.0f, then the tail dims are added with
// fused multiply-add. Sources are compiled with -ffp-contract=off so no other fusion happens.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace knng {

constexpr uint32_t kNoId = 0xFFFFFFFFu;
constexpr uint32_t kMaxAttempts = 8;     // split retries before the index-halving fallback
constexpr uint32_t kLargeChunk = 256;    // sample draws per FP64 partial sum (large branch)
constexpr uint32_t kMaxDim = 256;        // largest supported dimensionality
constexpr float kFltMax = 3.40282346638528859812e+38f;
constexpr float kFltMin = 1.17549435082228750797e-38f;

// ---- counter-based draws (replace thread_local mt19937 + uniform_int_distribution) ----------
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t round_key(uint64_t seed, uint32_t round) {
    return mix64(seed + 0x9E3779B97F4A7C15ull * (uint64_t)(round + 1));
}
__host__ __device__ inline uint64_t child_key(uint64_t key, uint32_t side) {
    return mix64(2 * key + 1 + side);
}
__host__ __device__ inline uint64_t attempt_key(uint64_t key, uint32_t attempt) {
    return mix64(key ^ (0xD1B54A32D192ED03ull * (uint64_t)(attempt + 1)));
}
// Uniform index in [0, n) for draw i of an attempt (multiply-shift of the high 32 bits).
__host__ __device__ inline uint32_t draw_below(uint64_t akey, uint32_t i, uint32_t n) {
    uint64_t h = mix64(akey + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1));
    return (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

// ---- reference-order arithmetic, one thread per row ------------------------------------------
// Rows are 16-byte aligned (row stride a multiple of 4 floats), padded with zeros past d.
__device__ inline float dot_ref(const float* __restrict__ a, const float* __restrict__ b,
                                uint32_t d) {
    const uint32_t main = d & ~7u;
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f, acc4 = 0.f, acc5 = 0.f, acc6 = 0.f,
          acc7 = 0.f;
    for (uint32_t i = 0; i < main; i += 8) {
        const float4 a0 = *reinterpret_cast<const float4*>(a + i);
        const float4 a1 = *reinterpret_cast<const float4*>(a + i + 4);
        const float4 b0 = *reinterpret_cast<const float4*>(b + i);
        const float4 b1 = *reinterpret_cast<const float4*>(b + i + 4);
        acc0 = __builtin_fmaf(a0.x, b0.x, acc0);
        acc1 = __builtin_fmaf(a0.y, b0.y, acc1);
        acc2 = __builtin_fmaf(a0.z, b0.z, acc2);
        acc3 = __builtin_fmaf(a0.w, b0.w, acc3);
        acc4 = __builtin_fmaf(a1.x, b1.x, acc4);
        acc5 = __builtin_fmaf(a1.y, b1.y, acc5);
        acc6 = __builtin_fmaf(a1.z, b1.z, acc6);
        acc7 = __builtin_fmaf(a1.w, b1.w, acc7);
    }
    float s = 0.0f;
    s = s + acc0; s = s + acc1; s = s + acc2; s = s + acc3;
    s = s + acc4; s = s + acc5; s = s + acc6; s = s + acc7;
    for (uint32_t i = main; i < d; ++i) s = __builtin_fmaf(a[i], b[i], s);
    return s;
}

__device__ inline float dist_ref(const float* __restrict__ a, const float* __restrict__ b,
                                 uint32_t d) {
    const uint32_t main = d & ~7u;
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f, acc4 = 0.f, acc5 = 0.f, acc6 = 0.f,
          acc7 = 0.f;
    for (uint32_t i = 0; i < main; i += 8) {
        const float4 a0 = *reinterpret_cast<const float4*>(a + i);
        const float4 a1 = *reinterpret_cast<const float4*>(a + i + 4);
        const float4 b0 = *reinterpret_cast<const float4*>(b + i);
        const float4 b1 = *reinterpret_cast<const float4*>(b + i + 4);
        float t;
        t = a0.x - b0.x; acc0 = __builtin_fmaf(t, t, acc0);
        t = a0.y - b0.y; acc1 = __builtin_fmaf(t, t, acc1);
        t = a0.z - b0.z; acc2 = __builtin_fmaf(t, t, acc2);
        t = a0.w - b0.w; acc3 = __builtin_fmaf(t, t, acc3);
        t = a1.x - b1.x; acc4 = __builtin_fmaf(t, t, acc4);
        t = a1.y - b1.y; acc5 = __builtin_fmaf(t, t, acc5);
        t = a1.z - b1.z; acc6 = __builtin_fmaf(t, t, acc6);
        t = a1.w - b1.w; acc7 = __builtin_fmaf(t, t, acc7);
    }
    float s = 0.0f;
    s = s + acc0; s = s + acc1; s = s + acc2; s = s + acc3;
    s = s + acc4; s = s + acc5; s = s + acc6; s = s + acc7;
    for (uint32_t i = main; i < d; ++i) {
        const float t = a[i] - b[i];
        s = __builtin_fmaf(t, t, s);
    }
    return s;
}

// Hyperplane between two centres (SolutionKmeans.hpp:162-164): between = 0.5f * (c1 + c2),
// coef = c1 - between, offset = dot(between, coef). Called by a whole block; `between` is block
// scratch of d floats; returns the offset to every thread.
__device__ inline float block_hyperplane(const float* __restrict__ c1,
                                         const float* __restrict__ c2, float* __restrict__ coef,
                                         float* __restrict__ between, uint32_t d, uint32_t ld,
                                         float* __restrict__ off_slot) {
    for (uint32_t j = threadIdx.x; j < ld; j += blockDim.x) {
        float b = 0.f, c = 0.f;
        if (j < d) {
            const float s = c1[j] + c2[j];
            b = 0.5f * s;
            c = c1[j] - b;
        }
        between[j] = b;
        coef[j] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) *off_slot = dot_ref(between, coef, d);
    __syncthreads();
    return *off_slot;
}

// ---- strict-weak (dist, id) order used by the sorted top-k lists -----------------------------
__device__ inline bool pair_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

// A point's top-k list held by one wave: lane l owns slots l and l + 64 (k <= 128), sorted
// ascending by (dist, id), empty slots (FLT_MAX, kNoId). `bound` is the k-th distance.
struct WaveList {
    float d0, d1;
    uint32_t i0, i1;
    float bound;

    __device__ inline void load(const uint32_t* gid, const float* gd, uint32_t k, uint32_t lane) {
        i0 = lane < k ? gid[lane] : kNoId;
        d0 = lane < k ? gd[lane] : kFltMax;
        i1 = lane + 64 < k ? gid[lane + 64] : kNoId;
        d1 = lane + 64 < k ? gd[lane + 64] : kFltMax;
        const uint32_t kl = k - 1;
        bound = __shfl(kl >= 64 ? d1 : d0, kl & 63);
    }
    __device__ inline void store(uint32_t* gid, float* gd, uint32_t k, uint32_t lane) const {
        if (lane < k) { gid[lane] = i0; gd[lane] = d0; }
        if (lane + 64 < k) { gid[lane + 64] = i1; gd[lane + 64] = d1; }
    }
    // Offer (nd, nid) to the list; wave-uniform arguments. The reference's filter is strict
    // `dist < bound` (KnnSets.hpp:152-157) and an id already present is rejected (:49, :59-72).
    __device__ inline bool insert(float nd, uint32_t nid, uint32_t k, uint32_t lane) {
        if (!(nd < bound)) return false;
        if (__ballot((lane < k && i0 == nid) || (lane + 64 < k && i1 == nid))) return false;
        const bool lt0 = lane < k && pair_less(d0, i0, nd, nid);
        const bool lt1 = lane + 64 < k && pair_less(d1, i1, nd, nid);
        const uint32_t pos = __popcll(__ballot(lt0)) + __popcll(__ballot(lt1));
        const float p0d = __shfl_up(d0, 1);
        const uint32_t p0i = __shfl_up(i0, 1);
        const float p1d_up = __shfl_up(d1, 1);
        const uint32_t p1i_up = __shfl_up(i1, 1);
        const float l63d = __shfl(d0, 63);
        const uint32_t l63i = __shfl(i0, 63);
        const float p1d = lane == 0 ? l63d : p1d_up;
        const uint32_t p1i = lane == 0 ? l63i : p1i_up;
        const uint32_t e0 = lane, e1 = lane + 64;
        if (e0 == pos) { d0 = nd; i0 = nid; }
        else if (e0 > pos) { d0 = p0d; i0 = p0i; }
        if (e1 == pos) { d1 = nd; i1 = nid; }
        else if (e1 > pos) { d1 = p1d; i1 = p1i; }
        const uint32_t kl = k - 1;
        bound = __shfl(kl >= 64 ? d1 : d0, kl & 63);
        return true;
    }
};

// ---- wave helpers and the 16-bit screening rows ----------------------------------------------
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Row stride of the 16-bit rows in LDS (elements); +8 staggers banks between rows.
__host__ __device__ inline uint32_t leaf_row_stride(uint32_t d) { return ((d + 15) & ~15u) + 8; }

__device__ inline uint16_t to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline uint32_t wave_excl_scan(uint32_t v, uint32_t lane) {
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    return x - v;
}

// Margin coefficient c of the exact screen: |approx - exact| <= c * (n_i + n_j).
// fp16: 2u + u^2 + accumulation (K<=128) + norm and final rounding, u = 2^-11 -> 9.9e-4; bf16:
// u = 2^-8 -> 7.9e-3. 25% slack on top.
template <uint32_t HALF>
struct HalfTraits;
template <>
struct HalfTraits<0> {
    static constexpr float kMargin = 1.25e-3f;
};
template <>
struct HalfTraits<1> {
    static constexpr float kMargin = 1.0e-2f;
};

// Converts one centred row (d <= 128; lane j holds dims j and j + 64) into the 16-bit screening
// layout: scaled by a power of two so that the largest |value| is below 2^14. Returns the FP32
// squared norm of the centred row and the inverse scale.
template <uint32_t HALF>
__device__ inline void screen_vals(float v0, float v1, uint32_t d, uint32_t rs,
                                   uint16_t* __restrict__ hrow, uint32_t lane, float& nrm,
                                   float& rinv) {
    const float m = wave_max(fmaxf(fabsf(v0), fabsf(v1)));
    nrm = wave_sum(__builtin_fmaf(v0, v0, v1 * v1));
    int e = 0;
    if (m > 0.f) frexpf(m, &e);
    const int sh = 14 - e;  // |v| * 2^sh < 2^14
    const float sc = ldexpf(1.0f, sh);
    v0 *= sc;
    v1 *= sc;
    uint16_t h0, h1;
    if (HALF == 0) {
        _Float16 f0 = (_Float16)v0, f1 = (_Float16)v1;
        h0 = __builtin_bit_cast(uint16_t, f0);
        h1 = __builtin_bit_cast(uint16_t, f1);
    } else {
        h0 = to_bf16_rne(v0);
        h1 = to_bf16_rne(v1);
    }
    hrow[lane] = lane < d ? h0 : 0;
    if (lane + 64 < rs) hrow[lane + 64] = lane + 64 < d ? h1 : 0;
    rinv = ldexpf(1.0f, -sh);
}

__device__ inline float4_t mfma16(half4_t a, half4_t b, float4_t c, uint32_t half_kind) {
    if (half_kind == 0) return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(short4_t, a),
                                                     __builtin_bit_cast(short4_t, b), c, 0, 0, 0);
}

}  // namespace knng
