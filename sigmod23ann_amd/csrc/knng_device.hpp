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

// Lane `src` (wave-uniform) of v, in every lane: v_readlane, a register read.
__device__ inline float lane_bcast(float v, uint32_t src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)src));
}
__device__ inline uint32_t lane_bcast(uint32_t v, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)src);
}
// Lane - 1 of v (lane 0 receives `fill`): DPP wave_shr:1, a register move with no LDS round trip.
__device__ inline uint32_t wave_shr1(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138, 0xF, 0xF, false);
}
__device__ inline float wave_shr1(float v, float fill) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138,
                                                      0xF, 0xF, false));
}

// A point's top-k list held by one wave: lane l owns slots l and l + 64 (k <= 128), sorted
// ascending by (dist, id); slots past the last filled one, and all slots >= k, hold the empty key.
// A slot is one 64-bit key, the distance's bits above the id: distances are non-negative, so
// unsigned key order is the (dist, id) order of pair_less and one 64-bit compare decides it. An
// insert is a few VALU, SALU, DPP and readlane operations with scalar branches only.
struct WaveList {
    static constexpr uint64_t kEmpty = ((uint64_t)0x7F7FFFFFu << 32) | 0xFFFFFFFFu;  // (FLT_MAX, kNoId)
    uint64_t k0, k1;
    float bound;  // k-th distance, wave-uniform

    __device__ static inline uint64_t key(float d, uint32_t id) {
        return ((uint64_t)__float_as_uint(d) << 32) | id;
    }
    __device__ inline void set_bound(uint32_t k) {
        const uint32_t kl = k - 1;
        const uint64_t kk = kl >= 64 ? k1 : k0;
        bound = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(kk >> 32), (int)(kl & 63)));
    }
    __device__ inline void load(const uint32_t* gid, const float* gd, uint32_t k, uint32_t lane) {
        k0 = lane < k ? key(gd[lane], gid[lane]) : kEmpty;
        k1 = lane + 64 < k ? key(gd[lane + 64], gid[lane + 64]) : kEmpty;
        set_bound(k);
    }
    __device__ inline void store(uint32_t* gid, float* gd, uint32_t k, uint32_t lane) const {
        if (lane < k) { gid[lane] = (uint32_t)k0; gd[lane] = __uint_as_float((uint32_t)(k0 >> 32)); }
        if (lane + 64 < k) { gid[lane + 64] = (uint32_t)k1; gd[lane + 64] = __uint_as_float((uint32_t)(k1 >> 32)); }
    }
    // Offer (nd, nid) to the list; wave-uniform arguments. The reference's filter is strict
    // `dist < bound` (KnnSets.hpp:152-157) and an id already present is rejected (:49, :59-72);
    // a present id carries the same distance, so its key equals the offered one.
    __device__ inline bool insert(float nd, uint32_t nid, uint32_t k, uint32_t lane) {
        if (!__builtin_amdgcn_readfirstlane(nd < bound)) return false;
        const uint64_t nk = key(nd, nid);
        if (__ballot(k0 == nk || k1 == nk)) return false;
        const uint32_t pos = __popcll(__ballot(k0 < nk)) + __popcll(__ballot(k1 < nk));
        // slot e takes slot e - 1's key for e > pos; slot 64 (lane 0's second) takes slot 63
        const uint32_t l63lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k0, 63);
        const uint32_t l63hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k0 >> 32), 63);
        const uint64_t s0 = ((uint64_t)wave_shr1((uint32_t)(k0 >> 32), 0u) << 32) | wave_shr1((uint32_t)k0, 0u);
        const uint64_t s1 = ((uint64_t)wave_shr1((uint32_t)(k1 >> 32), l63hi) << 32) | wave_shr1((uint32_t)k1, l63lo);
        const uint32_t e1 = lane + 64;
        k0 = lane == pos ? nk : (lane > pos ? s0 : k0);
        k1 = e1 >= k ? kEmpty : (e1 == pos ? nk : (e1 > pos ? s1 : k1));
        if (lane >= k) k0 = kEmpty;
        set_bound(k);
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
