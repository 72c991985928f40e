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
constexpr uint32_t kNewBit = 0x80000000u;  // NoN snapshot entry not in the previous pass's snapshot
constexpr uint32_t kExpBit = 0x40000000u;  // NoN snapshot entry whose neighbours are expanded
constexpr uint32_t kIdMask = 0x3FFFFFFFu;
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
#pragma unroll 4
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
#pragma unroll 4
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

// dist_ref (the reference's order: eight accumulators over dims j, j+8, ... below floor(d/8)*8,
// summed in order into 0.0f, then the tail with fused multiply-add) for a compile-time width D,
// fully unrolled so that every LDS load of the two rows is issued at once. Rows are 16-byte
// aligned.
template <uint32_t D>
__device__ inline float dist_ref_fixed(const float* __restrict__ a, const float* __restrict__ b) {
    constexpr uint32_t kMain = D & ~7u;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (uint32_t i = 0; i < kMain; i += 4) {
        const float4 x = *reinterpret_cast<const float4*>(a + i);
        const float4 y = *reinterpret_cast<const float4*>(b + i);
        const uint32_t j = i & 4u;
        float t;
        t = x.x - y.x; acc[j + 0] = __builtin_fmaf(t, t, acc[j + 0]);
        t = x.y - y.y; acc[j + 1] = __builtin_fmaf(t, t, acc[j + 1]);
        t = x.z - y.z; acc[j + 2] = __builtin_fmaf(t, t, acc[j + 2]);
        t = x.w - y.w; acc[j + 3] = __builtin_fmaf(t, t, acc[j + 3]);
    }
    float s = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) s = s + acc[j];
#pragma unroll
    for (uint32_t i = kMain; i < D; ++i) {
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
__device__ inline uint64_t lane_bcast64(uint64_t v, uint32_t src) {
    return ((uint64_t)lane_bcast((uint32_t)(v >> 32), src) << 32) | lane_bcast((uint32_t)v, src);
}
// Lane - 1 of v (lane 0 receives `fill`): DPP wave_shr:1, a register move with no LDS round trip.
__device__ inline uint32_t wave_shr1(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138, 0xF, 0xF, false);
}
__device__ inline float wave_shr1(float v, float fill) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138,
                                                      0xF, 0xF, false));
}

// v from lane ^ J (J a power of two below 64). J = 1, 2: DPP quad_perm; J = 4: DPP row_shl:4 /
// row_shr:4 selected by lane bit 2; J = 8: DPP row_ror:8; J = 16: ds_swizzle xor mode within 32
// lanes; J = 32: ds_bpermute.
template <uint32_t J>
__device__ inline uint32_t lane_xor(uint32_t v, uint32_t lane) {
    if constexpr (J == 1) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false);
    } else if constexpr (J == 2) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xF, 0xF, false);
    } else if constexpr (J == 4) {
        const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x104, 0xF, 0xF, false);
        const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x114, 0xF, 0xF, false);
        return (lane & 4) ? dn : up;
    } else if constexpr (J == 8) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x128, 0xF, 0xF, false);
    } else if constexpr (J == 16) {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);
    } else {
        static_assert(J == 32, "lane_xor: J must be a power of two below 64");
        return (uint32_t)__shfl_xor((int)v, 32);
    }
}
template <uint32_t J>
__device__ inline uint64_t lane_xor64(uint64_t v, uint32_t lane) {
    return ((uint64_t)lane_xor<J>((uint32_t)(v >> 32), lane) << 32) | lane_xor<J>((uint32_t)v, lane);
}
// One compare-exchange of a bitonic network across lanes at distance J: the pair's lower lane
// keeps the minimum when `asc`, the maximum otherwise.
template <uint32_t J>
__device__ inline uint64_t bitonic_step(uint64_t x, uint32_t lane, bool asc) {
    const uint64_t p = lane_xor64<J>(x, lane);
    const bool take_min = ((lane & J) == 0) == asc;
    return ((x < p) == take_min) ? x : p;
}
template <uint32_t SIZE>
__device__ inline uint64_t bitonic_phase_desc(uint64_t x, uint32_t lane) {
    // merge runs of SIZE / 2 into runs of SIZE, the run holding `lane` descending overall
    const bool asc = (lane & SIZE) != 0;
    if constexpr (SIZE >= 64) x = bitonic_step<32>(x, lane, asc);
    if constexpr (SIZE >= 32) x = bitonic_step<16>(x, lane, asc);
    if constexpr (SIZE >= 16) x = bitonic_step<8>(x, lane, asc);
    if constexpr (SIZE >= 8) x = bitonic_step<4>(x, lane, asc);
    if constexpr (SIZE >= 4) x = bitonic_step<2>(x, lane, asc);
    return bitonic_step<1>(x, lane, asc);
}
// 64 keys, one per lane, sorted descending (lane 0 the largest).
__device__ inline uint64_t sort64_desc(uint64_t x, uint32_t lane) {
    x = bitonic_phase_desc<2>(x, lane);
    x = bitonic_phase_desc<4>(x, lane);
    x = bitonic_phase_desc<8>(x, lane);
    x = bitonic_phase_desc<16>(x, lane);
    x = bitonic_phase_desc<32>(x, lane);
    return bitonic_phase_desc<64>(x, lane);
}

// A point's top-k list held by one wave: lane l owns slots l and l + 64 (k <= 128), sorted
// ascending by (dist, id); slots past the last filled one, and all slots >= k, hold the empty key.
// A slot is one 64-bit key, the distance's bits above the id: distances are non-negative, so
// unsigned key order is the (dist, id) order of pair_less and one 64-bit compare decides it. An
// insert is a few VALU, SALU, DPP and readlane operations with scalar branches only.
struct WaveList {
    static constexpr uint64_t kEmpty = ((uint64_t)0x7F7FFFFFu << 32) | 0xFFFFFFFFu;  // (FLT_MAX, kNoId)
    uint64_t k0, k1;
    uint64_t kth;  // key in slot k-1, wave-uniform
    float bound;   // its distance

    __device__ static inline uint64_t key(float d, uint32_t id) {
        return ((uint64_t)__float_as_uint(d) << 32) | id;
    }
    __device__ inline void set_bound(uint32_t k) {
        const uint32_t kl = k - 1;
        const uint64_t kk = kl >= 64 ? k1 : k0;
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(kk >> 32), (int)(kl & 63));
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)kk, (int)(kl & 63));
        kth = ((uint64_t)hi << 32) | lo;
        bound = __uint_as_float(hi);
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
    // Offer (nd, nid) to the list; wave-uniform arguments. The list keeps the k smallest keys
    // offered, with an id already present rejected (KnnSets.hpp:49, :59-72; a present id carries
    // the same distance, so its key equals the offered one). That set does not depend on the
    // order of offers; the reference's strict `dist < bound` (:152-157) differs only for a
    // distance equal to the k-th, where it keeps the first offered and this keeps the smaller id.
    __device__ inline bool insert(float nd, uint32_t nid, uint32_t k, uint32_t lane) {
        const uint64_t nk = key(nd, nid);
        if (!(nk < kth)) return false;
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

    // Offer up to 64 keys at once, one per lane (kEmpty where a lane has none); `scratch` is 128
    // keys of this wave's LDS. Same set as inserting them one by one: the k smallest distinct
    // keys. Up to `one_by_one` offers go through insert(); more go through a bitonic sort of the
    // batch, a half-cleaner against the list and a bitonic merge of the 128 smallest, then
    // adjacent equal keys (an id offered again) are dropped. Returns whether the list changed.
    __device__ inline bool offer64(uint64_t x, uint32_t k, uint32_t lane, uint64_t* scratch,
                                   uint32_t one_by_one = 4) {
        if (!(x < kth)) x = kEmpty;
        uint64_t m = __ballot(x != kEmpty);
        if (!m) return false;
        if ((uint32_t)__popcll(m) <= one_by_one) {
            bool changed = false;
            while (m) {
                const uint32_t src = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, (int)src);
                const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), (int)src);
                changed |= insert(__uint_as_float(hi), lo, k, lane);
            }
            return changed;
        }
        const uint64_t o0 = k0, o1 = k1;
        const uint64_t b = sort64_desc(x, lane);
        // list (ascending) ++ 64 empty ++ batch (descending) is bitonic. Its half-cleaner leaves
        // the 128 smallest, as a bitonic sequence, in (r0, r1); the next 64 are max(list slots
        // 64..127, batch) as a bitonic sequence, in r2. Duplicates are dropped only after the
        // merge, so the 192 sorted slots are kept: they hold every distinct key, and at most 64
        // of them are duplicates, which leaves at least 128 distinct ones.
        uint64_t r0 = k0, r1 = b < k1 ? b : k1, r2 = b < k1 ? k1 : b;
        const bool lt = r0 < r1;
        const uint64_t n0 = lt ? r0 : r1, n1 = lt ? r1 : r0;
        r0 = n0;
        r1 = n1;
        r0 = bitonic_step<32>(r0, lane, true); r1 = bitonic_step<32>(r1, lane, true); r2 = bitonic_step<32>(r2, lane, true);
        r0 = bitonic_step<16>(r0, lane, true); r1 = bitonic_step<16>(r1, lane, true); r2 = bitonic_step<16>(r2, lane, true);
        r0 = bitonic_step<8>(r0, lane, true);  r1 = bitonic_step<8>(r1, lane, true);  r2 = bitonic_step<8>(r2, lane, true);
        r0 = bitonic_step<4>(r0, lane, true);  r1 = bitonic_step<4>(r1, lane, true);  r2 = bitonic_step<4>(r2, lane, true);
        r0 = bitonic_step<2>(r0, lane, true);  r1 = bitonic_step<2>(r1, lane, true);  r2 = bitonic_step<2>(r2, lane, true);
        r0 = bitonic_step<1>(r0, lane, true);  r1 = bitonic_step<1>(r1, lane, true);  r2 = bitonic_step<1>(r2, lane, true);
        // drop a key equal to its predecessor (an id already in the list, or offered twice)
        auto prev = [&](uint64_t v, uint64_t before) {
            const uint32_t flo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)before, 63);
            const uint32_t fhi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(before >> 32), 63);
            return ((uint64_t)wave_shr1((uint32_t)(v >> 32), fhi) << 32) | wave_shr1((uint32_t)v, flo);
        };
        const uint64_t p0 = prev(r0, kEmpty), p1 = prev(r1, r0), p2 = prev(r2, r1);
        const bool dup0 = lane > 0 && r0 == p0 && r0 != kEmpty;
        const bool dup1 = r1 == p1 && r1 != kEmpty;
        const bool dup2 = r2 == p2 && r2 != kEmpty;
        const uint64_t dm0 = __ballot(dup0), dm1 = __ballot(dup1), dm2 = __ballot(dup2);
        if (dm0 | dm1) {
            const uint64_t below = (1ull << lane) - 1ull;
            const uint32_t d0 = __popcll(dm0), d1 = __popcll(dm1);
            const uint32_t q0 = lane - __popcll(dm0 & below);
            const uint32_t q1 = lane + 64 - d0 - __popcll(dm1 & below);
            const uint32_t q2 = lane + 128 - d0 - d1 - __popcll(dm2 & below);
            __builtin_amdgcn_wave_barrier();
            if (!dup0) scratch[q0] = r0;
            if (!dup1 && q1 < 128) scratch[q1] = r1;
            if (!dup2 && q2 < 128) scratch[q2] = r2;
            __builtin_amdgcn_wave_barrier();
            r0 = scratch[lane];
            r1 = scratch[lane + 64];
            __builtin_amdgcn_wave_barrier();
        }
        k0 = lane < k ? r0 : kEmpty;
        k1 = lane + 64 < k ? r1 : kEmpty;
        set_bound(k);
        return __ballot(k0 != o0 || k1 != o1) != 0;
    }
};

// ---- wave helpers and the 16-bit screening rows ----------------------------------------------
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
// Exclusive prefix sum over the 64 lanes, in registers: DPP row_shr:1/2/4/8 scans each row of 16
// lanes (a lane without a source adds 0), then row_bcast:15 and row_bcast:31 carry the row totals
// into the rows above (checked against a sequential scan on the device). Every lane must be active.
__device__ inline uint32_t wave_excl_scan(uint32_t v, uint32_t lane) {
    (void)lane;
    uint32_t x = v;
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x - v;
}

// ---- in-bucket limits ------------------------------------------------------------------------
// A limit is the m-th smallest of a row's in-bucket distance bounds. Each lane keeps the J smallest
// 16-bit keys it sees for a row (the top 16 bits of a non-negative float, which order as the
// floats do; 0x7F80 is +inf), and the limit is the m-th smallest key over a group of lanes, read
// back as the largest float with that key. The keys the lanes keep are a subset of all keys, so
// their m-th smallest is at least the m-th smallest of all: a valid upper bound.
constexpr uint32_t kKeyInf = 0x7F80u;
__device__ inline uint32_t key16(float v) { return __float_as_uint(fmaxf(v, 0.f)) >> 16; }
// key16(v) when ok (0 or 1), else kKeyInf, by arithmetic: no branch between the Gram's MFMA and
// the reads of its result
__device__ inline uint32_t key16_valid(float v, uint32_t ok) {
    return min(key16(v) | (0u - (ok ^ 1u)), kKeyInf);
}
__device__ inline float key16_upper(uint32_t x) {
    return x >= kKeyInf ? __builtin_inff() : __uint_as_float((x << 16) | 0xFFFFu);
}
template <uint32_t J>
__device__ inline void top_insert(uint32_t (&t)[J], uint32_t h) {
#pragma unroll
    for (uint32_t j = 0; j < J; ++j) {
        const uint32_t lo = min(t[j], h);
        h = max(t[j], h);
        t[j] = lo;
    }
}
// Keys kept per lane and row: the LDS leaf kernel (a row's keys over 16 lanes) and the big-bucket
// kernel (over 64 owners).
constexpr uint32_t kLimKeys = 16, kBigLimKeys = 4;
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
// Two rows' keys at once, one per 16-bit half (packed 16-bit min and max).
template <uint32_t J>
__device__ inline void top_insert2(u16x2 (&t)[J], u16x2 h) {
#pragma unroll
    for (uint32_t j = 0; j < J; ++j) {
        const u16x2 lo = __builtin_elementwise_min(t[j], h);
        h = __builtin_elementwise_max(t[j], h);
        t[j] = lo;
    }
}
// Sum over the 16 lanes of a DPP row (row_ror:8, 4, 2, 1): every lane of the row ends with it.
__device__ inline uint32_t row16_sum_u32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xF, 0xF, false);
    return v + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xF, 0xF, false);
}
// The m-th smallest key of a DPP row's 16 x J keys, for the two rows of packed key pairs (x0:
// the low halves, x1: the high halves): the smallest x with at least m keys at or below it, found
// one bit of x at a time from the top (x only grows while fewer than m keys lie below x + 2^bit;
// every key is at most kKeyInf, so x ends at most there). Both counts travel in one 32-bit sum
// (each at most 16 J < 2^16).
template <uint32_t J>
__device__ inline void row16_select2(const u16x2 (&t)[J], uint32_t m, uint32_t& x0, uint32_t& x1) {
    x0 = 0;
    x1 = 0;
#pragma unroll
    for (int bit = 14; bit >= 0; --bit) {
        const uint32_t p0 = x0 + (1u << bit) - 1u, p1 = x1 + (1u << bit) - 1u;
        uint32_t c = 0;
#pragma unroll
        for (uint32_t j = 0; j < J; ++j) c += ((uint32_t)t[j].x <= p0 ? 1u : 0u) + ((uint32_t)t[j].y <= p1 ? 0x10000u : 0u);
        c = row16_sum_u32(c);
        if ((c & 0xFFFFu) < m) x0 += 1u << bit;
        if ((c >> 16) < m) x1 += 1u << bit;
    }
}
// The same over the 64 lanes of a wave (J keys each); x is wave-uniform.
template <uint32_t J>
__device__ inline uint32_t wave_select(const uint32_t (&t)[J], uint32_t m) {
    uint32_t x = 0;
#pragma unroll
    for (int bit = 14; bit >= 0; --bit) {
        const uint32_t probe = x + (1u << bit) - 1u;
        uint32_t c = 0;
#pragma unroll
        for (uint32_t j = 0; j < J; ++j) c += (uint32_t)__popcll(__ballot(t[j] <= probe));
        if (c < m) x += 1u << bit;
    }
    return x;
}

// LeafArgs::half_kind of the FP32 Gram form: 16x16x4 FP32 MFMA over the FP32 rows (lane l gives
// row l & 15's dims 4 s + (l >> 4) at k-step s; the accumulator holds rows 4 (l >> 4) + i of
// column l & 15), d4 = ceil(d / 4) steps.
constexpr uint32_t kHalfF32 = 2;
__device__ inline float4_t gram_f32(const float* __restrict__ arow, const float* __restrict__ brow,
                                    uint32_t lk, uint32_t d4) {
    float4_t c = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t s = 0; s < d4; ++s)
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(arow[4 * s + lk], brow[4 * s + lk], c, 0, 0, 0);
    return c;
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
    if (lane < rs) hrow[lane] = lane < d ? h0 : 0;  // rs < 64 when d < 56
    if (lane + 64 < rs) hrow[lane + 64] = lane + 64 < d ? h1 : 0;
    rinv = ldexpf(1.0f, -sh);
}


typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// A lane's share of a 16-bit screening row (rows padded to 16 KS dims) for the 16x16 tile Gram on
// gfx950: k-steps of 32 dims on v_mfma_f32_16x16x32_{f16,bf16} (lane group g = lane >> 4 holds
// dims 32 m + 8 g .. + 7); when KS is odd, the last step takes its 16 dims in lane groups 0 and 1
// and zeros in groups 2 and 3 (a 16x16x16 MFMA for those dims gave lists that were wrong, and
// different from run to run, in the persistent leaf kernel once the in-bucket limit read the
// Gram). Rows and columns take the same dims per lane, so each product pairs the same dim of both.
template <uint32_t HALF, uint32_t KS>
struct HalfRow {
    static constexpr uint32_t K2 = KS / 2, K1 = KS % 2, KM = K2 + K1;
    uint4 w[KM];
    __device__ inline void load(const uint16_t* __restrict__ row, uint32_t g) {
#pragma unroll
        for (uint32_t m = 0; m < K2; ++m) w[m] = *reinterpret_cast<const uint4*>(row + 32 * m + 8 * g);
        if constexpr (K1 != 0)
            w[K2] = g < 2 ? *reinterpret_cast<const uint4*>(row + 32 * K2 + 8 * g) : make_uint4(0u, 0u, 0u, 0u);
    }
};

template <uint32_t HALF, uint32_t KS>
__device__ inline float4_t gram16(const HalfRow<HALF, KS>& a, const HalfRow<HALF, KS>& b) {
    float4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (uint32_t m = 0; m < HalfRow<HALF, KS>::KM; ++m) {
        if constexpr (HALF == 0)
            c = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8_t, a.w[m]),
                                                       __builtin_bit_cast(half8_t, b.w[m]), c, 0, 0, 0);
        else
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a.w[m]),
                                                        __builtin_bit_cast(bf16x8_t, b.w[m]), c, 0, 0, 0);
    }
    return c;
}

}  // namespace knng
