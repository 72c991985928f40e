// Leaf stage for gfx950: all pairs inside each leaf bucket (addCandidatesLessThan,
// KnnSets.hpp:131-160) merged into each point's top-k set (KnnSetScannableSimd, :26-127).
//
// One workgroup (8 or 16 waves) per bucket, rows resident in LDS:
//  phase 0  gather the bucket's rows, centre them on the bucket's first row, scale each row by a
//           power of two into fp16 (or bf16) range; FP32 squared norms of the centred rows.
//  phase 1  screening Gram on MFMA (16x16x32; fp32 accumulate), one 16-row strip per wave over
//           all columns: approximate distance n_i + n_j - 2 g_ij against the point's current
//           bound. Exact mode adds a rigorous error margin, so every pair whose exact FP32
//           distance beats the bound survives; survivors go to a per-row bitmask in LDS. Where
//           a row keeps many survivors (an empty list's first bucket), the strip is screened
//           again against the k-th smallest upper bound of the row's in-bucket distances.
//  phase 2  one wave per row: exact reference-order FP32 distance for each survivor, then a
//           wave-wide insert into the row's sorted (dist, id) list with id dedupe.
//
// The sorted list keeps the k smallest distinct ids offered (the reference's set semantics); ties
// at the k-th distance resolve by id instead of by arrival.
#include "knng_device.hpp"
#include "knng_kernels.hpp"

#include <cstdio>
#include <mutex>
#include <unordered_map>

namespace knng {

// Waves per workgroup: 8, or 16 for buckets above 128 points whose LDS (one workgroup per CU
// there) still fits, so that twice the waves hide the list and row loads of phase 2.
constexpr uint32_t kLeafWavesSmall = 8, kLeafWavesLarge = 16;



struct LeafLds {
    uint32_t b16, kp, rs, w16;
    uint32_t off_ids, off_bnd, off_nrm, off_rinv, off_mask, off_rows, off_rowbuf, off_cand, off_scr;
    uint32_t total;
    uint32_t part_rows;  // FP32 rows that fit the row region at once (phase 2 re-rank)
};

__host__ __device__ inline LeafLds leaf_lds_layout(uint32_t max_leaf, uint32_t d, uint32_t ld,
                                                  uint32_t waves = kLeafWavesSmall) {
    LeafLds L;
    L.b16 = (max_leaf + 15) & ~15u;
    L.kp = (d + 15) & ~15u;
    L.rs = leaf_row_stride(d);
    L.w16 = L.b16 / 16;
    uint32_t o = 0;
    L.off_ids = o;    o += L.b16 * 4;
    L.off_bnd = o;    o += L.b16 * 4;
    L.off_nrm = o;    o += L.b16 * 4;
    L.off_rinv = o;   o += L.b16 * 4;
    L.off_mask = o;   o += L.b16 * L.w16 * 2;
    o = (o + 15) & ~15u;
    L.off_rows = o;   o += L.b16 * L.rs * 2;
    o = (o + 15) & ~15u;
    L.part_rows = (L.b16 * L.rs * 2) / (ld * 4);
    L.off_rowbuf = o; o += waves * ld * 4;
    L.off_cand = o;   o += waves * L.b16 * 2;
    o = (o + 15) & ~15u;
    L.off_scr = o;    o += waves * 128 * 8;  // per-wave list scratch of the batched offer
    L.total = (o + 15) & ~15u;
    return L;
}

// KS = (d rounded up to 16) / 16 MFMA k-steps, a compile-time constant so the k-loop unrolls
// without guards and its LDS loads pipeline.
// One bucket, by the whole workgroup (W waves); a.max_leaf sets the LDS layout.
template <uint32_t HALF, uint32_t KS, uint32_t W>
__device__ inline void leaf_bucket(const LeafArgs& a, const LeafDesc lf) {
    constexpr uint32_t kLeafWaves = W, kLeafThreads = W * 64;
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t b = lf.size;
    if (b < 2) return;
    const uint32_t d = a.d, ld = a.ld, k = a.k;
    const LeafLds L = leaf_lds_layout(a.max_leaf, d, ld, W);
    const uint32_t b16 = (b + 15) & ~15u;
    const uint32_t nstrip = b16 / 16;
    uint32_t* s_ids = reinterpret_cast<uint32_t*>(lds + L.off_ids);
    float* s_bnd = reinterpret_cast<float*>(lds + L.off_bnd);
    float* s_nrm = reinterpret_cast<float*>(lds + L.off_nrm);
    float* s_rinv = reinterpret_cast<float*>(lds + L.off_rinv);
    uint16_t* s_mask = reinterpret_cast<uint16_t*>(lds + L.off_mask);
    uint16_t* s_rows = reinterpret_cast<uint16_t*>(lds + L.off_rows);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float* s_rowbuf = reinterpret_cast<float*>(lds + L.off_rowbuf) + wave * ld;
    uint16_t* s_cand = reinterpret_cast<uint16_t*>(lds + L.off_cand) + wave * L.b16;
    uint64_t* s_scr = reinterpret_cast<uint64_t*>(lds + L.off_scr) + wave * 128;

    __shared__ uint32_t s_stat[2];  // the workgroup's survivors and inserts
    if (tid < 2) s_stat[tid] = 0;

    // ---- phase 0: gather, centre, scale ---------------------------------------------------------
    for (uint32_t t = tid; t < b16; t += kLeafThreads) {
        const uint32_t id = t < b ? a.perm[lf.start + t] : kNoId;
        s_ids[t] = id;
        s_bnd[t] = t < b ? a.bound[id] : 0.f;
    }
    __syncthreads();
    const float* mu = a.X + (size_t)s_ids[0] * ld;
    const float mu0 = lane < d ? mu[lane] : 0.f;
    const float mu1 = lane + 64 < d ? mu[lane + 64] : 0.f;
    // four rows per wave at a time: their loads are in flight together
    for (uint32_t r0 = wave * 4; r0 < b16; r0 += kLeafWaves * 4) {
        float x0[4], x1[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t r = r0 + u;
            const float* x = a.X + (size_t)s_ids[r < b ? r : 0] * ld;
            x0[u] = lane < d ? x[lane] : 0.f;
            x1[u] = lane + 64 < d ? x[lane + 64] : 0.f;
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t r = r0 + u;
            if (r >= b16) break;
            uint16_t* hrow = s_rows + r * L.rs;
            if (r >= b) {
                for (uint32_t j = lane; j < L.rs; j += 64) hrow[j] = 0;
                if (lane == 0) {
                    s_nrm[r] = 0.f;
                    s_rinv[r] = 0.f;
                }
                continue;
            }
            float nrm, rinv;
            screen_vals<HALF>(lane < d ? x0[u] - mu0 : 0.f, lane + 64 < d ? x1[u] - mu1 : 0.f, d,
                              L.rs, hrow, lane, nrm, rinv);
            if (lane == 0) {
                s_nrm[r] = nrm;
                s_rinv[r] = rinv;
            }
        }
    }
    __syncthreads();

    // ---- phase 1: screening Gram --------------------------------------------------------------
    const bool exact = a.screen == kScreenExact;
    const float margin_c = HalfTraits<HALF>::kMargin;
    const uint32_t lrow = lane & 15, g4 = lane >> 4;
    // In-bucket limit (LeafArgs::rank_cap = m): a strip where some row keeps more than `trig`
    // survivors is screened again with each row's bound lowered to the m-th smallest of its
    // in-bucket distance bounds (the exact screen's upper bounds, or the approximate distances).
    // In the exact screen that only pays when many survivors are cut (an empty list's first
    // bucket), so it waits for m + 64; the re-rank cap of the approximate screens applies
    // whenever it binds.
    const uint32_t cap = a.rank_cap;
    const uint32_t trig = exact ? cap + 64 : cap;
    for (uint32_t s = wave; s < nstrip; s += kLeafWaves) {
        HalfRow<HALF, KS> af;
        af.load(s_rows + (16 * s + lrow) * L.rs, g4);
        float rn[4], rinv[4], lim[4];
        uint32_t rr[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            rr[i] = 16 * s + g4 * 4 + i;
            rn[i] = s_nrm[rr[i]];
            rinv[i] = s_rinv[rr[i]];
            lim[i] = s_bnd[rr[i]];
        }
        // one pass over the column tiles: the strip's survivor masks for the bounds lim; returns
        // the most survivors of a row of this lane group
        auto sweep = [&]() -> uint32_t {
            float thr_base[4];
            uint32_t cnt[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int i = 0; i < 4; ++i) thr_base[i] = exact ? lim[i] * (1.0f + 1.0f / 65536.0f) : lim[i];
            for (uint32_t t = 0; t < nstrip; ++t) {
                HalfRow<HALF, KS> bf;
                bf.load(s_rows + (16 * t + lrow) * L.rs, g4);
                const float4_t acc = gram16(af, bf);
                const uint32_t c = 16 * t + lrow;
                const float cn = s_nrm[c], cinv = s_rinv[c];
                uint64_t bits[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float g = acc[i] * (rinv[i] * cinv);
                    const float nsum = rn[i] + cn;
                    const float approx = nsum - 2.0f * g;
                    const float thr = exact ? thr_base[i] + margin_c * nsum : thr_base[i];
                    // <=: a zero distance (both rows at the centre) may meet a zero bound and win on its id
                    const bool ok = c < b && rr[i] < b && c != rr[i] && approx <= thr;
                    bits[i] = __ballot(ok);
                    cnt[i] += (uint32_t)__popcll((bits[i] >> (16 * g4)) & 0xFFFFull);
                }
                if (lane < 16) {
                    const uint32_t g = lane >> 2, i = lane & 3;
                    const uint64_t bw = i == 0 ? bits[0] : i == 1 ? bits[1] : i == 2 ? bits[2] : bits[3];
                    s_mask[(16 * s + lane) * L.w16 + t] = (uint16_t)(bw >> (16 * g));
                }
            }
            return max(max(cnt[0], cnt[1]), max(cnt[2], cnt[3]));
        };
        // each lane's 16 smallest keys of its rows' bounds (rows 2p and 2p + 1 in the halves of
        // pair p's registers), then the m-th smallest of the 256 keys of a row's lane group: with
        // m <= 160 a lane meets 10 of a row's m nearest on average, so few of them are missed
        auto limits = [&]() {
            constexpr uint32_t J = kLimKeys;
            u16x2 top[2][J];
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (uint32_t j = 0; j < J; ++j) top[p][j] = u16x2{(uint16_t)kKeyInf, (uint16_t)kKeyInf};
            for (uint32_t t = 0; t < nstrip; ++t) {
                HalfRow<HALF, KS> bf;
                bf.load(s_rows + (16 * t + lrow) * L.rs, g4);
                const float4_t acc = gram16(af, bf);
                const uint32_t c = 16 * t + lrow;
                const float cn = s_nrm[c], cinv = s_rinv[c];
                uint32_t key[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float g = acc[i] * (rinv[i] * cinv);
                    const float nsum = rn[i] + cn;
                    const float approx = nsum - 2.0f * g;
                    // the exact screen's upper bound: exact <= approx + c (n_i + n_j)
                    const float u = exact ? approx + margin_c * nsum : approx;
                    const uint32_t ok = (uint32_t)(c < b) & (uint32_t)(rr[i] < b) & (uint32_t)(c != rr[i]);
                    key[i] = key16_valid(u, ok);
                }
                top_insert2(top[0], u16x2{(uint16_t)key[0], (uint16_t)key[1]});
                top_insert2(top[1], u16x2{(uint16_t)key[2], (uint16_t)key[3]});
            }
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                uint32_t x0, x1;
                row16_select2(top[p], cap, x0, x1);
                lim[2 * p] = fminf(lim[2 * p], key16_upper(x0));
                lim[2 * p + 1] = fminf(lim[2 * p + 1], key16_upper(x1));
            }
        };
        const uint32_t most = sweep();
        if (cap != 0 && __ballot(most > trig) != 0) {
            limits();
            sweep();
        }
    }
    __syncthreads();

    // ---- phase 2: exact distances for survivors, sorted insert --------------------------------
    // The screen is done with the 16-bit rows, so their LDS region now holds FP32 rows, in column
    // parts of part_rows rows when the whole bucket does not fit; each part re-ranks the survivors
    // whose column lies in it.
    float* s_f32 = reinterpret_cast<float*>(lds + L.off_rows);
    __attribute__((address_space(3))) float* s_f32_lds = (__attribute__((address_space(3))) float*)(s_f32);
    const uint32_t ld4 = ld / 4;
    // t / ld4 = umulhi(t, ceil(2^32 / ld4)) exactly for t < 2^32 / ld4 (ld4 > 1)
    const uint32_t ld4_inv = ld4 > 1 ? (uint32_t)((0x100000000ull + ld4 - 1) / ld4) : 0u;
    uint32_t n_surv = 0, n_ins = 0;
    for (uint32_t c_lo = 0; c_lo < b; c_lo += L.part_rows) {
        const uint32_t c_hi = min(b, c_lo + L.part_rows);
        __syncthreads();
        // the part's rows go from global memory straight into LDS (global_load_lds_dwordx4: lane i
        // of a wave writes 16 bytes at the wave's base + 16 i), waited for before the barrier
        {
            const uint32_t tot4 = (c_hi - c_lo) * ld4;
            for (uint32_t t0 = tid - lane; t0 < tot4; t0 += kLeafThreads) {
                const uint32_t t = t0 + lane;
                if (t < tot4) {
                    const uint32_t rr = ld4 == 1 ? t : __umulhi(t, ld4_inv), q = t - rr * ld4;
                    __builtin_amdgcn_global_load_lds(
                        a.X + (size_t)s_ids[c_lo + rr] * ld + 4 * q,
                        (__attribute__((address_space(3))) void*)(s_f32_lds + 4 * t0), 16, 0, 0);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();
        // tiles overlapping [c_lo, c_hi): bits outside the part are masked off. The next row's
        // list and FP32 row are fetched while the current row is processed.
        WaveList nxt;
        float nx0 = 0.f, nx1 = 0.f;
        auto prefetch = [&](uint32_t rr) {
            const uint32_t id = s_ids[rr];
            nxt.load(a.knn_ids + (size_t)id * k, a.knn_d + (size_t)id * k, k, lane);
            const float* x = a.X + (size_t)id * ld;
            nx0 = lane < ld ? x[lane] : 0.f;
            nx1 = lane + 64 < ld ? x[lane + 64] : 0.f;
        };
        if (wave < b) prefetch(wave);
        for (uint32_t r = wave; r < b; r += kLeafWaves) {
            WaveList lst = nxt;
            const float x0 = nx0, x1 = nx1;
            if (r + kLeafWaves < b) prefetch(r + kLeafWaves);
            uint32_t w = 0;
            if (lane < nstrip) {
                w = s_mask[r * L.w16 + lane];
                const uint32_t t0 = 16 * lane;
                const uint32_t lo_b = c_lo > t0 ? min(16u, c_lo - t0) : 0u;
                const uint32_t hi_b = c_hi > t0 ? min(16u, c_hi - t0) : 0u;
                w &= ((1u << hi_b) - 1u) & ~((1u << lo_b) - 1u);
            }
            const uint32_t cnt = __popc(w);
            const uint32_t base = wave_excl_scan(cnt, lane);
            const uint32_t S = lane_bcast(base + cnt, 63);
            if (S == 0) continue;
            uint32_t o = base;
            while (w) {
                const uint32_t bit = __ffs(w) - 1;
                w &= w - 1;
                s_cand[o++] = (uint16_t)(16 * lane + bit);
            }
            if (lane < ld) s_rowbuf[lane] = x0;
            if (lane + 64 < ld) s_rowbuf[lane + 64] = x1;
            __builtin_amdgcn_wave_barrier();
            const uint32_t id_r = s_ids[r];
            bool changed = false;
            n_surv += S;
            for (uint32_t c0 = 0; c0 < S; c0 += 64) {
                uint64_t x = WaveList::kEmpty;
                if (c0 + lane < S) {
                    const uint32_t c = s_cand[c0 + lane];
                    const float* cr = s_f32 + (size_t)(c - c_lo) * ld;
                    // the reference's width (points[][112], 100 dims) fully unrolled
                    x = WaveList::key(d == 100 ? dist_ref_fixed<100>(s_rowbuf, cr) : dist_ref(s_rowbuf, cr, d),
                                      s_ids[c]);
                }
                n_ins += __popcll(__ballot((uint32_t)(x >> 32) < __float_as_uint(lst.bound)));
                changed |= lst.offer64(x, k, lane, s_scr);
            }
            if (changed) {
                lst.store(a.knn_ids + (size_t)id_r * k, a.knn_d + (size_t)id_r * k, k, lane);
                if (lane == 0) a.bound[id_r] = lst.bound;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (a.counters) {  // one pair of global atomics per workgroup: same-address atomics serialise
        if (lane == 0) {
            atomicAdd(&s_stat[0], n_surv);
            atomicAdd(&s_stat[1], n_ins);
        }
        __syncthreads();
        if (tid == 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.counters[0]), (unsigned long long)s_stat[0]);
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.counters[2]), (unsigned long long)s_stat[1]);
        }
    }
}

template <uint32_t HALF, uint32_t KS, uint32_t W>
__global__ __launch_bounds__(W * 64) void k_leaf(LeafArgs a) {
    leaf_bucket<HALF, KS, W>(a, a.leaves[blockIdx.x]);
}

// ---- buckets too large for LDS ------------------------------------------------------------------
// One workgroup (16 waves) per 64-row block of a bucket. Each wave owns 4 of the rows and keeps
// their sorted lists in registers for the whole kernel; the block's FP32 and 16-bit rows stay in
// LDS. The bucket's rows stream through LDS 64 at a time (16-bit screening rows, FP32 rows), the
// next column block's loads in flight while the current one is screened; for each column block the
// workgroup screens its 64 x 64 pairs on MFMA (one 16 x 16 tile per wave) (same scaling, centre and margin as k_leaf), then
// each wave re-ranks its rows' survivors exactly and offers them in one batch. Every row's list
// belongs to one workgroup, so there are no races; a pair (i, j) is computed once from each side,
// with the same distance bits both times.
constexpr uint32_t kBigThreads = 1024;
constexpr uint32_t kBigWaves = kBigThreads / 64;
constexpr uint32_t kBigRows = kLeafBigRows;
constexpr uint32_t kBigCols = 64;
constexpr uint32_t kBigRowsPerWave = kBigRows / kBigWaves;  // 4

struct BigLds {
    uint32_t rs, fs;  // 16-bit row stride (elements), FP32 row stride (floats)
    uint32_t off_arow, off_anrm, off_arinv, off_abnd, off_crow, off_cnrm, off_crinv, off_cid;
    uint32_t off_af32, off_cf32, off_mask, off_scr, off_pend, off_tlim, total;
};

__host__ __device__ inline BigLds big_lds_layout(uint32_t d, uint32_t ld) {
    BigLds L;
    L.rs = leaf_row_stride(d);
    // fs / 4 odd: the lanes of a phase-2 re-rank read consecutive rows without bank conflicts
    L.fs = ld % 8 == 0 ? ld + 4 : ld;
    uint32_t o = 0;
    L.off_arow = o;  o += kBigRows * L.rs * 2;
    o = (o + 15) & ~15u;
    L.off_anrm = o;  o += kBigRows * 4;
    L.off_arinv = o; o += kBigRows * 4;
    L.off_abnd = o;  o += kBigRows * 4;
    L.off_crow = o;  o += kBigCols * L.rs * 2;
    o = (o + 15) & ~15u;
    L.off_cnrm = o;  o += kBigCols * 4;
    L.off_crinv = o; o += kBigCols * 4;
    L.off_cid = o;   o += kBigCols * 4;
    o = (o + 15) & ~15u;
    L.off_af32 = o;  o += kBigRows * L.fs * 4;
    L.off_cf32 = o;  o += kBigCols * L.fs * 4;
    L.off_mask = o;  o += kBigRows * 8;
    L.off_scr = o;   o += kBigWaves * 128 * 8;
    L.off_pend = o;  o += kBigWaves * kBigRowsPerWave * 64 * 8;  // pending offers, 64 per owned row
    L.off_tlim = o;  o += kBigRows * 4;                          // the rows' in-bucket limits
    L.total = (o + 15) & ~15u;
    return L;
}

size_t leaf_big_lds_bytes(uint32_t d, uint32_t ld) { return big_lds_layout(d, ld).total; }

// One 64-row block of a bucket, by the whole workgroup (kBigThreads). CAP: the approximate
// screens' re-rank cap is compiled in (the exact forms' instances leave it out, and its registers).
template <uint32_t HALF, uint32_t KS, bool CAP>
__device__ inline void big_block(const LeafArgs& a, const BigBlock bb) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t b = bb.size, r0 = bb.row0;
    const uint32_t nrows = min(kBigRows, b - r0);
    const uint32_t d = a.d, ld = a.ld, k = a.k;
    const BigLds L = big_lds_layout(d, ld);
    uint16_t* s_arow = reinterpret_cast<uint16_t*>(lds + L.off_arow);
    float* s_anrm = reinterpret_cast<float*>(lds + L.off_anrm);
    float* s_arinv = reinterpret_cast<float*>(lds + L.off_arinv);
    float* s_abnd = reinterpret_cast<float*>(lds + L.off_abnd);
    uint16_t* s_crow = reinterpret_cast<uint16_t*>(lds + L.off_crow);
    float* s_cnrm = reinterpret_cast<float*>(lds + L.off_cnrm);
    float* s_crinv = reinterpret_cast<float*>(lds + L.off_crinv);
    uint32_t* s_cid = reinterpret_cast<uint32_t*>(lds + L.off_cid);
    float* s_cf32 = reinterpret_cast<float*>(lds + L.off_cf32);
    float* s_af32 = reinterpret_cast<float*>(lds + L.off_af32);
    uint16_t* s_mask = reinterpret_cast<uint16_t*>(lds + L.off_mask);  // 4 x 16 bits per row
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t* s_scr = reinterpret_cast<uint64_t*>(lds + L.off_scr) + wave * 128;
    const uint32_t* pm = a.perm + bb.start;
    const float* mu = a.X + (size_t)pm[0] * ld;  // centre: the bucket's first row, as in k_leaf
    const float mu0 = lane < d ? mu[lane] : 0.f;
    const float mu1 = lane + 64 < d ? mu[lane + 64] : 0.f;

    // ---- this wave's rows: lists in registers, 16-bit screening rows in LDS --------------------
    WaveList lst[kBigRowsPerWave];
    uint32_t pid[kBigRowsPerWave];
#pragma unroll
    for (uint32_t j = 0; j < kBigRowsPerWave; ++j) {
        const uint32_t rr = kBigRowsPerWave * wave + j;
        const bool valid = rr < nrows;
        pid[j] = valid ? __builtin_amdgcn_readfirstlane(pm[r0 + rr]) : kNoId;
        lst[j].k0 = lst[j].k1 = WaveList::kEmpty;
        lst[j].kth = 0;
        lst[j].bound = 0.f;
        if (valid) lst[j].load(a.knn_ids + (size_t)pid[j] * k, a.knn_d + (size_t)pid[j] * k, k, lane);
    }
#pragma unroll
    for (uint32_t j = 0; j < kBigRowsPerWave; ++j) {
        const uint32_t rr = kBigRowsPerWave * wave + j;
        uint16_t* hrow = s_arow + rr * L.rs;
        float* frow = s_af32 + rr * L.fs;
        if (rr < nrows) {
            const float* x = a.X + (size_t)pid[j] * ld;
            const float x0 = lane < ld ? x[lane] : 0.f, x1 = lane + 64 < ld ? x[lane + 64] : 0.f;
            if (lane < ld) frow[lane] = x0;
            if (lane + 64 < ld) frow[lane + 64] = x1;
            float nrm, rinv;
            if constexpr (HALF == kHalfF32) {  // the FP32 Gram reads the FP32 rows: their norms
                nrm = wave_sum(__builtin_fmaf(x0, x0, x1 * x1));
                rinv = 1.f;
            } else {
                screen_vals<HALF>(lane < d ? x0 - mu0 : 0.f, lane + 64 < d ? x1 - mu1 : 0.f, d,
                                  L.rs, hrow, lane, nrm, rinv);
            }
            if (lane == 0) {
                s_anrm[rr] = nrm;
                s_arinv[rr] = rinv;
                s_abnd[rr] = lst[j].bound;
            }
        } else {
            for (uint32_t t = lane; t < L.rs; t += 64) hrow[t] = 0;
            for (uint32_t t = lane; t < L.fs; t += 64) frow[t] = 0.f;
            if (lane == 0) {
                s_anrm[rr] = 0.f;
                s_arinv[rr] = 0.f;
                s_abnd[rr] = 0.f;
            }
        }
    }

    const bool exact = a.screen == kScreenExact;
    // FP32 Gram over K = 4 ceil(d/4) terms: |approx - exact| <= (K + 8) 2^-23 (n_i + n_j)
    const uint32_t d4 = (d + 3) / 4;
    const float margin_c = HALF == kHalfF32 ? (float)(4 * d4 + 8) * (1.0f / 8388608.0f)
                                            : HalfTraits<HALF == kHalfF32 ? 0 : HALF>::kMargin;
    const uint32_t lrow = lane & 15;
    const uint32_t strip = wave >> 2, tile = wave & 3;  // phase 1: 4 strips x 4 tiles
    bool changed[kBigRowsPerWave];
    // Offers below a row's bound wait in a pending batch of 64 (LDS) across column blocks and go
    // to the list by offer64 when the next block's would not fit, and at the end: the list keeps
    // the k smallest distinct keys offered, whatever the order and the bound they passed.
    uint64_t* s_pend = reinterpret_cast<uint64_t*>(lds + L.off_pend) + wave * kBigRowsPerWave * 64;
    uint32_t npend[kBigRowsPerWave];
#pragma unroll
    for (uint32_t j = 0; j < kBigRowsPerWave; ++j) {
        changed[j] = false;
        npend[j] = 0;
    }
    uint32_t n_surv = 0, n_ins = 0;

    // column staging: each wave fetches 4 rows of the next block into registers (fetch), then
    // writes them to LDS with their 16-bit screening rows (commit)
    constexpr uint32_t kColsPerWave = kBigCols / kBigWaves;  // 4
    float cx0[kColsPerWave], cx1[kColsPerWave];
    uint32_t cid[kColsPerWave];
    auto fetch = [&](uint32_t cb) {
#pragma unroll
        for (uint32_t u = 0; u < kColsPerWave; ++u) {
            const uint32_t cc = wave + kBigWaves * u;
            const bool valid = cb + cc < b;
            cid[u] = valid ? __builtin_amdgcn_readfirstlane(pm[cb + cc]) : kNoId;
            const float* x = a.X + (size_t)(valid ? cid[u] : pm[0]) * ld;
            cx0[u] = valid && lane < ld ? x[lane] : 0.f;
            cx1[u] = valid && lane + 64 < ld ? x[lane + 64] : 0.f;
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (uint32_t u = 0; u < kColsPerWave; ++u) {
            const uint32_t cc = wave + kBigWaves * u;
            uint16_t* hrow = s_crow + cc * L.rs;
            float* frow = s_cf32 + cc * L.fs;
            if (lane < ld) frow[lane] = cx0[u];
            if (lane + 64 < ld) frow[lane + 64] = cx1[u];
            if (cid[u] != kNoId) {
                float nrm, rinv;
                if constexpr (HALF == kHalfF32) {
                    nrm = wave_sum(__builtin_fmaf(cx0[u], cx0[u], cx1[u] * cx1[u]));
                    rinv = 1.f;
                } else {
                    screen_vals<HALF>(lane < d ? cx0[u] - mu0 : 0.f, lane + 64 < d ? cx1[u] - mu1 : 0.f, d,
                                      L.rs, hrow, lane, nrm, rinv);
                }
                if (lane == 0) {
                    s_cnrm[cc] = nrm;
                    s_crinv[cc] = rinv;
                }
            } else {
                for (uint32_t t = lane; t < L.rs; t += 64) hrow[t] = 0;
                if (lane == 0) {
                    s_cnrm[cc] = 0.f;
                    s_crinv[cc] = 0.f;
                }
            }
            if (lane == 0) s_cid[cc] = cid[u];
        }
    };

    // ---- in-bucket limits (LeafArgs::rank_cap = m) ---------------------------------------------
    // In the approximate screens (the re-rank cap) the block's rows get their bound lowered to the
    // m-th smallest of their approximate in-bucket distances before the screen. (The exact
    // screen's limit does not pay here: on the 10M x 128 Zipf round, whose rows all start with
    // empty lists, it cut the re-ranks but took the round from 477 to 564 ms, because each block
    // re-stages the whole bucket for it.) One extra pass over the column blocks screens the block's tiles as phase 1
    // does and keeps, per (tile, lane) owner of a row, its 4 smallest keys (64 owners: with m <= 160
    // an owner meets 2.5 of a row's m nearest on average);
    // the keys meet in LDS (the pending-offer area, unused until the main pass) and the wave that
    // owns a row selects its m-th smallest by ballot counts.
    float* s_tlim = reinterpret_cast<float*>(lds + L.off_tlim);
    if (lane < kBigRowsPerWave) s_tlim[kBigRowsPerWave * wave + lane] = __builtin_inff();
    __syncthreads();
    if constexpr (CAP && HALF != kHalfF32) {  // the FP32 Gram form is exact: no cap
        const uint32_t cap = a.rank_cap;
        auto limits = [&]() {
            constexpr uint32_t J = kBigLimKeys;
            uint32_t top[4][J];
    #pragma unroll
            for (int i = 0; i < 4; ++i)
    #pragma unroll
                for (uint32_t j = 0; j < J; ++j) top[i][j] = kKeyInf;
            HalfRow<HALF, KS> af;
            af.load(s_arow + (16 * strip + lrow) * L.rs, lane >> 4);
            uint32_t rr[4];
            float rn[4], rinv[4];
    #pragma unroll
            for (int i = 0; i < 4; ++i) {
                rr[i] = 16 * strip + (lane >> 4) * 4 + i;
                rn[i] = s_anrm[rr[i]];
                rinv[i] = s_arinv[rr[i]];
            }
            fetch(0);
            for (uint32_t c0 = 0; c0 < b; c0 += kBigCols) {
                const uint32_t nc = min(kBigCols, b - c0);
                __syncthreads();
                commit();
                __syncthreads();
                if (c0 + kBigCols < b) fetch(c0 + kBigCols);
                HalfRow<HALF, KS> bf;
                bf.load(s_crow + (16 * tile + lrow) * L.rs, lane >> 4);
                const float4_t acc = gram16(af, bf);
                const uint32_t c = 16 * tile + lrow;
                const float cn = s_cnrm[c], cinv = s_crinv[c];
    #pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float g = acc[i] * (rinv[i] * cinv);
                    const float nsum = rn[i] + cn;
                    const float approx = nsum - 2.0f * g;
                    const float u = exact ? approx + margin_c * nsum : approx;
                    const uint32_t ok = (uint32_t)(c < nc) & (uint32_t)(rr[i] < nrows) & (uint32_t)(r0 + rr[i] != c0 + c);
                    top_insert(top[i], key16_valid(u, ok));
                }
            }
            uint16_t* s_keys = reinterpret_cast<uint16_t*>(lds + L.off_pend);  // 64 rows x 64 owners x J
            static_assert(kBigRows * 64 * J * 2 <= kBigWaves * kBigRowsPerWave * 64 * 8, "keys fit the pending area");
            const uint32_t owner = 16 * tile + lrow;
    #pragma unroll
            for (int i = 0; i < 4; ++i)
    #pragma unroll
                for (uint32_t j = 0; j < J; ++j) s_keys[(rr[i] * 64 + owner) * J + j] = (uint16_t)top[i][j];
            __syncthreads();
    #pragma unroll 1
            for (uint32_t j2 = 0; j2 < kBigRowsPerWave; ++j2) {
                const uint32_t row = kBigRowsPerWave * wave + j2;
                uint32_t key[J];
    #pragma unroll
                for (uint32_t j = 0; j < J; ++j) key[j] = s_keys[(row * 64 + lane) * J + j];
                const float t = key16_upper(wave_select(key, cap));
                if (lane == 0 && row < nrows) {
                    s_tlim[row] = t;
                    s_abnd[row] = fminf(s_abnd[row], t);
                }
            }
            __syncthreads();  // the pending area is free again, the limits are in place
        };
        if (cap != 0 && !exact) limits();
    }
    fetch(0);

    for (uint32_t c0 = 0; c0 < b; c0 += kBigCols) {
        const uint32_t nc = min(kBigCols, b - c0);
        __syncthreads();  // the previous block's phase 2 is done with the column buffers
        commit();
        __syncthreads();
        if (c0 + kBigCols < b) fetch(c0 + kBigCols);

        // ---- phase 1: screen this wave's tile ------------------------------------------------
        {
            HalfRow<HALF == kHalfF32 ? 0 : HALF, KS> af;
            if constexpr (HALF != kHalfF32) af.load(s_arow + (16 * strip + lrow) * L.rs, lane >> 4);
            float rn[4], rinv[4], thr_base[4];
            uint32_t rr[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                rr[i] = 16 * strip + (lane >> 4) * 4 + i;
                rn[i] = s_anrm[rr[i]];
                rinv[i] = s_arinv[rr[i]];
                const float bnd = s_abnd[rr[i]];
                thr_base[i] = exact ? bnd * (1.0f + 1.0f / 65536.0f) : bnd;
            }
            {
                const uint32_t t = tile;
                float4_t acc;
                if constexpr (HALF == kHalfF32) {
                    acc = gram_f32(s_af32 + (16 * strip + lrow) * L.fs, s_cf32 + (16 * t + lrow) * L.fs,
                                   lane >> 4, d4);
                } else {
                    HalfRow<HALF, KS> bf;
                    bf.load(s_crow + (16 * t + lrow) * L.rs, lane >> 4);
                    acc = gram16(af, bf);
                }
                const uint32_t c = 16 * t + lrow;
                const float cn = s_cnrm[c], cinv = s_crinv[c];
                uint64_t bits[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float g = acc[i] * (rinv[i] * cinv);
                    const float nsum = rn[i] + cn;
                    const float approx = nsum - 2.0f * g;
                    const float thr = exact ? thr_base[i] + margin_c * nsum : thr_base[i];
                    const bool ok = c < nc && rr[i] < nrows && r0 + rr[i] != c0 + c && approx <= thr;
                    bits[i] = __ballot(ok);
                }
                if (lane < 16) {
                    const uint32_t g = lane >> 2, i = lane & 3;
                    const uint64_t bw = i == 0 ? bits[0] : i == 1 ? bits[1] : i == 2 ? bits[2] : bits[3];
                    s_mask[(16 * strip + lane) * 4 + t] = (uint16_t)(bw >> (16 * g));
                }
            }
        }
        __syncthreads();

        // ---- phase 2: exact re-rank of the wave's survivors, batched offers per row --------------
        // The block's survivors of the wave's rows are compacted into one list ((row << 6) | column,
        // in the wave's offer scratch), so each pass of 64 lanes re-ranks 64 of them whatever rows
        // they belong to; each row's results below its bound join the row's pending batch.
        {
            const bool last = c0 + kBigCols >= b;  // the last column block: the batches go out
            // a row's list selected in and out around an offer of its pending batch. With this one
            // offer site the compiler keeps the four lists in scratch memory between offers; a form
            // that keeps them in registers spills 53 VGPRs and ran the 10M x 128 Zipf leaf stage in
            // 520 ms against 437 ms
            auto offer_pending = [&](uint32_t j) {
                WaveList cur = lst[0];
                uint32_t np = npend[0];
#pragma unroll
                for (uint32_t jj = 1; jj < kBigRowsPerWave; ++jj)
                    if (j == jj) {
                        cur = lst[jj];
                        np = npend[jj];
                    }
                const uint64_t pd = lane < np ? s_pend[j * 64 + lane] : WaveList::kEmpty;
                const bool ch = cur.offer64(pd, k, lane, s_scr);
#pragma unroll
                for (uint32_t jj = 0; jj < kBigRowsPerWave; ++jj)
                    if (j == jj) {
                        lst[jj] = cur;
                        changed[jj] |= ch;
                        npend[jj] = 0;
                    }
            };
            uint64_t wm[kBigRowsPerWave];
            uint32_t tot = 0;
#pragma unroll
            for (uint32_t j = 0; j < kBigRowsPerWave; ++j) {
                const uint32_t rr = kBigRowsPerWave * wave + j;
                wm[j] = rr < nrows ? *reinterpret_cast<const uint64_t*>(s_mask + rr * 4) : 0ull;
                tot += __popcll(wm[j]);
            }
            // the block adds at most popc(wm[j]) offers to row j's batch: offer the batch first if
            // they might not fit
#pragma unroll 1
            for (uint32_t j = 0; j < kBigRowsPerWave; ++j) {
                uint64_t w = wm[0];
                uint32_t np = npend[0];
#pragma unroll
                for (uint32_t jj = 1; jj < kBigRowsPerWave; ++jj)
                    if (j == jj) {
                        w = wm[jj];
                        np = npend[jj];
                    }
                if (np + __popcll(w) > 64) offer_pending(j);
            }
            uint16_t* s_cand = reinterpret_cast<uint16_t*>(s_scr);  // <= 256 entries in 1 KiB
            const uint64_t below_l = (1ull << lane) - 1ull;
            __builtin_amdgcn_wave_barrier();
            {
                uint32_t base = 0;
#pragma unroll
                for (uint32_t j = 0; j < kBigRowsPerWave; ++j) {
                    if ((wm[j] >> lane) & 1ull) s_cand[base + __popcll(wm[j] & below_l)] = (uint16_t)((j << 6) | lane);
                    base += __popcll(wm[j]);
                }
            }
            __builtin_amdgcn_wave_barrier();
            for (uint32_t e0 = 0; e0 < tot; e0 += 64) {
                const uint32_t e = e0 + lane;
                uint64_t x = WaveList::kEmpty;
                uint32_t j = kBigRowsPerWave;
                if (e < tot) {
                    const uint32_t ent = s_cand[e], c = ent & 63u;
                    j = ent >> 6;
                    const float* ra = s_af32 + (size_t)(kBigRowsPerWave * wave + j) * L.fs;
                    const float* rc = s_cf32 + (size_t)c * L.fs;
                    const float dd = d == 16 * KS ? dist_ref_fixed<16 * KS>(ra, rc)
                                   : d == 100 ? dist_ref_fixed<100>(ra, rc) : dist_ref(ra, rc, d);
                    x = WaveList::key(dd, s_cid[c]);
                }
#pragma unroll
                for (uint32_t jj = 0; jj < kBigRowsPerWave; ++jj) {
                    const bool below = j == jj && x < lst[jj].kth;
                    const uint64_t m = __ballot(below);
                    n_ins += __popcll(m);
                    if (below) s_pend[jj * 64 + npend[jj] + __popcll(m & below_l)] = x;
                    npend[jj] += __popcll(m);
                }
            }
            n_surv += tot;
            __builtin_amdgcn_wave_barrier();
            if (last)
#pragma unroll 1
                for (uint32_t j = 0; j < kBigRowsPerWave; ++j) {
                    uint32_t np = npend[0];
#pragma unroll
                    for (uint32_t jj = 1; jj < kBigRowsPerWave; ++jj)
                        if (j == jj) np = npend[jj];
                    if (np) offer_pending(j);
                }
            if (lane == 0) {
#pragma unroll
                for (uint32_t j = 0; j < kBigRowsPerWave; ++j)
                    if (kBigRowsPerWave * wave + j < nrows)
                        s_abnd[kBigRowsPerWave * wave + j] = fminf(lst[j].bound, s_tlim[kBigRowsPerWave * wave + j]);
            }
        }
    }

#pragma unroll
    for (uint32_t j = 0; j < kBigRowsPerWave; ++j) {
        if (changed[j]) {
            lst[j].store(a.knn_ids + (size_t)pid[j] * k, a.knn_d + (size_t)pid[j] * k, k, lane);
            if (lane == 0) a.bound[pid[j]] = lst[j].bound;
        }
    }
    if (a.counters && lane == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.counters[0]), (unsigned long long)n_surv);
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.counters[2]), (unsigned long long)n_ins);
    }
}

template <uint32_t HALF, uint32_t KS, bool CAP>
__global__ __launch_bounds__(kBigThreads) void k_leaf_big(LeafArgs a, const BigBlock* __restrict__ blocks) {
    big_block<HALF, KS, CAP>(a, blocks[blockIdx.x]);
}

struct LeafQueue;  // knng_kernels.hpp

// ---- batched form: one persistent kernel over size-binned work queues -----------------------
// The round's work is one list: the 64-row blocks of the buckets above the LDS-resident limit
// first, then the LDS-resident buckets by decreasing size. A fixed grid of workgroups (the chip's
// resident slots, 16 waves each) takes items in list order from one device counter, so the
// longest items start first and no launch tail waits for a short one. LDS is sized once, for the
// limit bucket or a big-bucket block, whichever needs more, so the launch is the same every round
// and can be replayed from a graph.
__global__ void k_leaf_queue_reset(LeafQueue* q) {
    if (threadIdx.x == 0) q->next = 0;
}

// The two item kinds as separate functions: each gets its own register allocation, as in its
// stand-alone kernel, instead of one allocation for both.
template <uint32_t HALF, uint32_t KS, bool CAP>
__device__ __attribute__((noinline)) void big_block_item(const LeafArgs& a, const BigBlock bb) {
    big_block<HALF, KS, CAP>(a, bb);
}
template <uint32_t HALF, uint32_t KS>
__device__ __attribute__((noinline)) void leaf_bucket_item(const LeafArgs& a, const LeafDesc lf) {
    leaf_bucket<HALF, KS, kLeafWavesLarge>(a, lf);
}

template <uint32_t HALF, uint32_t KS, bool CAP>
__global__ __launch_bounds__(kBigThreads) void k_leaf_batched(LeafArgs a, const BigBlock* __restrict__ blocks,
                                                              LeafQueue* __restrict__ q) {
    static_assert(kBigThreads == kLeafWavesLarge * 64, "both item kinds take 16 waves");
    __shared__ uint32_t s_item;
    const uint32_t nbig = q->nbig, nleaf = q->nleaf;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&q->next, 1u);
        __syncthreads();
        const uint32_t it = s_item;
        if (it < nbig) big_block_item<HALF, KS, CAP>(a, blocks[it]);
        else if (it - nbig < nleaf) leaf_bucket_item<HALF, KS>(a, a.leaves[it - nbig]);
        __syncthreads();  // the item's LDS use and s_item read are over before the next item
        if (it >= nbig + nleaf) break;
    }
}

template <uint32_t HALF, uint32_t KS, bool CAP>
static void launch_leaf_big_t(const LeafArgs& a, const BigBlock* blocks, uint32_t nblocks, size_t lds,
                              hipStream_t st) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_leaf_big<HALF, KS, CAP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_leaf_big<HALF, KS, CAP><<<nblocks, kBigThreads, lds, st>>>(a, blocks);
}

template <uint32_t HALF, bool CAP>
static void launch_leaf_big_h(const LeafArgs& a, const BigBlock* blocks, uint32_t nblocks, size_t lds,
                              hipStream_t st) {
    switch ((a.d + 15) / 16) {
        case 1: launch_leaf_big_t<HALF, 1, CAP>(a, blocks, nblocks, lds, st); break;
        case 2: launch_leaf_big_t<HALF, 2, CAP>(a, blocks, nblocks, lds, st); break;
        case 3: launch_leaf_big_t<HALF, 3, CAP>(a, blocks, nblocks, lds, st); break;
        case 4: launch_leaf_big_t<HALF, 4, CAP>(a, blocks, nblocks, lds, st); break;
        case 5: launch_leaf_big_t<HALF, 5, CAP>(a, blocks, nblocks, lds, st); break;
        case 6: launch_leaf_big_t<HALF, 6, CAP>(a, blocks, nblocks, lds, st); break;
        case 7: launch_leaf_big_t<HALF, 7, CAP>(a, blocks, nblocks, lds, st); break;
        default: launch_leaf_big_t<HALF, 8, CAP>(a, blocks, nblocks, lds, st); break;
    }
}

// The approximate screens' re-rank cap is compiled into its own instances of the big-bucket kernel.
static bool leaf_cap(const LeafArgs& a) { return a.rank_cap != 0 && a.screen == kScreenApprox; }

void launch_leaf_big(const LeafArgs& a, const BigBlock* blocks, uint32_t nblocks, hipStream_t st) {
    if (!nblocks) return;
    const size_t lds = leaf_big_lds_bytes(a.d, a.ld);
    const bool cap = leaf_cap(a);
    if (a.half_kind == 0 && cap) launch_leaf_big_h<0, true>(a, blocks, nblocks, lds, st);
    else if (a.half_kind == 0) launch_leaf_big_h<0, false>(a, blocks, nblocks, lds, st);
    else if (a.half_kind == 1 && cap) launch_leaf_big_h<1, true>(a, blocks, nblocks, lds, st);
    else if (a.half_kind == 1) launch_leaf_big_h<1, false>(a, blocks, nblocks, lds, st);
    else launch_leaf_big_h<kHalfF32, false>(a, blocks, nblocks, lds, st);
}

size_t leaf_lds_bytes(uint32_t max_leaf, uint32_t d, uint32_t ld) {
    return leaf_lds_layout(max_leaf, d, ld).total;
}

uint32_t leaf_batched_limit(uint32_t d, uint32_t ld) {
    uint32_t lim = 0;
    for (uint32_t b = 16; b <= kMaxLeaf; b += 16)
        if (leaf_lds_layout(b, d, ld, kLeafWavesLarge).total <= 160 * 1024) lim = b;
    return lim;
}

size_t leaf_batched_lds(uint32_t d, uint32_t ld) {
    const uint32_t lim = leaf_batched_limit(d, ld);
    const size_t a = lim ? leaf_lds_layout(lim, d, ld, kLeafWavesLarge).total : 0;
    return std::max(a, leaf_big_lds_bytes(d, ld));
}

template <uint32_t HALF, uint32_t KS, bool CAP>
static void launch_leaf_batched_t(const LeafArgs& a, const BigBlock* blocks, LeafQueue* q, hipStream_t st) {
    const size_t lds = leaf_batched_lds(a.d, a.ld);
    auto kern = &k_leaf_batched<HALF, KS, CAP>;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBigThreads, lds);
    const uint32_t grid = (uint32_t)std::max(1, per_cu * cus);
    k_leaf_queue_reset<<<1, 64, 0, st>>>(q);
    kern<<<grid, kBigThreads, lds, st>>>(a, blocks, q);
}

template <uint32_t HALF, bool CAP>
static void launch_leaf_batched_h(const LeafArgs& a, const BigBlock* blocks, LeafQueue* q, hipStream_t st) {
    switch ((a.d + 15) / 16) {
        case 1: launch_leaf_batched_t<HALF, 1, CAP>(a, blocks, q, st); break;
        case 2: launch_leaf_batched_t<HALF, 2, CAP>(a, blocks, q, st); break;
        case 3: launch_leaf_batched_t<HALF, 3, CAP>(a, blocks, q, st); break;
        case 4: launch_leaf_batched_t<HALF, 4, CAP>(a, blocks, q, st); break;
        case 5: launch_leaf_batched_t<HALF, 5, CAP>(a, blocks, q, st); break;
        case 6: launch_leaf_batched_t<HALF, 6, CAP>(a, blocks, q, st); break;
        case 7: launch_leaf_batched_t<HALF, 7, CAP>(a, blocks, q, st); break;
        default: launch_leaf_batched_t<HALF, 8, CAP>(a, blocks, q, st); break;
    }
}

void launch_leaf_batched(const LeafArgs& a, const BigBlock* blocks, LeafQueue* q, hipStream_t st) {
    const bool cap = leaf_cap(a);
    if (a.half_kind == 0 && cap) launch_leaf_batched_h<0, true>(a, blocks, q, st);
    else if (a.half_kind == 0) launch_leaf_batched_h<0, false>(a, blocks, q, st);
    else if (cap) launch_leaf_batched_h<1, true>(a, blocks, q, st);
    else launch_leaf_batched_h<1, false>(a, blocks, q, st);
}

template <uint32_t HALF, uint32_t KS, uint32_t W>
static void launch_leaf_w(const LeafArgs& a, size_t lds, hipStream_t st) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_leaf<HALF, KS, W>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_leaf<HALF, KS, W><<<a.nleaves, W * 64, lds, st>>>(a);
}

// Workgroup-shape tuning (LeafArgs::tune). Bins above 128 points run as 8 or 16 waves per
// workgroup, with the same lists either way. The first launch of a (device, precision, width, bin)
// key does the real pass in the 16-wave shape, then one timed pass in each shape on the same
// leaves. A repeated pass of the exact screen offers the same (id, distance) pairs again, which the
// lists' id check rejects, so both timed passes see the same lists and leave them unchanged. The
// faster shape is kept for the rest of the process, and the two times are reported on stderr.
namespace {
std::mutex g_tune_mu;
std::unordered_map<uint64_t, uint32_t> g_tune;  // key -> waves per workgroup
}  // namespace

template <uint32_t HALF, uint32_t KS>
static void launch_leaf_t(const LeafArgs& a, hipStream_t st) {
    const size_t lds8 = leaf_lds_bytes(a.max_leaf, a.d, a.ld);
    const size_t lds16 = leaf_lds_layout(a.max_leaf, a.d, a.ld, kLeafWavesLarge).total;
    if (!(a.max_leaf > 128 && lds16 <= 160 * 1024)) {
        launch_leaf_w<HALF, KS, kLeafWavesSmall>(a, lds8, st);
        return;
    }
    if (!a.tune || a.screen != kScreenExact) {
        launch_leaf_w<HALF, KS, kLeafWavesLarge>(a, lds16, st);
        return;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    const uint32_t bin = a.max_leaf > 256 ? 1u : 0u;
    const uint64_t key = ((uint64_t)(uint32_t)dev << 32) | (HALF << 16) | (KS << 8) | bin;
    uint32_t w = 0;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        const auto it = g_tune.find(key);
        if (it != g_tune.end()) w = it->second;
    }
    if (w == kLeafWavesSmall) {
        launch_leaf_w<HALF, KS, kLeafWavesSmall>(a, lds8, st);
        return;
    }
    launch_leaf_w<HALF, KS, kLeafWavesLarge>(a, lds16, st);
    if (w) return;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    bool ok = true;
    for (auto& e : ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    float t8 = 0.f, t16 = 0.f;
    if (ok) {
        LeafArgs b = a;
        b.counters = nullptr;  // the statistics count the real pass only
        ok = hipEventRecord(ev[0], st) == hipSuccess;
        launch_leaf_w<HALF, KS, kLeafWavesSmall>(b, lds8, st);
        ok = ok && hipEventRecord(ev[1], st) == hipSuccess;
        launch_leaf_w<HALF, KS, kLeafWavesLarge>(b, lds16, st);
        ok = ok && hipEventRecord(ev[2], st) == hipSuccess && hipEventSynchronize(ev[2]) == hipSuccess &&
             hipEventElapsedTime(&t8, ev[0], ev[1]) == hipSuccess &&
             hipEventElapsedTime(&t16, ev[1], ev[2]) == hipSuccess;
    }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (!ok) return;  // no decision: the next launch of this key times again
    w = t8 < t16 ? kLeafWavesSmall : kLeafWavesLarge;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune.emplace(key, w);
    }
    std::fprintf(stderr, "[knng] leaf shape (%s, %u k-steps, leaves <= %u): 8 waves %.2f ms, 16 waves %.2f ms; %u kept\n",
                 HALF ? "bf16" : "fp16", KS, bin ? 512u : 256u, t8, t16, w);
}

template <uint32_t HALF>
static void launch_leaf_h(const LeafArgs& a, hipStream_t st) {
    switch ((a.d + 15) / 16) {
        case 1: launch_leaf_t<HALF, 1>(a, st); break;
        case 2: launch_leaf_t<HALF, 2>(a, st); break;
        case 3: launch_leaf_t<HALF, 3>(a, st); break;
        case 4: launch_leaf_t<HALF, 4>(a, st); break;
        case 5: launch_leaf_t<HALF, 5>(a, st); break;
        case 6: launch_leaf_t<HALF, 6>(a, st); break;
        case 7: launch_leaf_t<HALF, 7>(a, st); break;
        default: launch_leaf_t<HALF, 8>(a, st); break;
    }
}

void launch_leaf(const LeafArgs& a, hipStream_t st) {
    if (!a.nleaves) return;
    if (a.half_kind == 0) launch_leaf_h<0>(a, st);
    else launch_leaf_h<1>(a, st);
}

// ---- top-k state helpers ----------------------------------------------------------------------
__global__ void k_init_knn(uint32_t* ids, float* dists, float* bound, uint32_t n, uint32_t k) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (size_t)n * k) {
        ids[i] = kNoId;
        dists[i] = kFltMax;
    }
    if (i < n) bound[i] = kFltMax;
}

void launch_init_knn(uint32_t* ids, float* dists, float* bound, uint32_t n, uint32_t k,
                     hipStream_t st) {
    const size_t tot = (size_t)n * k;
    if (tot) k_init_knn<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(ids, dists, bound, n, k);
}

// bound[p] = the k-th distance of p's sorted list; and the statistics counters set to zero.
__global__ void k_bound_from_lists(const float* dists, uint32_t n, uint32_t k, float* bound,
                                   uint64_t* counters, uint32_t ncounters) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bound[i] = dists[i * k + k - 1];
    if (i < ncounters) counters[i] = 0;
}

__global__ void k_zero_u64(uint64_t* p, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

void launch_bound_from_lists(const float* dists, uint32_t n, uint32_t k, float* bound,
                             uint64_t* counters, uint32_t ncounters, hipStream_t st) {
    const size_t tot = std::max<size_t>(n, ncounters);
    if (tot) k_bound_from_lists<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(dists, n, k, bound,
                                                                               counters, ncounters);
}

void launch_zero_u64(uint64_t* p, uint32_t count, hipStream_t st) {
    if (count) k_zero_u64<<<(count + 255) / 256, 256, 0, st>>>(p, count);
}

// finalize + padResult (KnnSets.hpp:113-126, :334-344): lists are already sorted by (dist, id);
// empty slots become id 1.
// Lists of width kl -> outputs of width k <= kl (each list's first k: the k nearest found).
__global__ void k_finalize(const uint32_t* ids, const float* dists, uint32_t n, uint32_t kl,
                           uint32_t k, uint32_t* out_ids, float* out_d) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * k) return;
    const size_t src = (i / k) * kl + i % k;
    const uint32_t id = ids[src];
    out_ids[i] = id == kNoId ? 1u : id;
    if (out_d) out_d[i] = dists[src];
}

void launch_finalize(const uint32_t* ids, const float* dists, uint32_t n, uint32_t kl, uint32_t k,
                     uint32_t* out_ids, float* out_d, hipStream_t st) {
    const size_t tot = (size_t)n * k;
    if (tot) k_finalize<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(ids, dists, n, kl, k, out_ids, out_d);
}

}  // namespace knng
