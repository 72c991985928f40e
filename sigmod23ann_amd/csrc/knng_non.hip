// Neighbours-of-neighbours pass (topUpSingle, SolutionKmeans.hpp:480-555) for gfx950.
//
// Snapshot: every point's k id slots, sorted by id (the reference sorts too, :495); unfilled slots
// read as id 0, as the reference's zero-initialised current_ids do (KnnSets.hpp:29). Point p is
// offered every id in the snapshot lists of its snapshot neighbours except itself.
//
// Ordered form (topUpSingleOrdered, :393-477): p expands only its neighbours at distance ranks
// >= rank_lo (the reference uses 50), each into its filled neighbours. The snapshot keeps all of a
// list (it is also the r side and the membership set) and marks the expanded entries (kExpBit);
// an unfilled slot reads as the point's own id, which every membership test rejects.
//
// Reverse form (the fast path). The offers (p <- r) with q in S[p] and r in S[q] are grouped by q:
// a workgroup screens all pairs (p in RS[q], r in S[q]) of a q with an FP32 MFMA Gram, drops
// pairs with r already in S[p] (a membership mask from a hash of S[q]'s ids) and pairs that
// cannot beat p's bound, re-ranks the rest in exact reference-order FP32 and appends the ones below
// the bound to p's candidate buffer. A second kernel merges each buffer into its list (the list
// insert rejects the duplicates that arrive through several q). The pre-pass bound admits a
// superset of what the reference's live bound admits, so the merged sets are the same. A point whose
// buffer overflows is redone by the direct kernel.
//
// Direct form: one wave per point walks its 2-hop candidates itself (used for overflow and as the
// small-input path).
//
// Incremental passes: a snapshot entry carries bit 31 when its id was not in the previous pass's
// snapshot of the same point (every entry of a first pass is new). An offer p <- r through q whose
// two hops (q in S[p], r in S[q]) are both old was already made in the previous pass and can only
// be rejected again (bounds only shrink, lists keep the k smallest keys), so a later pass screens
// new rows of RS[q] against all of S[q] and old rows against S[q]'s new entries only.
#include "knng_device.hpp"
#include <algorithm>

#include "knng_kernels.hpp"

namespace knng {

// ---- snapshot sorted by id -------------------------------------------------------------------
// Bitonic sort of 128 keys held two per lane (element lane in k0, element lane + 64 in k1),
// ascending. Exchange partners lane ^ j: DPP quad permutes for j = 1, 2, swizzles within 32 lanes
// for j < 32, a permute for 32 and an in-lane exchange for 64.
template <uint32_t J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    if constexpr (J == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    else if constexpr (J == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    else if constexpr (J < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (int)(0x1F | (J << 10)));
    else return (uint32_t)__shfl_xor((int)v, (int)J);
}

// merge step j of stage s: sequences of s elements, ascending where element & s is 0 (for
// s = 64: k0's half ascending, k1's descending; s = 128: all ascending)
template <uint32_t S, uint32_t J>
__device__ __forceinline__ void bitonic_sub(uint32_t& k0, uint32_t& k1, uint32_t lane) {
    const bool lower = (lane & J) == 0;
    const bool asc0 = S >= 64 || (lane & S) == 0;
    const bool asc1 = S >= 128 || (S < 64 && (lane & S) == 0);
    const uint32_t o0 = xor_lane<J>(k0), o1 = xor_lane<J>(k1);
    k0 = lower == asc0 ? min(k0, o0) : max(k0, o0);
    k1 = lower == asc1 ? min(k1, o1) : max(k1, o1);
    if constexpr (J > 1) bitonic_sub<S, J / 2>(k0, k1, lane);
}

template <uint32_t S>
__device__ __forceinline__ void bitonic_stage(uint32_t& k0, uint32_t& k1, uint32_t lane) {
    if constexpr (S == 128) {
        const uint32_t lo = min(k0, k1);
        k1 = max(k0, k1);
        k0 = lo;
        bitonic_sub<128, 32>(k0, k1, lane);
    } else {
        bitonic_sub<S, S / 2>(k0, k1, lane);
        bitonic_stage<S * 2>(k0, k1, lane);
    }
}

// ---- snapshot ---------------------------------------------------------------------------------
// rcnt (nullable): each entry that goes to a reverse list (expanded, and not a repeat of an id
// that sorts before it, i.e. the repeated id 0 of an unfilled list) is counted for RS[id], the
// same entries k_rs_fill places; rnew (nullable) counts the new ones. pkey/pval (nullable,
// without rcnt): the same entries as (id, p) pairs at the entry's snapshot slot, every other slot
// as (0xFFFFFFFF, p), for a sort by id that yields the reverse lists.
__global__ __launch_bounds__(256) void k_non_snapshot_sorted(const uint32_t* __restrict__ ids,
                                                             uint32_t* __restrict__ snap,
                                                             uint32_t n, uint32_t k,
                                                             const uint32_t* __restrict__ prev,
                                                             uint32_t rank_lo,
                                                             uint32_t* __restrict__ rcnt,
                                                             uint32_t* __restrict__ rnew,
                                                             uint32_t* __restrict__ pkey,
                                                             uint32_t* __restrict__ pval) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= n) return;
    const uint32_t* src = ids + (size_t)p * k;
    uint32_t v0 = lane < k ? src[lane] : 0xFFFFFFFFu;
    uint32_t v1 = lane + 64 < k ? src[lane + 64] : 0xFFFFFFFFu;
    const bool un0 = lane < k && v0 == kNoId, un1 = lane + 64 < k && v1 == kNoId;
    if (un0) v0 = rank_lo ? p : 0u;
    if (un1) v1 = rank_lo ? p : 0u;
    // slots are in (dist, id) order, so a slot is its rank
    const bool x0 = !rank_lo || (!un0 && lane >= rank_lo);
    const bool x1 = !rank_lo || (!un1 && lane + 64 >= rank_lo);
    // sort key id << 1 | expanded (ids < 2^30); slots past k sort last. Equal ids (the repeated
    // fill id of an unfilled list) carry equal flags, so their order among themselves is moot.
    uint32_t k0 = lane < k ? v0 << 1 | (uint32_t)x0 : 0xFFFFFFFFu;
    uint32_t k1 = lane + 64 < k ? v1 << 1 | (uint32_t)x1 : 0xFFFFFFFFu;
    bitonic_stage<2>(k0, k1, lane);
    v0 = k0 >> 1;
    v1 = k1 >> 1;
    const uint32_t xb0 = k0 & 1u ? kExpBit : 0u, xb1 = k1 & 1u ? kExpBit : 0u;
    // rep: the previous sorted slot holds the same id (slot lane's predecessor is lane - 1's,
    // slot 64's is slot 63)
    const uint32_t pr0 = (uint32_t)__shfl((int)v0, (int)((lane + 63) & 63));
    const uint32_t pr1 = (uint32_t)__shfl((int)v1, (int)((lane + 63) & 63));
    const bool rep0 = lane > 0 && pr0 == v0;
    const bool rep1 = (lane > 0 ? pr1 : pr0) == v1;
    // new: not in the previous snapshot of this point (all new without one)
    uint32_t f0 = kNewBit, f1 = kNewBit;
    if (prev) {
        const uint32_t* pp = prev + (size_t)p * k;
        const uint32_t q0 = lane < k ? pp[lane] & kIdMask : 0xFFFFFFFFu;
        const uint32_t q1 = lane + 64 < k ? pp[lane + 64] & kIdMask : 0xFFFFFFFFu;
        bool h0 = false, h1 = false;
        for (uint32_t j = 0; j < 64; ++j) {
            const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)q0, (int)j);
            const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)q1, (int)j);
            h0 |= (a == v0) | (b == v0);
            h1 |= (a == v1) | (b == v1);
        }
        f0 = h0 ? 0u : kNewBit;
        f1 = h1 ? 0u : kNewBit;
    }
    uint32_t* dst = snap + (size_t)p * k;
    if (lane < k) dst[lane] = v0 | f0 | xb0;
    if (lane + 64 < k) dst[lane + 64] = v1 | f1 | xb1;
    if (pkey) {
        if (lane < k) {
            pkey[(size_t)p * k + lane] = xb0 && !rep0 ? v0 : 0xFFFFFFFFu;
            pval[(size_t)p * k + lane] = p;
        }
        if (lane + 64 < k) {
            pkey[(size_t)p * k + lane + 64] = xb1 && !rep1 ? v1 : 0xFFFFFFFFu;
            pval[(size_t)p * k + lane + 64] = p;
        }
    }
    if (rcnt) {
        if (lane < k && xb0 && !rep0) {
            atomicAdd(&rcnt[v0], 1u);
            if (rnew && f0) atomicAdd(&rnew[v0], 1u);
        }
        if (lane + 64 < k && xb1 && !rep1) {
            atomicAdd(&rcnt[v1], 1u);
            if (rnew && f1) atomicAdd(&rnew[v1], 1u);
        }
    }
}

// ---- reverse lists: RS[q] = { p : q in S[p] } ----------------------------------------------------
// Incremental passes (full passes sort (id, p) pairs instead): RS[q] holds the p with a new entry q
// first (rnew[q] of them), then the rest.
// One wave per snapshot row (slots lane, lane + 64; k <= 128), four rows per workgroup. An entry
// is placed when it is expanded and its id differs from the previous slot's (a repeated id 0 of
// an unfilled list is placed once), the entries launch_non_snapshot counted.
__global__ __launch_bounds__(256) void k_rs_fill(const uint32_t* __restrict__ snap, uint32_t n, uint32_t k,
                                                 const uint64_t* __restrict__ roff, const uint32_t* __restrict__ rcnt,
                                                 uint32_t* __restrict__ rfill, uint32_t* __restrict__ rfill_old,
                                                 uint32_t* __restrict__ rs) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= n) return;
    const uint32_t* row = snap + (size_t)p * k;
    const uint32_t v0 = lane < k ? row[lane] : 0u, v1 = lane + 64 < k ? row[lane + 64] : 0u;
    const uint32_t q0 = v0 & kIdMask, q1 = v1 & kIdMask;
    // the previous slot's id: lane - 1 of the same half; slot 63 for slot 64
    const uint32_t up0 = __shfl_up(q0, 1), up1 = __shfl_up(q1, 1), last0 = __shfl(q0, 63);
    const bool ok0 = lane < k && (v0 & kExpBit) && !(lane > 0 && up0 == q0);
    const bool ok1 = lane + 64 < k && (v1 & kExpBit) && (lane == 0 ? last0 : up1) != q1;
    auto place = [&](uint32_t v, uint32_t q) {
        const uint32_t pos = (!rfill_old || (v & kNewBit)) ? atomicAdd(&rfill[q], 1u)
                                                           : rcnt[q] - 1u - atomicAdd(&rfill_old[q], 1u);
        rs[roff[q] + pos] = p;
    };
    if (ok0) place(v0, q0);
    if (ok1) place(v1, q1);
}

// ---- exclusive scan of n u32 counts into n+1 u64 offsets ----------------------------------------
constexpr uint32_t kScanPer = 16, kScanThreads = 256, kScanBlock = kScanPer * kScanThreads;

__global__ __launch_bounds__(kScanThreads) void k_scan_sums(const uint32_t* __restrict__ in,
                                                            uint32_t n, uint64_t* __restrict__ sums) {
    __shared__ uint64_t red[kScanThreads];
    const size_t b0 = (size_t)blockIdx.x * kScanBlock;
    uint64_t s = 0;
    for (uint32_t j = 0; j < kScanPer; ++j) {
        const size_t i = b0 + (size_t)j * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t h = kScanThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void k_scan_bases(uint64_t* __restrict__ sums, uint32_t nb) {
    __shared__ uint64_t part[1024];
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = min(nb, b + per);
    uint64_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += sums[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t h = 1; h < 1024; h <<= 1) {
        const uint64_t v = threadIdx.x >= h ? part[threadIdx.x - h] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = part[threadIdx.x] - s;
    for (uint32_t i = b; i < e; ++i) {
        const uint64_t v = sums[i];
        sums[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* __restrict__ in,
                                                             uint32_t n,
                                                             const uint64_t* __restrict__ bases,
                                                             uint64_t* __restrict__ out) {
    __shared__ uint64_t part[kScanThreads];
    const size_t b0 = (size_t)blockIdx.x * kScanBlock + (size_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
    uint64_t s = 0;
    for (uint32_t j = 0; j < kScanPer; ++j) {
        v[j] = b0 + j < n ? in[b0 + j] : 0;
        s += v[j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t h = 1; h < kScanThreads; h <<= 1) {
        const uint64_t x = threadIdx.x >= h ? part[threadIdx.x - h] : 0;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    uint64_t run = bases[blockIdx.x] + part[threadIdx.x] - s;
    for (uint32_t j = 0; j < kScanPer; ++j) {
        if (b0 + j < n) out[b0 + j] = run;
        run += v[j];
    }
}

__global__ void k_scan_total(const uint32_t* __restrict__ in, uint32_t n, uint64_t* __restrict__ out) {
    out[n] = out[n - 1] + in[n - 1];
}

// ---- reverse screening kernel --------------------------------------------------------------------
// Persistent workgroups, each taking its q one at a time. FP32 rows of S[q] and 32-row chunks of
// RS[q] sit in LDS, each loaded as its
// first ceil(d/4) float4 (the dims beyond, up to that bound, are zero in every buffer the builder
// binds); an FP32 MFMA Gram (16x16x4) over ceil(d/4) steps
// screens every pair with a rigorous margin; pairs with r already in S[p] are masked by
// a per-chunk membership bitmap (S(q)'s ids sit in a bucketed hash table: one lookup is one
// 16-byte LDS read and a few integer operations, with no probe loop for the lanes of a wave to
// diverge on); the survivors queue up per wave and are re-ranked exactly: a short queue eight
// lanes per survivor (lane j owns the reference's accumulator j, then the lanes are summed in
// order), a longer one a lane per survivor. The next chunk's rows and snapshots are fetched into registers while the current chunk
// is screened. Accepted offers collect in an LDS list and go to the candidate buffers once, after
// the last chunk: an atomic that returns a value waits for every load issued before it, so
// appending inside the chunk loop would stall on the next chunk's fetch.
constexpr uint32_t kNrThreads = 512;  // 8 waves, dealt (16-row strip, 16-column tile) jobs in turn
constexpr uint32_t kNrChunk = 32;     // reverse-neighbour rows per chunk
constexpr uint32_t kNrMaxK4 = 32;     // ld / 4 for ld <= 128 (larger ld takes the direct kernel)
constexpr uint32_t kNrIds = 512;      // reverse-neighbour ids staged at a time
constexpr uint32_t kNrQueue = 320;    // per-wave survivor queue; flushed before it could overflow
constexpr uint32_t kNrPfRows = kNrChunk * kNrMaxK4 / kNrThreads;  // float4 per thread: dims < ld <= 128
constexpr uint32_t kNrPfSnap = kNrChunk * 128 / kNrThreads;        // snapshot ids per thread
constexpr uint32_t kNrSurv = 512;     // accepted offers held in LDS per workgroup before appending
typedef __attribute__((address_space(3))) uint32_t kLdsU32;
// consecutive q per workgroup. Workgroups end often enough for the dispatcher to interleave the
// previous sub-pass's merge, which runs beside on another stream (a persistent grid holding every
// slot for the whole launch delayed the merges: 1554 ms of NoN per 10M build against 1338)
constexpr uint32_t kNrRun = 32;
static_assert(kNrIds == kNrThreads, "a q's first reverse ids are staged one per thread");
static_assert(kNrThreads / 16 == kNrChunk, "16 threads commit each chunk row (one wave holds a row)");
static_assert(kNrChunk == 32 && (kNrThreads / 64) % 2 == 0,
              "a chunk is one or two 16-row strips, and a wave keeps its strip from job to job");

__host__ __device__ inline uint32_t nr_row_stride(uint32_t ld) { return ld % 8 == 0 ? ld + 4 : ld; }

// S(q)'s ids (<= 128) hashed into 128 buckets of 4 slots. A key whose bucket is full is left out:
// membership then reads "absent", the offer goes through, and the list's id check rejects it.
constexpr uint32_t kNrBuckets = 128;
constexpr uint32_t kNrSlots = 4;
constexpr uint32_t kNrHash = kNrBuckets * kNrSlots;
// S(q) rows staged per thread (k <= 128 rows of <= 33 float4)
constexpr uint32_t kNrBRows = (128 * (kNrMaxK4 + 1) + kNrThreads - 1) / kNrThreads;

// LDS: the fixed-size arrays first, at compile-time offsets (they fold into the LDS instructions'
// immediate offsets and take no scalar registers), then the arrays sized by k and ld.
constexpr uint32_t kNrOffHkey = 0;                                   // hash entries (16-byte buckets)
constexpr uint32_t kNrOffHcnt = kNrOffHkey + kNrHash * 4;            // bucket fill counts
constexpr uint32_t kNrOffRsid = kNrOffHcnt + kNrBuckets * 4;         // reverse ids, current / next q
constexpr uint32_t kNrOffRsid2 = kNrOffRsid + kNrIds * 4;
constexpr uint32_t kNrOffNsq = kNrOffRsid2 + kNrIds * 4;             // next q's snapshot slots
constexpr uint32_t kNrOffNbn = kNrOffNsq + 128 * 4;                  // ... and their norms
constexpr uint32_t kNrOffMeta = kNrOffNbn + 128 * 4;                 // next q: q, roff[q], roff[q+1]
constexpr uint32_t kNrOffAid = kNrOffMeta + 32;                      // chunk rows: ids, bounds, norms
constexpr uint32_t kNrOffAbnd = kNrOffAid + kNrChunk * 4;
constexpr uint32_t kNrOffAnrm = kNrOffAbnd + kNrChunk * 4;
constexpr uint32_t kNrOffMask = kNrOffAnrm + kNrChunk * 4;           // membership words, 4 per row
constexpr uint32_t kNrOffQueue = kNrOffMask + kNrChunk * 4 * 4;      // per-wave survivor queues
constexpr uint32_t kNrOffSpid = (kNrOffQueue + (kNrThreads / 64) * kNrQueue * 2 + 15) & ~15u;
constexpr uint32_t kNrOffSrid = kNrOffSpid + kNrSurv * 4;            // accepted offers
constexpr uint32_t kNrOffSd = kNrOffSrid + kNrSurv * 4;
constexpr uint32_t kNrOffCnt = kNrOffSd + kNrSurv * 4;               // [0] accepted, [1] screened, [2] new
constexpr uint32_t kNrOffXq = (kNrOffCnt + 16 + 15) & ~15u;         // q's FP32 row (16-bit screens)
constexpr uint32_t kNrOffBscl = kNrOffXq + 128 * 4;                  // S(q)'s row scale exponents
constexpr uint32_t kNrOffAscl = kNrOffBscl + 128 * 4;                // chunk rows' scales
constexpr uint32_t kNrFixed = kNrOffAscl + kNrChunk * 4;
static_assert(kNrOffMeta % 8 == 0, "s_meta holds 64-bit words");

struct NrLds {
    uint32_t kb16, ldp, k32, rs16;
    uint32_t a16u;  // screen 3: 16-byte units per fp16 chunk row (odd: 16 rows' reads hit distinct banks)
    uint32_t off_bid, off_bnrm, off_brow, off_arow, off_brinv, off_b16, off_a16, off_arinv, off_a16f, total;
};

// FP32 Gram (screen16 = 0): S(q)'s FP32 rows sit in LDS for the Gram and the exact re-rank, k of
// them: the Gram's last column tile also reads the rows up to kb16, which lie in the chunk rows that
// follow (those columns are never offered). Screen 3 (fp16 Gram of the FP32 rows) adds the chunk's
// rows converted to fp16 once per chunk (the Gram's A operand).
// 16-bit screens: S(q)'s rows sit in LDS only as 16-bit rows centred on q (the exact re-rank reads
// their FP32 rows from global memory), beside the chunk's FP32 and 16-bit rows, so two workgroups
// still share a CU.
__host__ __device__ inline NrLds nr_layout(uint32_t k, uint32_t ld, uint32_t d, uint32_t screen16) {
    const bool f16a = screen16 == 3;
    if (screen16 == 3) screen16 = 0;  // fp16 Gram of the FP32 rows in LDS: the FP32 layout
    NrLds L;
    L.kb16 = (k + 15) & ~15u;
    L.ldp = nr_row_stride(ld);
    L.k32 = (d + 31) / 32;
    L.rs16 = 32 * L.k32 + 8;  // halves; +8 staggers the banks of consecutive rows
    uint32_t o = kNrFixed;
    L.off_bid = o;  o += L.kb16 * 4;
    L.off_bnrm = o; o += L.kb16 * 4;
    L.off_brinv = o; o += screen16 ? L.kb16 * 4 : 0;
    o = (o + 15) & ~15u;
    L.off_brow = o; o += screen16 ? 0 : k * L.ldp * 4;
    L.off_b16 = o;  o += screen16 ? L.kb16 * L.rs16 * 2 : 0;
    o = (o + 15) & ~15u;
    L.off_arow = o; o += kNrChunk * L.ldp * 4;
    L.off_a16 = o;  o += screen16 ? kNrChunk * L.rs16 * 2 : 0;
    L.off_arinv = o; o += screen16 ? kNrChunk * 4 : 0;
    const uint32_t d4 = (d + 3) / 4;
    L.a16u = ((d4 + 1) / 2) | 1u;
    o = (o + 15) & ~15u;
    L.off_a16f = o; o += f16a ? kNrChunk * L.a16u * 16 : 0;
    // 16-bit screens stage S(q)'s FP32 rows from off_b16 on, over the 16-bit and chunk regions
    if (screen16) o = o > L.off_b16 + L.kb16 * L.ldp * 4 ? o : L.off_b16 + L.kb16 * L.ldp * 4;
    L.total = (o + 15) & ~15u;
    return L;
}

size_t non_reverse_lds(uint32_t k, uint32_t ld, uint32_t d, uint32_t screen16) {
    return nr_layout(k, ld, d, screen16).total;
}
// For these widths the 16-bit rows (rs16 = 32 ceil(d / 32) + 8 halves) are at most as long in
// bytes as half an FP32 row (rs16 <= 2 ldp), which the in-place conversion of the staged rows needs.
bool non_screen16_supported(uint32_t d) {
    const uint32_t ks = (d + 3) / 4;
    return ks == 16 || ks == 25 || ks == 32;
}

// Membership hash of an id: h = id * odd constant is a bijection of 32-bit words, so the bucket
// (h's top 7 bits) and h's low 25 bits identify the id. An entry holds (h << 7) | column (column
// < 128); an empty slot holds 0xFFFFFFFF. For a query h, entry ^ (h << 7) is the column when the
// entry is the query's, at least 128 for another id's entry and at least 127 for an empty slot, so
// the minimum over the bucket's 4 slots, when below 127, is the column (a column of 127, k = 128
// only, reads as absent, which is always safe). Id 0 is never inserted, and its query lands in
// bucket 0 with h = 0, where no entry is below 128.
__device__ inline uint32_t nr_hash(uint32_t id) { return id * 0x9E3779B1u; }

// The 8 partial sums of a lane group (lanes 8g .. 8g+7) added in lane order into 0.0f, valid in
// the group's first lane. A group never crosses a 16-lane DPP row, so lane 8g reads lane 8g + j
// with row_shl:j: register moves, no LDS round trip.
__device__ inline float lane8_sum_in_order(float acc) {
    const int v = __float_as_int(acc);
    float s = 0.0f + acc;
    s = s + __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x101, 0xF, 0xF, false));
    s = s + __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x102, 0xF, 0xF, false));
    s = s + __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x103, 0xF, 0xF, false));
    s = s + __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x104, 0xF, 0xF, false));
    s = s + __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x105, 0xF, 0xF, false));
    s = s + __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x106, 0xF, 0xF, false));
    s = s + __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x107, 0xF, 0xF, false));
    return s;
}

// Reference-order distance of two rows, computed by the 8 lanes of a lane group: lane j of the
// group accumulates dims j, j+8, ... below floor(d/8)*8 with fused multiply-add; the 8 partial sums
// are added in lane order into 0.0f; the tail dims are added with fused multiply-add. The group's
// first lane (lane % 8 == 0) returns dist_ref(a, b, d); the other lanes return partial values.
__device__ inline float dist_ref_lane8(const float* __restrict__ a, const float* __restrict__ b,
                                       uint32_t d, uint32_t lane) {
    const uint32_t main = d & ~7u, l8 = lane & 7u;
    float acc = 0.f;
    for (uint32_t i = l8; i < main; i += 8) {
        const float t = a[i] - b[i];
        acc = __builtin_fmaf(t, t, acc);
    }
    float s = lane8_sum_in_order(acc);
    for (uint32_t i = main; i < d; ++i) {
        const float t = a[i] - b[i];
        s = __builtin_fmaf(t, t, s);
    }
    return s;
}

// dist_ref_lane8 for a compile-time width D: every load of the loops is issued at once.
template <uint32_t D>
__device__ inline float dist_ref_lane8_fixed(const float* __restrict__ a, const float* __restrict__ b,
                                             uint32_t lane) {
    constexpr uint32_t kMain = D & ~7u;
    const uint32_t l8 = lane & 7u;
    float acc = 0.f;
#pragma unroll
    for (uint32_t i = 0; i < kMain; i += 8) {
        const float t = a[i + l8] - b[i + l8];
        acc = __builtin_fmaf(t, t, acc);
    }
    float s = lane8_sum_in_order(acc);
#pragma unroll
    for (uint32_t i = kMain; i < D; ++i) {
        const float t = a[i] - b[i];
        s = __builtin_fmaf(t, t, s);
    }
    return s;
}

// FP32 Gram of a 16x16 tile over K = 4 KS dims, one accumulation chain (an MFMA that takes the
// previous one's result as its accumulator issues back to back). arow and brow are the lane's rows
// (row lane & 15) in LDS; k-step s of the MFMA sums the products of lane group lk = lane >> 4. Any
// assignment of the dims to (k-step, lane group) that takes each dim once gives the same Gram up to
// rounding, and the screen's error bound holds for every summation order. With KS a compile-time
// constant, group lk takes dims 16m + 4lk .. 16m + 4lk + 3 for m < KS / 4 (one 16-byte LDS read per
// operand for four k-steps) and dim 16 (KS / 4) + 4s + lk for the last KS % 4 k-steps s; KS = 0 is
// the generic runtime loop over dims 4s + lk.
template <uint32_t KS>
__device__ inline float4_t gram_tile_f32(const float* __restrict__ arow,
                                         const float* __restrict__ brow, uint32_t lk, uint32_t ksteps) {
    float4_t c0 = {0.f, 0.f, 0.f, 0.f};
    if constexpr (KS > 0) {
        constexpr uint32_t M = KS / 4, R = KS % 4;
        float4 av[M], bv[M];
        float ar[R > 0 ? R : 1], br[R > 0 ? R : 1];
#pragma unroll
        for (uint32_t m = 0; m < M; ++m) {
            av[m] = *reinterpret_cast<const float4*>(arow + 16 * m + 4 * lk);
            bv[m] = *reinterpret_cast<const float4*>(brow + 16 * m + 4 * lk);
        }
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            ar[r] = arow[16 * M + 4 * r + lk];
            br[r] = brow[16 * M + 4 * r + lk];
        }
#pragma unroll
        for (uint32_t m = 0; m < M; ++m) {
            c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m].x, bv[m].x, c0, 0, 0, 0);
            c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m].y, bv[m].y, c0, 0, 0, 0);
            c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m].z, bv[m].z, c0, 0, 0, 0);
            c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[m].w, bv[m].w, c0, 0, 0, 0);
        }
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[r], br[r], c0, 0, 0, 0);
    } else {
        for (uint32_t kk = 0; kk < ksteps; ++kk)
            c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(arow[4 * kk + lk], brow[4 * kk + lk], c0, 0, 0, 0);
    }
    return c0;
}

// ---- 16-bit screens centred on q -----------------------------------------------------------------
// A row's 16 threads (one DPP row of a wave) each hold 8 of its dims: 4 sc .. 4 sc + 3 and
// 64 + 4 sc .. 64 + 4 sc + 3. They centre their dims on q's row in FP32, reduce the row's largest
// |value| and its squared norm over the 16 lanes (DPP row rotations: all lanes end with the
// result), scale the row by a power of two so that its largest |value| is below 2^14 (as the leaf
// stage's rows), and store it as 16-bit values. Relative rounding per element therefore holds for
// every element of the row, which is what the screen's margin assumes.
constexpr uint32_t kNrSPass = 4;  // S(q) rows converted 32 at a time: k <= 128

template <int CTRL>
__device__ inline float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
// row_ror:8, 4, 2, 1 (DPP controls 0x128, 0x124, 0x122, 0x121)
__device__ inline float row16_max(float v) {
    v = fmaxf(v, dpp_f<0x128>(v));
    v = fmaxf(v, dpp_f<0x124>(v));
    v = fmaxf(v, dpp_f<0x122>(v));
    return fmaxf(v, dpp_f<0x121>(v));
}
__device__ inline float row16_sum(float v) {
    v = v + dpp_f<0x128>(v);
    v = v + dpp_f<0x124>(v);
    v = v + dpp_f<0x122>(v);
    return v + dpp_f<0x121>(v);
}

template <uint32_t SCR>
__device__ inline uint16_t to16(float f) {
    if constexpr (SCR == 1) return __builtin_bit_cast(uint16_t, (_Float16)f);
    else return to_bf16_rne(f);
}

// Centres (v0, v1) on (x0, x1), the thread's dims of a row and of q's row; `valid` false gives a zero
// row. Writes the 16-bit dims at row + 4 sc and row + 64 + 4 sc (below rs16 - 8 only), returns the
// FP32 squared norm of the centred row and the inverse of its scale. All 16 threads of the row
// must call it together.
template <uint32_t SCR>
__device__ inline void centre_row16(float4 v0, float4 v1, float4 x0, float4 x1, bool valid,
                                    uint32_t sc, uint32_t d, uint32_t rs16,
                                    uint16_t* __restrict__ row, float& nrm, float& rinv) {
    float c[8] = {v0.x - x0.x, v0.y - x0.y, v0.z - x0.z, v0.w - x0.w,
                  v1.x - x1.x, v1.y - x1.y, v1.z - x1.z, v1.w - x1.w};
    float m = 0.f, s2 = 0.f;
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) {
        const uint32_t dim = (e < 4 ? 4 * sc : 64 + 4 * sc) + (e & 3);
        if (!valid || dim >= d) c[e] = 0.f;
        m = fmaxf(m, fabsf(c[e]));
        s2 = __builtin_fmaf(c[e], c[e], s2);
    }
    m = row16_max(m);
    nrm = row16_sum(s2);
    int ex = 0;
    if (m > 0.f) frexpf(m, &ex);
    const int sh = 14 - ex;
    const float scl = ldexpf(1.0f, sh);
    rinv = ldexpf(1.0f, -sh);
    uint32_t w[4];
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e)
        w[e] = (uint32_t)to16<SCR>(c[2 * e] * scl) | ((uint32_t)to16<SCR>(c[2 * e + 1] * scl) << 16);
    const uint32_t lim = rs16 - 8;
    if (4 * sc < lim) *reinterpret_cast<uint2*>(row + 4 * sc) = make_uint2(w[0], w[1]);
    if (64 + 4 * sc < lim) *reinterpret_cast<uint2*>(row + 64 + 4 * sc) = make_uint2(w[2], w[3]);
}

// A row's fp16 scale exponent from its squared norm n: 2^(14 - e) with sqrt(n) < 2^e, so every
// |value| of the row, at most sqrt(n), scales below 2^14, and the largest, at least sqrt(n / d),
// stays far above fp16's subnormal range.
__device__ inline float scale_exp(float n) {
    int e = 0;
    if (n > 0.f) frexpf(sqrtf(n), &e);
    return (float)(14 - e);
}

// fp16 Gram of a 16x16 tile straight from FP32 rows in LDS: lane group g = lane >> 4 takes dims
// 32 m + 8 g .. + 7 of k-step m for the row (arow) and the column (brow) alike, each row scaled by
// its power of two (sa, sb: exact) into fp16 range, or taken as it is (UNSC: every row's values
// already lie where the margin holds unscaled); dims at or past 4 KS read as zero. Both operands
// are converted per tile: holding the strip's converted rows across tiles costs registers the
// kernel does not have (measured: spills, 313 ms per launch against 217).
template <uint32_t KS, uint32_t UNSC>
__device__ inline float4_t gram_tile_f16s(const float* __restrict__ arow, const float* __restrict__ brow,
                                          float sa, float sb, uint32_t g) {
    if constexpr (UNSC != 0) {
        sa = 1.0f;  // a multiply by a constant 1 folds away
        sb = 1.0f;
    }
    constexpr uint32_t K32 = (4 * KS + 31) / 32;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    float4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (uint32_t m = 0; m < K32; ++m) {
        const uint32_t q0 = 8 * m + 2 * g;  // float4 index of the lane's first dims
        // past 4 KS dims: load an in-row float4 and select zero (a select per value, where a
        // guarded load per value made a branch and an LDS wait of each)
        const bool v0 = q0 < KS, v1 = q0 + 1 < KS;
        const uint32_t i0 = v0 ? q0 : 0u, i1 = v1 ? q0 + 1 : 0u;
        float4 a0 = reinterpret_cast<const float4*>(arow)[i0], a1 = reinterpret_cast<const float4*>(arow)[i1];
        float4 b0 = reinterpret_cast<const float4*>(brow)[i0], b1 = reinterpret_cast<const float4*>(brow)[i1];
        if (!v0) {
            a0 = z;
            b0 = z;
        }
        if (!v1) {
            a1 = z;
            b1 = z;
        }
        const half8_t ha = {(_Float16)(a0.x * sa), (_Float16)(a0.y * sa), (_Float16)(a0.z * sa), (_Float16)(a0.w * sa),
                            (_Float16)(a1.x * sa), (_Float16)(a1.y * sa), (_Float16)(a1.z * sa), (_Float16)(a1.w * sa)};
        const half8_t hb = {(_Float16)(b0.x * sb), (_Float16)(b0.y * sb), (_Float16)(b0.z * sb), (_Float16)(b0.w * sb),
                            (_Float16)(b1.x * sb), (_Float16)(b1.y * sb), (_Float16)(b1.z * sb), (_Float16)(b1.w * sb)};
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ha, hb, c, 0, 0, 0);
    }
    return c;
}

// gram_tile_f16s with the row operand already fp16 in LDS: a16 points at the lane's row, 8 halves
// per 16-byte unit, dims 8 u .. 8 u + 7 in unit u (zero past d); units at or past ceil(KS / 2) read
// as zero.
template <uint32_t KS, uint32_t UNSC>
__device__ inline float4_t gram_tile_f16a(const uint4* __restrict__ a16, const float* __restrict__ brow,
                                          float sb, uint32_t g) {
    if constexpr (UNSC != 0) sb = 1.0f;
    constexpr uint32_t K32 = (4 * KS + 31) / 32;
    constexpr uint32_t U = (KS + 1) / 2;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    float4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (uint32_t m = 0; m < K32; ++m) {
        const uint32_t ua = 4 * m + g;
        const uint32_t q0 = 8 * m + 2 * g;  // float4 index of the lane's first column dims
        const bool va = ua < U, v0 = q0 < KS, v1 = q0 + 1 < KS;
        uint4 av = a16[va ? ua : 0u];
        if (!va) av = make_uint4(0u, 0u, 0u, 0u);
        float4 b0 = reinterpret_cast<const float4*>(brow)[v0 ? q0 : 0u];
        float4 b1 = reinterpret_cast<const float4*>(brow)[v1 ? q0 + 1 : 0u];
        if (!v0) b0 = z;
        if (!v1) b1 = z;
        const half8_t hb = {(_Float16)(b0.x * sb), (_Float16)(b0.y * sb), (_Float16)(b0.z * sb), (_Float16)(b0.w * sb),
                            (_Float16)(b1.x * sb), (_Float16)(b1.y * sb), (_Float16)(b1.z * sb), (_Float16)(b1.w * sb)};
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8_t, av), hb, c, 0, 0, 0);
    }
    return c;
}

template <uint32_t SCR>
__device__ inline float4_t mfma32(const uint16_t* a, const uint16_t* b, float4_t c) {
    const uint4 av = *reinterpret_cast<const uint4*>(a), bv = *reinterpret_cast<const uint4*>(b);
    if constexpr (SCR == 1)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8_t, av),
                                                      __builtin_bit_cast(half8_t, bv), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, av),
                                                       __builtin_bit_cast(bf16x8_t, bv), c, 0, 0, 0);
}

template <uint32_t KS, uint32_t SCR, uint32_t UNSC>
__global__ __launch_bounds__(kNrThreads) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_non_reverse(NonRevArgs a) {
    static_assert(UNSC == 0 || SCR == 3, "only the fp16 Gram of the FP32 rows has an unscaled form");
    static_assert(SCR == 0 || KS > 0, "the 16-bit screens take a compile-time width");
    constexpr uint32_t K32 = SCR ? (4 * KS + 31) / 32 : 1;
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t k = a.k, d = a.d, ld = a.ld;
    // XCD-aware schedule: the workgroups of one XCD (blockIdx % 8) share one contiguous stretch of
    // this launch's q list, each taking a run of kNrRun consecutive q of it, so a workgroup's
    // consecutive q are neighbours in the list (tree order) and find each other's rows in the
    // XCD's L2. A q's metadata (its position, reverse-list offsets, snapshot slots and first
    // reverse ids) is loaded while the previous q is processed, so only the row gathers start cold.
    const uint32_t per = (a.nq + 7) / 8, xcd = blockIdx.x & 7, nslot = 1;
    uint32_t idx = xcd * per + (blockIdx.x >> 3) * kNrRun;
    const uint32_t i_end = min(min(a.nq, (xcd + 1) * per), idx + kNrRun);
    if (idx >= i_end) return;
    const NrLds L = nr_layout(k, ld, d, SCR);
    const uint32_t rs16 = L.rs16;
    uint16_t* s_b16 = reinterpret_cast<uint16_t*>(lds + L.off_b16);
    uint16_t* s_a16 = reinterpret_cast<uint16_t*>(lds + L.off_a16);
    float* s_brinv = reinterpret_cast<float*>(lds + L.off_brinv);
    float* s_arinv = reinterpret_cast<float*>(lds + L.off_arinv);
    uint4* s_a16f = reinterpret_cast<uint4*>(lds + L.off_a16f);  // screen 3: the chunk rows in fp16
    const uint32_t ldp = L.ldp, d4 = (d + 3) / 4, ldp4 = ldp / 4;
    // t / ldp4 = umulhi(t, ceil(2^32 / ldp4)) exactly for t < 2^32 / ldp4 (ldp4 > 1)
    const uint32_t ldp4_inv = ldp4 > 1 ? (uint32_t)((0x100000000ull + ldp4 - 1) / ldp4) : 0u;
    uint32_t* s_bid = reinterpret_cast<uint32_t*>(lds + L.off_bid);
    float* s_bnrm = reinterpret_cast<float*>(lds + L.off_bnrm);
    float* s_brow = reinterpret_cast<float*>(lds + L.off_brow);
    uint32_t* s_hkey = reinterpret_cast<uint32_t*>(lds + kNrOffHkey);
    uint32_t* s_hcnt = reinterpret_cast<uint32_t*>(lds + kNrOffHcnt);
    // reverse ids of the current q and of the next one (the two swap every q)
    uint32_t* s_rsbuf0 = reinterpret_cast<uint32_t*>(lds + kNrOffRsid);
    uint32_t* s_rsbuf1 = reinterpret_cast<uint32_t*>(lds + kNrOffRsid2);
    uint32_t* s_nsq = reinterpret_cast<uint32_t*>(lds + kNrOffNsq);  // next S(q): snapshot slots
    float* s_nbn = reinterpret_cast<float*>(lds + kNrOffNbn);        // ... and their norms
    float* s_bscl = reinterpret_cast<float*>(lds + kNrOffBscl);
    float* s_ascl = reinterpret_cast<float*>(lds + kNrOffAscl);
    constexpr bool kFold = SCR == 3;                // fp16 Gram of the FP32 LDS rows, folded test
    constexpr bool kScaled = SCR == 3 && UNSC == 0;  // ... with per-row power-of-two scales
    uint64_t* s_meta = reinterpret_cast<uint64_t*>(lds + kNrOffMeta);  // next q: q, roff[q], roff[q+1]
    uint32_t* s_aid = reinterpret_cast<uint32_t*>(lds + kNrOffAid);
    float* s_abnd = reinterpret_cast<float*>(lds + kNrOffAbnd);
    float* s_anrm = reinterpret_cast<float*>(lds + kNrOffAnrm);
    float* s_arow = reinterpret_cast<float*>(lds + L.off_arow);
    uint32_t* s_mask = reinterpret_cast<uint32_t*>(lds + kNrOffMask);
    uint32_t* s_spid = reinterpret_cast<uint32_t*>(lds + kNrOffSpid);
    uint32_t* s_srid = reinterpret_cast<uint32_t*>(lds + kNrOffSrid);
    float* s_sd = reinterpret_cast<float*>(lds + kNrOffSd);
    uint32_t* s_cnt = reinterpret_cast<uint32_t*>(lds + kNrOffCnt);  // [0] accepted, [1] screened
    const uint32_t tid0 = threadIdx.x;


    // Pipeline. While q is processed, the next q's position (two ahead), reverse-list offsets,
    // snapshot slots, their norms and first reverse ids arrive in LDS (global_load_lds: lane i of
    // a wave writes at the wave's base + 4 i or 16 i, no registers held). In q's epilogue the next
    // q's S is ordered and hashed, its rows are put in flight to LDS and, in full passes, its first
    // chunk is fetched into registers, so those loads overlap q's candidate appends.
    const bool early_fetch = a.rnew == nullptr;  // the first chunk's extent needs no new-entry count
    float4 pf_row[kNrPfRows];
    uint32_t pf_snap[kNrPfSnap];
    uint32_t pf_id = kNoId, pf_na = 0;
    float pf_bnd = 0.f, pf_nrm = 0.f;
    // ids[sr] for chunk rows sr < na (16 threads per row): bound, norm, row, snapshot -> registers
    auto fetch_rows = [&](uint32_t tid, const uint32_t* ids, uint32_t na) {
        const uint32_t sr = tid >> 4, sc = tid & 15;
        pf_na = na;
        pf_id = sr < na ? ids[sr] : kNoId;
        if (sc == 0 && sr < na) {
            pf_bnd = a.bound[pf_id];
            pf_nrm = a.norms[pf_id];
        }
        const float* xr = a.X + (size_t)(sr < na ? pf_id : 0) * ld;
#pragma unroll
        for (uint32_t u = 0; u < kNrPfRows; ++u) {
            const uint32_t q4 = sc + 16 * u;
            pf_row[u] = (sr < na && q4 < d4) ? reinterpret_cast<const float4*>(xr)[q4]
                                             : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const uint32_t* sp = a.snap + (size_t)(sr < na ? pf_id : 0) * k;
#pragma unroll
        for (uint32_t u = 0; u < kNrPfSnap; ++u) {
            const uint32_t j = sc + 16 * u;
            pf_snap[u] = (sr < na && j < k) ? sp[j] : 0u;
        }
    };
    // 16-bit screens: S(q)'s FP32 rows are staged in LDS (the 16-bit rows' and the chunk rows'
    // regions, free between q), with q's row in s_xq, then centred on q and stored as 16-bit rows
    // before the q's first chunk is committed. Pass u moves rows 32 u .. 32 u + 31, 16 threads per
    // row (dims 4 sc .. and 64 + 4 sc ..); its 16-bit rows end at 2 rs16 32 (u + 1) bytes, below
    // where pass u + 1's FP32 rows start (4 ldp 32 (u + 1), as rs16 <= 2 ldp), so with the reads of
    // all passes done before one barrier, the writes clobber nothing still to be read.
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float* s_stage = reinterpret_cast<float*>(lds + L.off_b16);
    float* s_xq = reinterpret_cast<float*>(lds + kNrOffXq);
    auto convert_s16 = [&](uint32_t tid) {
        if constexpr (SCR == 1 || SCR == 2) {
            const uint32_t sr = tid >> 4, sc = tid & 15;
            const float4 x0 = sc < d4 ? reinterpret_cast<const float4*>(s_xq)[sc] : zero4;
            const float4 x1 = 16 + sc < d4 ? reinterpret_cast<const float4*>(s_xq)[16 + sc] : zero4;
            float4 v[kNrSPass][2];
#pragma unroll
            for (uint32_t u = 0; u < kNrSPass; ++u) {
                const uint32_t r = 32 * u + sr;
                const bool ok = r < k;
                const float4* src = reinterpret_cast<const float4*>(s_stage + (ok ? r : 0u) * ldp);
                v[u][0] = ok && sc < d4 ? src[sc] : zero4;
                v[u][1] = ok && 16 + sc < d4 ? src[16 + sc] : zero4;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t u = 0; u < kNrSPass; ++u) {
                const uint32_t r = 32 * u + sr;
                if (r < L.kb16) {
                    float nrm, rinv;
                    centre_row16<SCR>(v[u][0], v[u][1], x0, x1, r < k, sc, d, rs16, s_b16 + r * rs16,
                                      nrm, rinv);
                    if (sc == 0) {
                        s_bnrm[r] = r < k ? nrm : 0.f;
                        s_brinv[r] = r < k ? rinv : 0.f;
                    }
                }
            }
        }
    };
    // S from its snapshot slots and norms (s_nsq, s_nbn): ids and norms with the new entries first
    // (wave 0; slots lane, lane + 64; k <= 128), s_cnt[2] = its new-entry count; the hash cleared
    auto prep_order = [&](uint32_t tid) {
        const uint32_t lane = tid & 63, wave = tid >> 6;
        for (uint32_t t = tid; t < kNrHash; t += kNrThreads) s_hkey[t] = kNoId;
        if (tid < kNrBuckets) s_hcnt[tid] = 0;
        if (wave == 0) {
            const uint32_t v0 = lane < k ? s_nsq[lane] : 0u, v1 = lane + 64 < k ? s_nsq[lane + 64] : 0u;
            const float bn0 = lane < k ? s_nbn[lane] : 0.f, bn1 = lane + 64 < k ? s_nbn[lane + 64] : 0.f;
            const float bs0 = kScaled ? scale_exp(bn0) : 0.f;
            const float bs1 = kScaled ? scale_exp(bn1) : 0.f;
            const bool n0 = lane < k && (v0 & kNewBit), n1 = lane + 64 < k && (v1 & kNewBit);
            const uint64_t m0 = __ballot(n0), m1 = __ballot(n1);
            const uint64_t below = (1ull << lane) - 1ull;
            const uint32_t nnew = __popcll(m0) + __popcll(m1);
            const uint32_t b0 = __popcll(m0 & below), b1 = __popcll(m0) + __popcll(m1 & below);
            if (lane < k) {
                const uint32_t pos = n0 ? b0 : nnew + lane - b0, id = v0 & kIdMask;
                s_bid[pos] = id;
                s_bnrm[pos] = bn0;
                if constexpr (kScaled) s_bscl[pos] = bs0;
            }
            if (lane + 64 < k) {
                const uint32_t pos = n1 ? b1 : nnew + lane + 64 - b1, id = v1 & kIdMask;
                s_bid[pos] = id;
                s_bnrm[pos] = bn1;
                if constexpr (kScaled) s_bscl[pos] = bs1;
            }
            if (lane == 0) s_cnt[2] = nnew;
        }
        for (uint32_t t = k + tid; t < L.kb16; t += kNrThreads) {
            s_bid[t] = kNoId;
            s_bnrm[t] = 0.f;
            if constexpr (kScaled) s_bscl[t] = 0.f;
        }
    };
    // after prep_order and a barrier: S's ids hashed, S's rows in flight (to LDS for the FP32
    // Gram, to registers for the 16-bit screens, with q's row qq)
    auto prep_rows = [&](uint32_t tid, uint32_t qq) {
        const uint32_t wave = tid >> 6;
        // Id 0 may be the padding of an unfilled list rather than a member, so it is never hashed:
        // offers of id 0 always go through (the list's id check settles them).
        if (tid < k) {
            const uint32_t id = s_bid[tid];
            if (id != 0) {
                const uint32_t h = nr_hash(id), b = h >> 25;
                const uint32_t slot = atomicAdd(&s_hcnt[b], 1u);
                if (slot < kNrSlots) s_hkey[b * kNrSlots + slot] = (h << 7) | tid;
            }
        }
        // S's rows go from global memory straight into LDS (global_load_lds_dwordx4: lane i of a
        // wave writes 16 bytes at the wave's base + 16 i), every load in flight at once and no
        // registers held. The data is waited for (vmcnt) before the barrier that precedes its
        // first read. Slots past a row's first ceil(d/4) float4 and rows past k are left as they
        // are: the Gram reads dims below 4 ceil(d/4) only, a Gram column depends on its own row
        // only and columns past k are never offered, and the exact distances read dims below d of
        // rows below k.
        if constexpr (SCR == 1 || SCR == 2) {  // q's row, for centring
            if (wave == 0 && (tid & 63) < d4)
                __builtin_amdgcn_global_load_lds(a.X + (size_t)qq * ld + 4 * (tid & 63),
                                                 (__attribute__((address_space(3))) void*)s_xq, 16, 0, 0);
        }
        float* s_dst = SCR == 1 || SCR == 2 ? s_stage : s_brow;
        __attribute__((address_space(3))) float* s_dst_lds = (__attribute__((address_space(3))) float*)(s_dst);
        const uint32_t tot = L.kb16 * ldp4;
#pragma unroll
        for (uint32_t u = 0; u < kNrBRows; ++u) {
            const uint32_t t = tid + u * kNrThreads;
            const uint32_t r = ldp4 == 1 ? t : __umulhi(t, ldp4_inv), q4 = t - r * ldp4;
            if (t < tot && r < k && q4 < d4) {
                __builtin_amdgcn_global_load_lds(
                    a.X + (size_t)s_bid[r] * ld + 4 * q4,
                    (__attribute__((address_space(3))) void*)(s_dst_lds + 4 * (u * kNrThreads + wave * 64)),
                    16, 0, 0);
            }
        }
    };

    // the first q: its metadata straight into LDS, then prepared like every later one
    uint32_t q = a.order[(size_t)idx * a.order_stride];
    uint64_t a0 = a.roff[q], a1 = a.roff[q + 1];
    for (uint32_t t = tid0; t < k; t += kNrThreads) {
        const uint32_t v = a.snap[(size_t)q * k + t];
        s_nsq[t] = v;
        s_nbn[t] = a.norms[v & kIdMask];
    }
    for (uint32_t t = tid0; t < min(a1 - a0, (uint64_t)kNrIds); t += kNrThreads) s_rsbuf0[t] = a.rs[a0 + t];
    uint32_t nq_l = 0;  // the next q's position (lane 0 of each wave)
    if ((tid0 & 63) == 0 && idx + nslot < i_end) nq_l = a.order[(size_t)(idx + nslot) * a.order_stride];
    __syncthreads();
    prep_order(tid0);
    __syncthreads();
    if (a1 > a0) {
        prep_rows(tid0, q);
        if (early_fetch) fetch_rows(tid0, s_rsbuf0, (uint32_t)min(a1 - a0, (uint64_t)kNrChunk));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    uint32_t par = 0;

    for (;;) {
    // Per-thread indices are re-derived each q from an opaque copy of the thread id: hoisted out
    // of this loop they would stay live across it and spill, and a spill reload waits for every
    // load in flight, prefetches included.
    uint32_t tid = tid0;
    asm volatile("" : "+v"(tid));
    const uint32_t lane = tid & 63, wave = tid >> 6;
    if constexpr (SCR == 1 || SCR == 2) {  // S(q)'s staged rows -> 16-bit rows centred on q
        if (a1 > a0) {
            convert_s16(tid);
            __syncthreads();
        }
    }
    uint16_t* s_queue = reinterpret_cast<uint16_t*>(lds + kNrOffQueue) + wave * kNrQueue;
    // commit map: 16 threads per chunk row
    const uint32_t sr = tid >> 4, sc = tid & 15;
    uint32_t* s_rsid = par ? s_rsbuf1 : s_rsbuf0;  // RS(q)'s first kNrIds entries, staged
    uint32_t* s_nrs = par ? s_rsbuf0 : s_rsbuf1;   // the next q's
    const uint32_t nidx = idx + nslot;
    const bool has_next = nidx < i_end;
    const uint32_t nrs = (uint32_t)(a1 - a0);
    const uint32_t nn = a.rnew ? min(a.rnew[q], nrs) : nrs;  // RS(q)'s new rows come first
    // next q: reverse-list offsets (lane 0 of each wave) and snapshot slots (wave 0, into s_nsq)
    uint32_t nq = 0;
    uint64_t nr0 = 0, nr1 = 0;
    if (has_next) {
        nq = lane_bcast(nq_l, 0);
        if (lane == 0) {
            nr0 = a.roff[nq];
            nr1 = a.roff[nq + 1];
        }
        if (wave == 0) {
            const uint32_t* sq = a.snap + (size_t)nq * k;
            if (lane < k) __builtin_amdgcn_global_load_lds(sq + lane, (kLdsU32*)s_nsq, 4, 0, 0);
            if (lane + 64 < k) __builtin_amdgcn_global_load_lds(sq + 64 + lane, (kLdsU32*)(s_nsq + 64), 4, 0, 0);
        }
    }
    auto fetch = [&](uint32_t c_off, uint32_t na) { fetch_rows(tid, s_rsid + c_off, na); };
    // next q: its first reverse ids (into s_nrs) and its S's norms, once its offsets and snapshot
    // slots have landed
    auto prefetch_rs = [&]() {
        if (!has_next) return;
        __builtin_amdgcn_s_waitcnt(0);  // the offsets, and s_nsq for wave 0
        const uint64_t na0 = lane_bcast64(nr0, 0), na1 = lane_bcast64(nr1, 0);
        if (tid == 0) {  // through LDS rather than registers live across the chunk loop
            s_meta[0] = nq;
            s_meta[1] = na0;
            s_meta[2] = na1;
        }
        if (tid < min(na1 - na0, (uint64_t)kNrIds))
            __builtin_amdgcn_global_load_lds(a.rs + na0 + tid, (kLdsU32*)(s_nrs + wave * 64), 4, 0, 0);
        if (wave == 0) {
            if (lane < k)
                __builtin_amdgcn_global_load_lds(a.norms + (s_nsq[lane] & kIdMask), (kLdsU32*)s_nbn, 4, 0, 0);
            if (lane + 64 < k)
                __builtin_amdgcn_global_load_lds(a.norms + (s_nsq[lane + 64] & kIdMask), (kLdsU32*)(s_nbn + 64), 4, 0, 0);

        }
    };
    // old rows pair with S(q)'s new entries only; with none of either there is nothing to offer
    const uint32_t ntile_new = (s_cnt[2] + 15) / 16;
    const uint32_t ne = ntile_new ? nrs : nn;
    if (tid < 2) s_cnt[tid] = 0;
    uint32_t g0 = 0;  // first reverse entry of the staged ids group
    uint32_t n_scr = 0, nsurv = 0;
    if (ne > 0) {
    // chunks never straddle the new/old boundary
    auto cend = [&](uint32_t c) { return c < nn ? min(c + kNrChunk, nn) : min(c + kNrChunk, ne); };

    // ---- chunk commit: registers -> LDS + membership ---------------------------------------
    // A row's 16 threads are one wave's lanes: they zero the row's membership words and then set
    // bits in them, in the wave's LDS order, with no barrier between.
    auto commit = [&]() {
        if (sc < 4) s_mask[sr * 4 + sc] = 0;
        if (sc == 0) {
            s_aid[sr] = pf_id;
            s_abnd[sr] = sr < pf_na ? pf_bnd : 0.f;
            s_anrm[sr] = sr < pf_na ? pf_nrm : 0.f;
            if constexpr (kScaled) s_ascl[sr] = sr < pf_na ? scale_exp(pf_nrm) : 0.f;
        }
#pragma unroll
        for (uint32_t u = 0; u < kNrPfRows; ++u) {
            const uint32_t q4 = sc + 16 * u;
            if (q4 < ldp4) reinterpret_cast<float4*>(s_arow)[sr * ldp4 + q4] = pf_row[u];
        }
        if constexpr (SCR == 3) {  // the Gram's row operand: the row in fp16 (scaled by its power of two)
            static_assert(kNrPfRows == 2, "a thread holds dims 4 sc .. and 64 + 4 sc ..");
            float sa_c = 1.f;
            if constexpr (kScaled) sa_c = ldexpf(1.0f, (int)scale_exp(__shfl(pf_nrm, (int)(tid & ~15u) & 63)));
            uint16_t* h = reinterpret_cast<uint16_t*>(s_a16f + sr * L.a16u);
            auto pk = [&](float x, float y) {
                return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(x * sa_c)) |
                       ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)(y * sa_c)) << 16);
            };
            const uint32_t nh = 8 * ((4 * d4 + 7) / 8);  // halves of the row's units below ceil(KS / 2)
            if (4 * sc < nh)
                *reinterpret_cast<uint2*>(h + 4 * sc) = make_uint2(pk(pf_row[0].x, pf_row[0].y), pk(pf_row[0].z, pf_row[0].w));
            if (64 + 4 * sc < nh)
                *reinterpret_cast<uint2*>(h + 64 + 4 * sc) = make_uint2(pk(pf_row[1].x, pf_row[1].y), pk(pf_row[1].z, pf_row[1].w));
        }
        if constexpr (SCR == 1 || SCR == 2) {  // the row centred on q, 16-bit, with its centred norm
            static_assert(kNrPfRows == 2, "a thread holds dims 4 sc .. and 64 + 4 sc ..");
            float nrm, rinv;
            const bool v = sr < pf_na;
            const float4 x0 = sc < d4 ? reinterpret_cast<const float4*>(s_xq)[sc] : zero4;
            const float4 x1 = 16 + sc < d4 ? reinterpret_cast<const float4*>(s_xq)[16 + sc] : zero4;
            centre_row16<SCR>(pf_row[0], pf_row[1], x0, x1, v, sc, d, rs16, s_a16 + sr * rs16,
                              nrm, rinv);
            if (sc == 0) {
                s_anrm[sr] = v ? nrm : 0.f;
                s_arinv[sr] = v ? rinv : 0.f;
            }
        }
        // member[sr][j]: S(q)[j] is in S(p): the list's id check would reject it. Every bucket read
        // is issued before the first is used.
        if (sr < pf_na) {
            uint4 kb[kNrPfSnap];
            uint32_t q7[kNrPfSnap];
#pragma unroll
            for (uint32_t u = 0; u < kNrPfSnap; ++u) {
                const uint32_t h = nr_hash(pf_snap[u] & kIdMask);
                q7[u] = h << 7;
                kb[u] = reinterpret_cast<const uint4*>(s_hkey)[h >> 25];
            }
#pragma unroll
            for (uint32_t u = 0; u < kNrPfSnap; ++u) {
                const uint32_t m = min(min(kb[u].x ^ q7[u], kb[u].y ^ q7[u]), min(kb[u].z ^ q7[u], kb[u].w ^ q7[u]));
                if (m < 127u) atomicOr(&s_mask[sr * 4 + (m >> 5)], 1u << (m & 31));
            }
        }
    };
    if (!early_fetch) fetch(0, cend(0));
    commit();  // S(q)'s rows and hash were made ready before the loop head's barrier
    __syncthreads();

    // FP32 Gram: |approx - exact| <= (K + 8) 2^-23 (n_p + n_r) over K = 4 ceil(d/4) terms.
    // 16-bit Gram of rows centred on q (n_p, n_r their centred norms): the leaf stage's bound for
    // fp16 (bf16 screens are only used without a margin).
    const bool exact16 = a.exact != 0;
    const float margin_c = SCR == 0 ? (float)(4 * d4 + 8) * (1.0f / 8388608.0f)
                                    : HalfTraits<SCR == 2 ? 1 : 0>::kMargin;
    const uint32_t ksteps = d4;
    const uint32_t ntile = L.kb16 / 16;
    const uint32_t lrow = lane & 15, lk = lane >> 4;

    for (uint32_t c0 = 0; c0 < ne;) {
        const uint32_t c1 = cend(c0);
        const uint32_t na = c1 - c0;
        // the next chunk's loads fly while this chunk is screened
        const uint32_t n0 = c1, n1 = n0 < ne ? cend(n0) : n0;
        const bool next_in_group = n0 < ne && n1 <= g0 + kNrIds;
        if (next_in_group) fetch(n0 - g0, n1 - n0);
        // old rows pair with S(q)'s new columns only (the first ntile_new tiles)
        const uint32_t tt = c0 >= nn ? ntile_new : ntile;

        uint32_t qn = 0;
        auto flush = [&]() {
            __builtin_amdgcn_wave_barrier();
            // up to 8 survivors: eight lanes each (one short pass); more: one lane each, every LDS
            // load of a distance issued at once. Both give dist_ref's value.
            const bool wide = qn <= 8;
            for (uint32_t e0 = 0; e0 < qn; e0 += wide ? 8 : 64) {
                const uint32_t e = e0 + (wide ? lane >> 3 : lane);
                const bool act = e < qn;
                const uint32_t ent = act ? s_queue[e] : 0u;
                const uint32_t i = ent >> 7, cc = ent & 127;
                // S(q)'s FP32 rows: in LDS for the FP32 Gram, in global memory for the 16-bit screens
                const float* rb = SCR == 0 || SCR == 3 ? s_brow + cc * ldp : a.X + (size_t)s_bid[cc] * ld;
                const float* ra = s_arow + i * ldp;
                float ex = 0.f;
                if (wide) {
                    if constexpr (KS > 0) {
                        if (d == 4 * KS) ex = dist_ref_lane8_fixed<4 * KS>(rb, ra, lane);
                        else ex = dist_ref_lane8(rb, ra, d, lane);
                    } else {
                        ex = dist_ref_lane8(rb, ra, d, lane);
                    }
                } else if (act) {
                    if constexpr (KS > 0) {
                        if (d == 4 * KS) ex = dist_ref_fixed<4 * KS>(rb, ra);
                        else ex = dist_ref(rb, ra, d);
                    } else {
                        ex = dist_ref(rb, ra, d);
                    }
                }
                // <=: a distance equal to the bound can still win on its id in the merge
                const bool keep = act && (!wide || (lane & 7) == 0) && ex <= s_abnd[i];
                const uint64_t bal = __ballot(keep);
                if (bal) {
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&s_cnt[0], (uint32_t)__popcll(bal));
                    base = lane_bcast(base, 0);
                    if (keep) {
                        const uint32_t at = base + __popcll(bal & ((1ull << lane) - 1ull));
                        const uint32_t pid = s_aid[i];
                        if (at < kNrSurv) {
                            s_spid[at] = pid;
                            s_srid[at] = s_bid[cc];
                            s_sd[at] = ex;
                        } else {  // list full: append directly
                            const uint32_t slot = atomicAdd(&a.ncnt[pid], 1u);
                            if (slot < a.cap) {
                                a.nbuf_d[(size_t)pid * a.cap + slot] = ex;
                                a.nbuf_id[(size_t)pid * a.cap + slot] = s_bid[cc];
                            }
                        }
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            qn = 0;
        };
        // Jobs (16-row strip, 16-column tile) are dealt to the waves in turn, tile-major: with two
        // strips a wave keeps one strip (the wave count is even), and a chunk of one strip still
        // spreads its tiles over every wave.
        const uint32_t nstrip = (na + 15) / 16;
        const uint32_t wv = __builtin_amdgcn_readfirstlane(wave);
        if (wv < nstrip * tt) {
            const uint32_t strip = nstrip == 2 ? (wv & 1u) : 0u;
            float rn[4], thr0[4];
            uint32_t rr[4], pid[4];
            bool rok[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                rr[i] = 16 * strip + (lane >> 4) * 4 + i;
                rok[i] = rr[i] < na;
                rn[i] = s_anrm[rr[i]];
                pid[i] = s_aid[rr[i]];
                thr0[i] = SCR == 0 || exact16 ? s_abnd[rr[i]] * (1.0f + 1.0f / 65536.0f) : s_abnd[rr[i]];
                // fp16 Gram of the scaled rows (exact form): the test n_p + n_r - 2 g <= thr0 +
                // c (n_p + n_r) is evaluated as -2 g + (1 - c) n_r <= thr0 - (1 - c) n_p, the
                // row's side once per strip; the reordering's FP32 rounding (~2^-23 of the terms)
                // is far inside the margin's slack
                if constexpr (kFold) thr0[i] = thr0[i] - (1.0f - margin_c) * rn[i];
            }
            const float* arow = s_arow + (16 * strip + lrow) * ldp;
            // 16-bit screens: the strip's rows centred on q (lane group lk >> 2 takes halves
            // 8 (lk >> 2) .. + 7 of each 32-dim k-step, for the rows and the columns alike), the
            // inverse scales, and -2 folded into the rows' inverse scales (powers of two: exact)
            const uint16_t* a16 = s_a16 + (16 * strip + lrow) * rs16 + 8 * (lane >> 4);
            float m2r[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                m2r[i] = kScaled ? -2.0f * ldexpf(1.0f, -(int)s_ascl[rr[i]]) : kFold ? -2.0f
                                  : SCR ? -2.0f * s_arinv[rr[i]] : 0.f;
            for (uint32_t j = wv; j < nstrip * tt; j += kNrThreads / 64) {
                const uint32_t t = nstrip == 2 ? j >> 1 : j;
                if (qn + 256 > kNrQueue) flush();
                // membership words of this tile's columns (c >> 5 = t >> 1 for the whole tile)
                uint32_t mw[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) mw[i] = s_mask[rr[i] * 4 + (t >> 1)];
                float4_t acc;
                if constexpr (SCR == 0) {
                    acc = gram_tile_f32<KS>(arow, s_brow + (16 * t + lrow) * ldp, lk, ksteps);
                } else if constexpr (SCR == 3) {
                    const float sb = kScaled ? ldexpf(1.0f, (int)s_bscl[16 * t + lrow]) : 1.f;
                    acc = gram_tile_f16a<KS, UNSC>(s_a16f + (16 * strip + lrow) * L.a16u,
                                                   s_brow + (16 * t + lrow) * ldp, sb, lane >> 4);
                } else {
                    const uint16_t* b16 = s_b16 + (16 * t + lrow) * rs16 + 8 * (lane >> 4);
                    acc = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (uint32_t m = 0; m < K32; ++m) acc = mfma32<SCR>(a16 + 32 * m, b16 + 32 * m, acc);
                }
                const uint32_t c = 16 * t + lrow, cb = c & 31;
                const bool cok = c < k;
                const float cn = s_bnrm[c];  // zero past k
                const float cinv = kScaled ? ldexpf(1.0f, -(int)s_bscl[c]) : kFold ? 1.f : SCR ? s_brinv[c] : 0.f;
                const uint32_t rid = s_bid[c];
                // compact the survivors of this tile into the wave's queue as (row << 7 | col)
                const uint32_t qn_before = qn;
                const float cnm = (1.0f - margin_c) * cn;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float nsum = rn[i] + cn;
                    // FP32 Gram: 2 acc is exact. 16-bit: acc * cinv * (-2 rinv) is exact scaling.
                    const float approx = SCR == 0 ? __builtin_fmaf(-2.0f, acc[i], nsum)
                                         : kFold ? __builtin_fmaf(kScaled ? acc[i] * cinv : acc[i], m2r[i], cnm)
                                                    : __builtin_fmaf(acc[i] * cinv, m2r[i], nsum);
                    const float thr = kFold ? thr0[i]
                                      : SCR == 0 || exact16 ? __builtin_fmaf(margin_c, nsum, thr0[i]) : thr0[i];
                    // <=: a distance equal to a zero bound can still win on its id
                    // every term evaluated (bitwise &): a short-circuit chain here became
                    // branches around the approximate distance
                    const bool ok = cok & rok[i] & (rid != pid[i]) & !((mw[i] >> cb) & 1u) &
                                    (approx <= thr);
                    const uint64_t bal = __ballot(ok);
                    if (ok) {
                        const uint32_t pos = qn + __popcll(bal & ((1ull << lane) - 1ull));
                        s_queue[pos] = (uint16_t)((rr[i] << 7) | c);
                    }
                    qn += __popcll(bal);
                }
                n_scr += qn - qn_before;
            }
        }
        if (qn) flush();
        if (c0 == 0) prefetch_rs();  // the next q's offsets and snapshot slots have had a chunk to land
        __syncthreads();
        if (n0 < ne) {
            if (!next_in_group) {  // next ids group, then a direct fetch
                g0 = n0;
                for (uint32_t t = tid; t < min(nrs - g0, kNrIds); t += kNrThreads)
                    s_rsid[t] = a.rs[a0 + g0 + t];
                __syncthreads();
                fetch(0, n1 - n0);
            }
            commit();
            __syncthreads();
        }
        c0 = c1;
    }
    // the chunk loop ends with a barrier: the list is complete
    nsurv = s_cnt[0];
    } else {
        __syncthreads();  // every wave has read the s_meta entries prefetch_rs overwrites
        prefetch_rs();
    }
    // q's accepted offers go to the candidate buffers
    auto append = [&]() {
        for (uint32_t t = tid; t < min(nsurv, kNrSurv); t += kNrThreads) {
            const uint32_t pid = s_spid[t];
            const uint32_t slot = atomicAdd(&a.ncnt[pid], 1u);
            if (slot < a.cap) {
                a.nbuf_d[(size_t)pid * a.cap + slot] = s_sd[t];
                a.nbuf_id[(size_t)pid * a.cap + slot] = s_srid[t];
            }
        }
        if (a.counters) {  // diagnostics only: two same-address atomics per workgroup
            if (lane == 0 && n_scr) atomicAdd(&s_cnt[1], n_scr);
            __syncthreads();
            if (tid == 0) {
                atomicAdd(reinterpret_cast<unsigned long long*>(&a.counters[0]), (unsigned long long)s_cnt[1]);
                atomicAdd(reinterpret_cast<unsigned long long*>(&a.counters[1]), (unsigned long long)nsurv);
            }
        }
    };
    // this q's chunk registers are dead from here on: say so, so that they are not carried
    // around the loop on the paths that do not fetch the next q's first chunk
#pragma unroll
    for (uint32_t u = 0; u < kNrPfRows; ++u) pf_row[u] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (uint32_t u = 0; u < kNrPfSnap; ++u) pf_snap[u] = 0u;
    pf_id = kNoId;
    pf_na = 0;
    pf_bnd = pf_nrm = 0.f;
    if (!has_next) {
        append();
        break;
    }
    // The next q: S ordered and hashed, its rows in flight to LDS, its first chunk fetched (always
    // assigned, empty when there is none or it is fetched at the loop head, so that nothing of
    // this q's registers stays live across the loop), and the position of the q after it.
    __builtin_amdgcn_s_waitcnt(0);  // its snapshot slots, norms and reverse ids
    prep_order(tid);
    __syncthreads();
    const uint32_t nrs1 = (uint32_t)(s_meta[2] - s_meta[1]);
    if (nrs1 > 0) prep_rows(tid, (uint32_t)s_meta[0]);
    fetch_rows(tid, s_nrs, early_fetch ? min(nrs1, kNrChunk) : 0u);
    if (lane == 0 && nidx + nslot < i_end) nq_l = a.order[(size_t)(nidx + nslot) * a.order_stride];
    append();  // the atomics' round trips overlap the loads above
    __builtin_amdgcn_s_waitcnt(0);  // the next q's rows, first chunk and position have landed
    __syncthreads();                // ... and the next q reuses the LDS
    idx = nidx;
    q = (uint32_t)s_meta[0];
    a0 = s_meta[1];
    a1 = s_meta[2];
    par ^= 1u;
    }
}

// One wave per point: merge its candidate buffer into the list; overflowed points are queued for
// the direct kernel instead.
__global__ __launch_bounds__(256) void k_non_merge(NonRevArgs a, uint32_t* __restrict__ ovf_cnt,
                                                   uint32_t* __restrict__ ovf_list,
                                                   uint32_t* __restrict__ ovf_flag) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= a.n) return;
    // the count, the list and the first 64 candidate slots are loaded together (slots past the
    // count hold stale values and are masked below), so a point waits out one memory latency
    const uint32_t k = a.k;
    uint32_t* gid = a.knn_ids + (size_t)p * k;
    float* gd = a.knn_d + (size_t)p * k;
    const size_t cb = (size_t)p * a.cap;
    const uint32_t m = a.ncnt[p];
    const uint32_t id_l = lane < a.cap ? a.nbuf_id[cb + lane] : 0u;
    const float d_l = lane < a.cap ? a.nbuf_d[cb + lane] : 0.f;
    const uint32_t lid0 = lane < k ? gid[lane] : kNoId, lid1 = lane + 64 < k ? gid[lane + 64] : kNoId;
    const float ld0 = lane < k ? gd[lane] : kFltMax, ld1 = lane + 64 < k ? gd[lane + 64] : kFltMax;
    if (m == 0) return;
    if (m > a.cap) {
        if (lane == 0 && atomicExch(&ovf_flag[p], 1u) == 0u) ovf_list[atomicAdd(ovf_cnt, 1u)] = p;
        return;
    }
    __shared__ uint64_t scr[4][128];
    WaveList lst;
    lst.k0 = lane < k ? WaveList::key(ld0, lid0) : WaveList::kEmpty;
    lst.k1 = lane + 64 < k ? WaveList::key(ld1, lid1) : WaveList::kEmpty;
    lst.set_bound(k);
    bool changed = false;
    for (uint32_t c0 = 0; c0 < m; c0 += 64) {
        uint64_t x = WaveList::kEmpty;
        if (c0 + lane < m)
            x = c0 == 0 ? WaveList::key(d_l, id_l)
                        : WaveList::key(a.nbuf_d[cb + c0 + lane], a.nbuf_id[cb + c0 + lane]);
        changed |= lst.offer64(x, k, lane, scr[threadIdx.x >> 6]);
    }
    if (changed) {
        lst.store(gid, gd, k, lane);
        if (lane == 0) a.bound[p] = lst.bound;
    }
}

// ---- direct kernel ---------------------------------------------------------------------------------
constexpr uint32_t kNonWaves = 4;

__global__ __launch_bounds__(kNonWaves * 64) void k_non_direct(NonArgs a, const uint32_t* plist,
                                                               uint32_t np) {
    __shared__ __align__(16) float rowbuf[kNonWaves][kMaxDim];
    __shared__ uint64_t scr[kNonWaves][128];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t w = blockIdx.x * kNonWaves + wave;
    if (w >= np) return;
    const uint32_t p = plist ? plist[w] : w;
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
        const uint32_t q = sp[a2] & kIdMask;
        if (!(sp[a2] & kExpBit)) continue;
        if (a2 > 0 && (sp[a2 - 1] & kIdMask) == q) continue;  // repeated id 0 of an unfilled list
        const uint32_t* sq = a.snap + (size_t)q * k;
        for (uint32_t c0 = 0; c0 < k; c0 += 64) {
            uint64_t x = WaveList::kEmpty;
            if (c0 + lane < k) {
                const uint32_t cid = sq[c0 + lane] & kIdMask;
                if (cid != p) x = WaveList::key(dist_ref(a.X + (size_t)cid * ld, xr, d), cid);
            }
            changed |= lst.offer64(x, k, lane, scr[wave]);
        }
    }
    if (changed) {
        lst.store(gid, gd, k, lane);
        if (lane == 0) a.bound[p] = lst.bound;
    }
}

// FP32 squared norms for the screen (any summation order: the screen's margin covers it). Sixteen
// lanes per row, float4 loads, so a wave reads four whole rows at a time.
__global__ __launch_bounds__(256) void k_row_norms(const float* __restrict__ X, uint32_t ld,
                                                   uint32_t d, uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * 16 + (threadIdx.x >> 4), sc = threadIdx.x & 15;
    const uint32_t d4 = (d + 3) / 4;
    float s = 0.f;
    if (i < n) {
        const float4* x = reinterpret_cast<const float4*>(X + (size_t)i * ld);
        for (uint32_t q = sc; q < d4; q += 16) {  // the row's first ceil(d/4) float4 only
            const float4 v = x[q];
            const uint32_t j = 4 * q;
            if (j + 0 < d) s = __builtin_fmaf(v.x, v.x, s);
            if (j + 1 < d) s = __builtin_fmaf(v.y, v.y, s);
            if (j + 2 < d) s = __builtin_fmaf(v.z, v.z, s);
            if (j + 3 < d) s = __builtin_fmaf(v.w, v.w, s);
        }
    }
    for (uint32_t o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (i < n && sc == 0) out[i] = s;
}

void launch_row_norms(const float* X, uint32_t ld, uint32_t d, uint32_t n, float* out,
                      hipStream_t st) {
    if (n) k_row_norms<<<(n + 15) / 16, 256, 0, st>>>(X, ld, d, n, out);
}

// The smallest nonzero and the largest squared norm, as the bits of non-negative floats (whose
// unsigned order is their order): out[0] starts at +inf, out[1] at 0. A grid of a few hundred
// blocks, so two same-address atomics per block.
__global__ void k_norm_range_init(uint32_t* out) {
    out[0] = 0x7F800000u;
    out[1] = 0u;
}
__global__ __launch_bounds__(256) void k_norm_range(const float* __restrict__ nrm, uint32_t n,
                                                    uint32_t* __restrict__ out) {
    uint32_t lo = 0x7F800000u, hi = 0u;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const uint32_t b = __float_as_uint(nrm[i]);
        if (b != 0u) lo = min(lo, b);
        hi = max(hi, b);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    __shared__ uint32_t s_lo[4], s_hi[4];
    if ((threadIdx.x & 63) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        atomicMin(&out[0], lo);
        atomicMax(&out[1], hi);
    }
}

void launch_norm_range(const float* norms, uint32_t n, uint32_t* out, hipStream_t st) {
    k_norm_range_init<<<1, 1, 0, st>>>(out);
    if (n) k_norm_range<<<std::min<uint32_t>(512, (n + 255) / 256), 256, 0, st>>>(norms, n, out);
}

// ---- launchers ----------------------------------------------------------------------------------------
void launch_non_snapshot(const uint32_t* ids, uint32_t* snap, uint32_t n, uint32_t k,
                         const uint32_t* prev, uint32_t rank_lo, uint32_t* rcnt, uint32_t* rnew,
                         uint32_t* pkey, uint32_t* pval, hipStream_t st) {
    if (!n) return;
    if (rcnt) (void)hipMemsetAsync(rcnt, 0, (size_t)n * 4, st);
    if (rcnt && rnew) (void)hipMemsetAsync(rnew, 0, (size_t)n * 4, st);
    k_non_snapshot_sorted<<<(n + 3) / 4, 256, 0, st>>>(ids, snap, n, k, prev, rank_lo, rcnt,
                                                       rcnt ? rnew : nullptr,
                                                       rcnt ? nullptr : pkey, rcnt ? nullptr : pval);
}

// roff[q] = the first position of a key >= q in the sorted keys (q <= n): the reverse list of q
// spans [roff[q], roff[q+1]); keys past the last list are 0xFFFFFFFF.
__global__ __launch_bounds__(256) void k_roff_search(const uint32_t* __restrict__ keys, uint64_t count,
                                                     uint32_t n, uint64_t* __restrict__ roff) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q > n) return;
    uint64_t lo = 0, hi = count;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] < q) lo = mid + 1;
        else hi = mid;
    }
    roff[q] = lo;
}

void launch_roff_search(const uint32_t* keys, uint64_t count, uint32_t n, uint64_t* roff, hipStream_t st) {
    k_roff_search<<<(n + 1 + 255) / 256, 256, 0, st>>>(keys, count, n, roff);
}

void launch_non_direct(const NonArgs& a, const uint32_t* plist, uint32_t np, hipStream_t st) {
    if (np) k_non_direct<<<(np + kNonWaves - 1) / kNonWaves, kNonWaves * 64, 0, st>>>(a, plist, np);
}

size_t scan_scratch_elems(uint32_t n) { return (n + kScanBlock - 1) / kScanBlock + 1; }

void launch_exclusive_scan(const uint32_t* in, uint32_t n, uint64_t* out, uint64_t* scratch,
                           hipStream_t st) {
    const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb == 0) {
        (void)hipMemsetAsync(out, 0, sizeof(uint64_t), st);
        return;
    }
    k_scan_sums<<<nb, kScanThreads, 0, st>>>(in, n, scratch);
    k_scan_bases<<<1, 1024, 0, st>>>(scratch, nb);
    k_scan_apply<<<nb, kScanThreads, 0, st>>>(in, n, scratch, out);
    k_scan_total<<<1, 1, 0, st>>>(in, n, out);
}

void launch_reverse_lists(const uint32_t* snap, uint32_t n, uint32_t k, uint32_t* rcnt,
                          uint32_t* rnew, uint64_t* roff, uint64_t* scratch, uint32_t* rfill,
                          uint32_t* rfill_old, uint32_t* rs, hipStream_t st) {
    if (!n || !k) return;
    const unsigned g = (n + 3) / 4;
    // rnew == nullptr: no new/old split (every entry filled from the front)
    (void)hipMemsetAsync(rfill, 0, (size_t)n * 4, st);
    if (rnew) (void)hipMemsetAsync(rfill_old, 0, (size_t)n * 4, st);
    launch_exclusive_scan(rcnt, n, roff, scratch, st);
    k_rs_fill<<<g, 256, 0, st>>>(snap, n, k, roff, rcnt, rfill, rnew ? rfill_old : nullptr, rs);
}

void launch_non_reverse(const NonRevArgs& a, hipStream_t st) {
    if (!a.nq) return;
    const uint32_t scr = non_screen16_supported(a.d) ? a.screen16 : 0u;
    const size_t lds = non_reverse_lds(a.k, a.ld, a.d, scr);
    auto go = [&](auto kernel) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        // one workgroup per run of kNrRun q in each XCD's stretch, a multiple of the 8 XCDs
        const uint32_t grid = 8 * (((a.nq + 7) / 8 + kNrRun - 1) / kNrRun);
        kernel<<<grid, kNrThreads, lds, st>>>(a);
    };
    const uint32_t unsc = scr == 3 && a.unscaled ? 1000u : 0u;
    switch ((a.d + 3) / 4 + 100 * scr + unsc) {  // the FP32 Gram's k-steps, the screen, unscaled rows
        case 16: go(&k_non_reverse<16, 0, 0>); break;
        case 25: go(&k_non_reverse<25, 0, 0>); break;
        case 32: go(&k_non_reverse<32, 0, 0>); break;
        case 116: go(&k_non_reverse<16, 1, 0>); break;
        case 125: go(&k_non_reverse<25, 1, 0>); break;
        case 132: go(&k_non_reverse<32, 1, 0>); break;
        case 216: go(&k_non_reverse<16, 2, 0>); break;
        case 225: go(&k_non_reverse<25, 2, 0>); break;
        case 232: go(&k_non_reverse<32, 2, 0>); break;
        case 316: go(&k_non_reverse<16, 3, 0>); break;
        case 325: go(&k_non_reverse<25, 3, 0>); break;
        case 332: go(&k_non_reverse<32, 3, 0>); break;
        case 1316: go(&k_non_reverse<16, 3, 1>); break;
        case 1325: go(&k_non_reverse<25, 3, 1>); break;
        case 1332: go(&k_non_reverse<32, 3, 1>); break;
        default: go(&k_non_reverse<0, 0, 0>); break;
    }
}

void launch_non_merge(const NonRevArgs& a, uint32_t* ovf_cnt, uint32_t* ovf_list, uint32_t* ovf_flag,
                      hipStream_t st) {
    if (a.n) k_non_merge<<<(a.n + 3) / 4, 256, 0, st>>>(a, ovf_cnt, ovf_list, ovf_flag);
}

}  // namespace knng
