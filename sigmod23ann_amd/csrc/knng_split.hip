// Level-synchronous recursive 2-means split (splitKmeansBinaryProcess, SolutionKmeans.hpp:142-311)
// for gfx950. Every node of one tree level is processed by the same few launches:
//
//   start_pair   one block per node: farthest pair among the sampled rows (kmeansStartVecs,
//                :68-138) and the first hyperplane.
//   lloyd_chunk  one block per (node, chunk): side test of the chunk's rows and FP64 per-side sums
//                in entry order (:166-175 small branch, the whole range is one chunk; :218-249
//                large branch, chunks of 256 sample draws).
//   lloyd_finish one block per node: chunk partials added in chunk order, centres = sum / count
//                (:182-185), next hyperplane.
//   side_count   one block per tile of <= 1024 positions: final side test over the full range
//                (:189-197, :262-289) on rows staged through LDS by coalesced loads, one flag
//                byte per position, lo count per tile.
//   scan / node_lo / scatter / copy_back: stable partition of each node's positions (lo first).
//
// The side test reads one 400-byte row per position and is HBM bound; everything else is small.
#include "knng_device.hpp"
#include "knng_kernels.hpp"

namespace knng {

constexpr uint32_t kSplitThreads = 256;
constexpr uint32_t kMaxStartRows = 1024;

// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kSplitThreads) void k_start_pair(
    const float* __restrict__ X, uint32_t ld, uint32_t d, const uint32_t* __restrict__ perm,
    NodeDesc* __restrict__ nodes, float* __restrict__ c1, float* __restrict__ c2,
    float* __restrict__ coef, float* __restrict__ off) {
    NodeDesc& nd = nodes[blockIdx.x];
    if (nd.status != kNodeActive) return;
    __shared__ uint32_t rows[kMaxStartRows];
    __shared__ float best_d[kSplitThreads];
    __shared__ uint32_t best_u[kSplitThreads];
    __shared__ __align__(16) float between[kMaxDim];
    __shared__ float off_slot;
    const uint32_t s = nd.s;
    for (uint32_t t = threadIdx.x; t < s; t += blockDim.x)
        rows[t] = perm[nd.start + draw_below(nd.akey, t, nd.size)];
    __syncthreads();
    float bd = kFltMin;
    uint32_t bu = 0xFFFFFFFFu;
    const uint32_t total = s * s;
    for (uint32_t u = threadIdx.x; u < total; u += blockDim.x) {
        const uint32_t i = u / s, j = u - i * s;
        if (j <= i) continue;
        const float dd = dist_ref(X + (size_t)rows[i] * ld, X + (size_t)rows[j] * ld, d);
        if (dd > bd) {  // first maximum in (i, j) order within this thread (u ascending)
            bd = dd;
            bu = u;
        }
    }
    best_d[threadIdx.x] = bd;
    best_u[threadIdx.x] = bu;
    __syncthreads();
    for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const float od = best_d[threadIdx.x + h];
            const uint32_t ou = best_u[threadIdx.x + h];
            const float md = best_d[threadIdx.x];
            const uint32_t mu = best_u[threadIdx.x];
            if (ou != 0xFFFFFFFFu && (mu == 0xFFFFFFFFu || od > md || (od == md && ou < mu))) {
                best_d[threadIdx.x] = od;
                best_u[threadIdx.x] = ou;
            }
        }
        __syncthreads();
    }
    const uint32_t u = best_u[0];
    if (u == 0xFFFFFFFFu) {  // no pair beyond FLT_MIN: the reference dereferences null here
        if (threadIdx.x == 0) nd.status = kNodeFailed;
        return;
    }
    const float* ri = X + (size_t)rows[u / s] * ld;
    const float* rj = X + (size_t)rows[u % s] * ld;
    float* a = c1 + (size_t)blockIdx.x * ld;
    float* b = c2 + (size_t)blockIdx.x * ld;
    for (uint32_t j = threadIdx.x; j < ld; j += blockDim.x) {
        a[j] = ri[j];
        b[j] = rj[j];
    }
    __syncthreads();
    const float o = block_hyperplane(a, b, coef + (size_t)blockIdx.x * ld, between, d, ld, &off_slot);
    if (threadIdx.x == 0) off[blockIdx.x] = o;
}

// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kSplitThreads) void k_lloyd_chunk(
    const float* __restrict__ X, uint32_t ld, uint32_t d, const uint32_t* __restrict__ perm,
    const NodeDesc* __restrict__ nodes, const ChunkDesc* __restrict__ chunks,
    const float* __restrict__ coef, const float* __restrict__ off, double* __restrict__ part,
    uint32_t* __restrict__ pcnt) {
    const ChunkDesc ch = chunks[blockIdx.x];
    const NodeDesc nd = nodes[ch.node];
    if (nd.status != kNodeActive) return;
    __shared__ uint32_t rows[kMaxSmallNode];
    __shared__ uint8_t flags[kMaxSmallNode];
    __shared__ __align__(16) float cf[kMaxDim];
    __shared__ uint32_t nlo;
    const bool large = nd.m != 0;
    for (uint32_t j = threadIdx.x; j < ld; j += blockDim.x) cf[j] = coef[(size_t)ch.node * ld + j];
    if (threadIdx.x == 0) nlo = 0;
    __syncthreads();
    const float o = off[ch.node];
    uint32_t mylo = 0;
    for (uint32_t e = threadIdx.x; e < ch.count; e += blockDim.x) {
        const uint32_t idx = ch.begin + e;
        const uint32_t pos = nd.start + (large ? draw_below(nd.akey, idx, nd.size) : idx);
        const uint32_t r = perm[pos];
        rows[e] = r;
        const bool lo = dot_ref(cf, X + (size_t)r * ld, d) >= o;
        flags[e] = lo;
        mylo += lo;
    }
    atomicAdd(&nlo, mylo);
    __syncthreads();
    // FP64 sums in entry order, one thread per (side, dim). The loads of 16 entries are issued
    // before their adds, which stay in entry order.
    constexpr uint32_t kAhead = 16;
    for (uint32_t t = threadIdx.x; t < 2 * ld; t += blockDim.x) {
        const uint32_t side = t / ld, j = t - side * ld;
        const uint8_t want = side == 0 ? 1 : 0;
        double acc = 0.0;
        if (j < d) {
            const float* xj = X + j;
            uint32_t e = 0;
            for (; e + kAhead <= ch.count; e += kAhead) {
                float v[kAhead];
#pragma unroll
                for (uint32_t u = 0; u < kAhead; ++u) v[u] = xj[(size_t)rows[e + u] * ld];
#pragma unroll
                for (uint32_t u = 0; u < kAhead; ++u)
                    if (flags[e + u] == want) acc += (double)v[u];
            }
            for (; e < ch.count; ++e)
                if (flags[e] == want) acc += (double)xj[(size_t)rows[e] * ld];
        }
        part[((size_t)blockIdx.x * 2 + side) * ld + j] = acc;
    }
    if (threadIdx.x == 0) {
        pcnt[blockIdx.x * 2 + 0] = nlo;
        pcnt[blockIdx.x * 2 + 1] = ch.count - nlo;
    }
}

__global__ __launch_bounds__(kSplitThreads) void k_lloyd_finish(
    NodeDesc* __restrict__ nodes, const double* __restrict__ part,
    const uint32_t* __restrict__ pcnt, float* __restrict__ c1, float* __restrict__ c2,
    float* __restrict__ coef, float* __restrict__ off, uint32_t d, uint32_t ld) {
    NodeDesc& nd = nodes[blockIdx.x];
    if (nd.status != kNodeActive) return;
    __shared__ uint64_t cnt[2];
    __shared__ __align__(16) float between[kMaxDim];
    __shared__ float off_slot;
    if (threadIdx.x < 2) {
        uint64_t c = 0;
        for (uint32_t k = 0; k < nd.nchunks; ++k) c += pcnt[(nd.chunk0 + k) * 2 + threadIdx.x];
        cnt[threadIdx.x] = c;
    }
    __syncthreads();
    if (cnt[0] == 0 || cnt[1] == 0) {  // an empty side restarts the node (:177-179, :251-253)
        if (threadIdx.x == 0) nd.status = kNodeFailed;
        return;
    }
    float* a = c1 + (size_t)blockIdx.x * ld;
    float* b = c2 + (size_t)blockIdx.x * ld;
    for (uint32_t t = threadIdx.x; t < 2 * ld; t += blockDim.x) {
        const uint32_t side = t / ld, j = t - side * ld;
        float v = 0.f;
        if (j < d) {
            double acc = 0.0;
            for (uint32_t k = 0; k < nd.nchunks; ++k)
                acc += part[((size_t)(nd.chunk0 + k) * 2 + side) * ld + j];
            v = (float)(acc / (double)cnt[side]);
        }
        (side == 0 ? a : b)[j] = v;
    }
    __syncthreads();
    const float o = block_hyperplane(a, b, coef + (size_t)blockIdx.x * ld, between, d, ld, &off_slot);
    if (threadIdx.x == 0) off[blockIdx.x] = o;
}

// ---------------------------------------------------------------------------------------------
// The final side test reads every position's row once. A tile's rows are staged into LDS in passes
// of kSideRows with global_load_lds (16 bytes per lane, consecutive lanes on consecutive 16-byte
// words of a row, so a wave's load instruction reads 1 KB of whole rows), then each thread takes one
// staged row through the reference-order dot. LDS rows of ld floats (ld + 4 when ld is a multiple
// of 8) put the 16-byte reads of 16 consecutive lanes in distinct banks.
constexpr uint32_t kSideThreads = 128;
constexpr uint32_t kSideRows = kSideThreads;  // one row per thread and pass

__host__ __device__ inline uint32_t side_row_stride(uint32_t ld) { return ld % 8 == 0 ? ld + 4 : ld; }

__global__ __launch_bounds__(kSideThreads) void k_side_count(
    const float* __restrict__ X, uint32_t ld, uint32_t d, const uint32_t* __restrict__ perm,
    const NodeDesc* __restrict__ nodes, const TileDesc* __restrict__ tiles,
    const float* __restrict__ coef, const float* __restrict__ off, uint8_t* __restrict__ flags,
    uint32_t* __restrict__ tile_lo) {
    extern __shared__ __align__(16) float s_rows[];  // kSideRows rows of side_row_stride(ld)
    __shared__ __align__(16) float cf[kMaxDim];
    __shared__ uint32_t s_id[kSideRows];
    __shared__ uint32_t nlo;
    const TileDesc tl = tiles[blockIdx.x];
    const NodeDesc nd = nodes[tl.node];
    if (nd.status == kNodeFailed) {
        if (threadIdx.x == 0) tile_lo[blockIdx.x] = 0;
        return;
    }
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t j = tid; j < ld; j += kSideThreads) cf[j] = coef[(size_t)tl.node * ld + j];
    if (tid == 0) nlo = 0;
    const float o = off[tl.node];
    const uint32_t ldp = side_row_stride(ld), ldp4 = ldp / 4, d4 = (d + 3) / 4;
    // t / ldp4 = umulhi(t, ceil(2^32 / ldp4)) exactly for t < 2^32 / ldp4 (ldp4 > 1)
    const uint32_t inv = ldp4 > 1 ? (uint32_t)((0x100000000ull + ldp4 - 1) / ldp4) : 0u;
    __attribute__((address_space(3))) float* s_rows_lds = (__attribute__((address_space(3))) float*)(s_rows);
    uint32_t mylo = 0;
    const uint32_t base = nd.start + tl.begin;
    for (uint32_t s0 = 0; s0 < tl.count; s0 += kSideRows) {
        const uint32_t ns = min(kSideRows, tl.count - s0);
        s_id[tid] = tid < ns ? perm[base + s0 + tid] : 0u;
        __syncthreads();  // ids in; the previous pass's rows all read
        const uint32_t tot = ns * ldp4;
        for (uint32_t t0 = 0; t0 < tot; t0 += kSideThreads) {
            const uint32_t t = t0 + tid;
            const uint32_t r = ldp4 == 1 ? t : __umulhi(t, inv), q4 = t - r * ldp4;
            if (t < tot && q4 < d4)
                __builtin_amdgcn_global_load_lds(
                    X + (size_t)s_id[r] * ld + 4 * q4,
                    (__attribute__((address_space(3))) void*)(s_rows_lds + 4 * (t0 + wave * 64)), 16, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid < ns) {
            const bool lo = dot_ref(cf, s_rows + tid * ldp, d) >= o;
            flags[base + s0 + tid] = lo;
            mylo += lo;
        }
    }
    (void)lane;
    atomicAdd(&nlo, mylo);
    __syncthreads();
    if (tid == 0) tile_lo[blockIdx.x] = nlo;
}

size_t side_count_lds(uint32_t ld) { return (size_t)kSideRows * side_row_stride(ld) * sizeof(float); }

// ---- side test on a 16-bit copy of the rows -------------------------------------------------------
// The same decision, dot_ref(coef, x, d) >= off, with each row read as fp16 (ld16 halves, zero past
// d: half the bytes of the FP32 row). A thread forms s~ = sum x~_i c_i and A = sum |x~_i c_i| over
// the converted row x~ (FP32 arithmetic, any order) and bounds the distance to the reference's FP32
// value s: the conversion moves each value by at most u |x~_i| (u = 2^-11 / (1 - 2^-11)), or 2^-25 in
// fp16's subnormal range; each FP32 sum is within gamma_d = d 2^-24 / (1 - d 2^-24) of its exact
// value relative to the sum of its absolute terms. For d <= 256,
//   |s~ - s| <= (u + 2 gamma_d (1 + u)) / (1 - gamma_d) A + 2^-25 (1 + gamma_d) sum |c_i|
//            <= 5.2e-4 A + 3.0e-8 sum |c_i|,
// and the test takes E = 5.4e-4 A + 3.2e-8 sum |c_i| + 2^-22 (|s~| + |off|), which also covers the
// rounding of s~ - off and of E. Where |s~ - off| <= E (or a value overflowed fp16: s~ and A are
// then infinite or NaN, and neither test holds) the row's FP32 values decide through dot_ref itself,
// after the tile's last pass, so the flags equal the FP32 test's bit for bit.
constexpr uint32_t kSideTileMax = kTileSize;  // positions per tile (deferred rows at most)

__host__ __device__ inline uint32_t side16_units(uint32_t ld16) { return ld16 / 8 + 1; }  // +1: bank stagger

__global__ __launch_bounds__(kSideThreads) void k_side_count16(
    const float* __restrict__ X, uint32_t ld, const uint16_t* __restrict__ X16, uint32_t ld16,
    uint32_t d, const uint32_t* __restrict__ perm, const NodeDesc* __restrict__ nodes,
    const TileDesc* __restrict__ tiles, const float* __restrict__ coef, const float* __restrict__ off,
    uint8_t* __restrict__ flags, uint32_t* __restrict__ tile_lo) {
    extern __shared__ __align__(16) uint4 s_rows16[];  // kSideRows rows of side16_units(ld16) x 8 halves
    __shared__ __align__(16) float cf[kMaxDim];
    __shared__ uint32_t s_id[kSideRows];
    __shared__ uint16_t s_defer[kSideTileMax];  // tile positions left to the FP32 rows
    __shared__ uint32_t nlo, ndefer;
    __shared__ float s_csum;
    const TileDesc tl = tiles[blockIdx.x];
    const NodeDesc nd = nodes[tl.node];
    if (nd.status == kNodeFailed) {
        if (threadIdx.x == 0) tile_lo[blockIdx.x] = 0;
        return;
    }
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // coef past d reads as zero: the fp16 rows are zero there, and 0 * 0 adds nothing
    for (uint32_t j = tid; j < kMaxDim; j += kSideThreads) cf[j] = j < d ? coef[(size_t)tl.node * ld + j] : 0.f;
    if (tid == 0) nlo = ndefer = 0;
    __syncthreads();
    if (wave == 0) {
        float c = 0.f;
        for (uint32_t j = lane; j < kMaxDim; j += 64) c += fabsf(cf[j]);
        c = wave_sum(c);
        if (lane == 0) s_csum = c;
    }
    const float o = off[tl.node];
    const uint32_t rs = side16_units(ld16), d8 = (d + 7) / 8;
    // t / rs = umulhi(t, ceil(2^32 / rs)) exactly for t < 2^32 / rs (rs > 1)
    const uint32_t inv = (uint32_t)((0x100000000ull + rs - 1) / rs);
    __attribute__((address_space(3))) uint4* s_rows_lds = (__attribute__((address_space(3))) uint4*)(s_rows16);
    uint32_t mylo = 0;
    const uint32_t base = nd.start + tl.begin;
    for (uint32_t s0 = 0; s0 < tl.count; s0 += kSideRows) {
        const uint32_t ns = min(kSideRows, tl.count - s0);
        s_id[tid] = tid < ns ? perm[base + s0 + tid] : 0u;
        __syncthreads();  // ids in (and s_csum); the previous pass's rows all read
        const uint32_t tot = ns * rs;
        for (uint32_t t0 = 0; t0 < tot; t0 += kSideThreads) {
            const uint32_t t = t0 + tid;
            const uint32_t r = __umulhi(t, inv), q = t - r * rs;
            if (t < tot && q < d8)
                __builtin_amdgcn_global_load_lds(
                    X16 + (size_t)s_id[r] * ld16 + 8 * q,
                    (__attribute__((address_space(3))) void*)(s_rows_lds + t0 + wave * 64), 16, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid < ns) {
            const uint4* row = s_rows16 + tid * rs;
            float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            float aab[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            for (uint32_t q = 0; q < d8; ++q) {
                const uint4 w = row[q];
                const uint32_t wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (uint32_t e = 0; e < 8; ++e) {
                    const float x = (float)__builtin_bit_cast(_Float16, (uint16_t)(wv[e >> 1] >> (16 * (e & 1))));
                    const float c = cf[8 * q + e];
                    acc[e] = __builtin_fmaf(x, c, acc[e]);
                    aab[e] = __builtin_fmaf(fabsf(x), fabsf(c), aab[e]);
                }
            }
            float sv = 0.f, av = 0.f;
#pragma unroll
            for (uint32_t e = 0; e < 8; ++e) {
                sv += acc[e];
                av += aab[e];
            }
            const float e_b = 5.4e-4f * av + 3.2e-8f * s_csum + 0x1p-22f * (fabsf(sv) + fabsf(o));
            const float diff = sv - o;
            if (diff > e_b || diff < -e_b) {
                const bool lo = diff > e_b;
                flags[base + s0 + tid] = lo;
                mylo += lo;
            } else {
                s_defer[atomicAdd(&ndefer, 1u)] = (uint16_t)(s0 + tid);
            }
        }
    }
    __syncthreads();
    // the deferred rows: the reference-order FP32 dot of the row itself
    for (uint32_t i = tid; i < ndefer; i += kSideThreads) {
        const uint32_t pos = s_defer[i];
        const bool lo = dot_ref(cf, X + (size_t)perm[base + pos] * ld, d) >= o;
        flags[base + pos] = lo;
        mylo += lo;
    }
    atomicAdd(&nlo, mylo);
    __syncthreads();
    if (tid == 0) tile_lo[blockIdx.x] = nlo;
}

size_t side_count16_lds(uint32_t ld16) { return (size_t)kSideRows * side16_units(ld16) * 16; }

// X16[i][j] = fp16(X[i][j]) (round to nearest even) for j < d, 0 for d <= j < ld16; one thread per
// 8 halves.
__global__ __launch_bounds__(256) void k_rows_to_half(const float* __restrict__ X, uint32_t ld, uint32_t d,
                                                      uint32_t n, uint16_t* __restrict__ X16, uint32_t ld16) {
    const uint32_t units = ld16 / 8;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (uint64_t)n * units) return;
    const uint32_t r = (uint32_t)(t / units), u = (uint32_t)(t - (uint64_t)r * units);
    const uint32_t d4 = (d + 3) / 4;
    const float4* xr = reinterpret_cast<const float4*>(X + (size_t)r * ld);
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 a = 2 * u < d4 ? xr[2 * u] : z, b = 2 * u + 1 < d4 ? xr[2 * u + 1] : z;
    const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t w[4];
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
        // dims past d inside the last float4 are the rows' zero padding
        const uint16_t lo16 = __builtin_bit_cast(uint16_t, (_Float16)(8 * u + 2 * e < d ? v[2 * e] : 0.f));
        const uint16_t hi16 = __builtin_bit_cast(uint16_t, (_Float16)(8 * u + 2 * e + 1 < d ? v[2 * e + 1] : 0.f));
        w[e] = (uint32_t)lo16 | ((uint32_t)hi16 << 16);
    }
    reinterpret_cast<uint4*>(X16 + (size_t)r * ld16)[u] = make_uint4(w[0], w[1], w[2], w[3]);
}

void launch_rows_to_half(const float* X, uint32_t ld, uint32_t d, uint32_t n, uint16_t* X16, uint32_t ld16,
                         hipStream_t st) {
    const uint64_t tot = (uint64_t)n * (ld16 / 8);
    if (tot) k_rows_to_half<<<(uint32_t)((tot + 255) / 256), 256, 0, st>>>(X, ld, d, n, X16, ld16);
}

// Exclusive scan of tile lo counts over all tiles of the level (one block; tiles <= ~1e6).
__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t* __restrict__ tile_lo,
                                                     uint32_t ntiles,
                                                     uint32_t* __restrict__ tile_base) {
    __shared__ uint32_t part[1024];
    const uint32_t per = (ntiles + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = min(ntiles, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += tile_lo[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t h = 1; h < 1024; h <<= 1) {
        uint32_t v = threadIdx.x >= h ? part[threadIdx.x - h] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t i = b; i < e; ++i) {
        tile_base[i] = run;
        run += tile_lo[i];
    }
    if (threadIdx.x == 1023) tile_base[ntiles] = part[1023];
}

__global__ void k_node_lo(NodeDesc* __restrict__ nodes, uint32_t nnodes,
                          const uint32_t* __restrict__ tile_base) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    NodeDesc& nd = nodes[i];
    if (nd.status == kNodeFailed) return;
    const uint32_t lo = tile_base[nd.tile0 + nd.ntiles] - tile_base[nd.tile0];
    nd.lo = lo;
    if ((lo == 0 || lo == nd.size) && nd.status == kNodeActive) nd.status = kNodeFailed;
}

// Stable scatter: lo positions keep their order at the front, hi positions after them.
__global__ __launch_bounds__(kSplitThreads) void k_scatter(
    const uint32_t* __restrict__ perm, uint32_t* __restrict__ perm_out,
    const NodeDesc* __restrict__ nodes, const TileDesc* __restrict__ tiles,
    const uint8_t* __restrict__ flags, const uint32_t* __restrict__ tile_base) {
    const TileDesc tl = tiles[blockIdx.x];
    const NodeDesc nd = nodes[tl.node];
    const uint32_t base = nd.start + tl.begin;
    constexpr uint32_t kPer = kTileSize / kSplitThreads;
    __shared__ uint32_t scan[kSplitThreads];
    const bool identity = nd.status == kNodeFailed || nd.lo == 0 || nd.lo == nd.size;
    uint8_t f[kPer];
    uint32_t cnt = 0;
    const uint32_t e0 = threadIdx.x * kPer;
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t e = e0 + q;
        f[q] = e < tl.count ? flags[base + e] : 0;
        cnt += f[q];
    }
    scan[threadIdx.x] = cnt;
    __syncthreads();
    for (uint32_t h = 1; h < kSplitThreads; h <<= 1) {
        uint32_t v = threadIdx.x >= h ? scan[threadIdx.x - h] : 0;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    const uint32_t lo_tile = tile_base[blockIdx.x] - tile_base[nd.tile0];
    uint32_t rank = lo_tile + scan[threadIdx.x] - cnt;  // lo entries before e0 within the node
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t e = e0 + q;
        if (e >= tl.count) break;
        const uint32_t pos = base + e;
        uint32_t dst;
        if (identity) {
            dst = pos;
        } else if (f[q]) {
            dst = nd.start + rank;
            ++rank;
        } else {
            dst = nd.start + nd.lo + (tl.begin + e - rank);
        }
        perm_out[dst] = perm[pos];
    }
}

__global__ __launch_bounds__(kSplitThreads) void k_copy_back(
    uint32_t* __restrict__ perm, const uint32_t* __restrict__ perm_tmp,
    const NodeDesc* __restrict__ nodes, const TileDesc* __restrict__ tiles) {
    const TileDesc tl = tiles[blockIdx.x];
    const uint32_t base = nodes[tl.node].start + tl.begin;
    for (uint32_t e = threadIdx.x; e < tl.count; e += blockDim.x) perm[base + e] = perm_tmp[base + e];
}

__global__ void k_iota(uint32_t* __restrict__ perm, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[i] = i;
}

// ---------------------------------------------------------------------------------------------
// Host launchers.
void launch_iota(uint32_t* perm, uint32_t n, hipStream_t st) {
    if (n) k_iota<<<(n + 255) / 256, 256, 0, st>>>(perm, n);
}

void launch_split_prepare(const SplitLevelArgs& a, hipStream_t st) {
    if (!a.run_lloyd) return;
    k_start_pair<<<a.nnodes, kSplitThreads, 0, st>>>(a.X, a.ld, a.d, a.perm, a.nodes, a.c1, a.c2,
                                                     a.coef, a.off);
    for (uint32_t it = 0; it < a.iterations; ++it) {
        if (a.nchunks)
            k_lloyd_chunk<<<a.nchunks, kSplitThreads, 0, st>>>(a.X, a.ld, a.d, a.perm, a.nodes,
                                                              a.chunks, a.coef, a.off, a.part,
                                                              a.pcnt);
        k_lloyd_finish<<<a.nnodes, kSplitThreads, 0, st>>>(a.nodes, a.part, a.pcnt, a.c1, a.c2,
                                                           a.coef, a.off, a.d, a.ld);
    }
}

void launch_split_side(const SplitLevelArgs& a, hipStream_t st) {
    if (a.X16) {
        static const bool attr16 = [] {
            return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_side_count16),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)side_count16_lds(kMaxDim)) == hipSuccess;
        }();
        (void)attr16;
        k_side_count16<<<a.ntiles, kSideThreads, side_count16_lds(a.ld16), st>>>(
            a.X, a.ld, a.X16, a.ld16, a.d, a.perm, a.nodes, a.tiles, a.coef, a.off, a.flags, a.tile_lo);
        return;
    }
    const size_t lds = side_count_lds(a.ld);
    static const bool attr_set = [] {
        return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_side_count),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)side_count_lds(kMaxDim)) == hipSuccess;
    }();
    (void)attr_set;
    k_side_count<<<a.ntiles, kSideThreads, lds, st>>>(a.X, a.ld, a.d, a.perm, a.nodes, a.tiles,
                                                      a.coef, a.off, a.flags, a.tile_lo);
}

void launch_split_partition(const SplitLevelArgs& a, hipStream_t st) {
    k_scan_tiles<<<1, 1024, 0, st>>>(a.tile_lo, a.ntiles, a.tile_base);
    k_node_lo<<<(a.nnodes + 255) / 256, 256, 0, st>>>(a.nodes, a.nnodes, a.tile_base);
    k_scatter<<<a.ntiles, kSplitThreads, 0, st>>>(a.perm, a.perm_tmp, a.nodes, a.tiles, a.flags,
                                                  a.tile_base);
    k_copy_back<<<a.ntiles, kSplitThreads, 0, st>>>(a.perm, a.perm_tmp, a.nodes, a.tiles);
}

}  // namespace knng
