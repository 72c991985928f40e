// Host driver and C ABI of the gfx950 KNN-graph builder (constructResult,
// SolutionKmeans.hpp:558-616, re-designed around device-resident state).
//
// A build keeps the points, the permutation, the split level tables and the per-point sorted
// top-k lists on the device for its whole life. Per round the host walks the split tree one level
// at a time: it uploads the level's node/chunk/tile tables, launches the level kernels, reads back
// each node's lo size and status, and derives the next level (children below leaf_max become
// leaves). Then one leaf launch covers every bucket of the round. Sample sizes are computed on the
// host with the C library's log/pow (SolutionKmeans.hpp:50-58, :70, :105), so they match a CPU
// run of the same formulas exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/knng.h"
#include "knng_device.hpp"
#include "knng_exact.hpp"
#include "knng_kernels.hpp"

namespace knng {
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define KNNG_HIP(expr)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? KNNG_ERR_NOMEM : KNNG_ERR_HIP,           \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

#define KNNG_TRY(expr)          \
    do {                        \
        int r_ = (expr);        \
        if (r_ != KNNG_OK) return r_; \
    } while (0)

// Device bytes held by the buffers a thread allocates, when the thread names a counter (a serving
// worker does, so that admission sees what its builder keeps between jobs). A buffer remembers the
// counter it was allocated under and gives its bytes back to it when freed.
thread_local std::atomic<uint64_t>* tl_device_account = nullptr;

// Grow-only device buffer.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    std::atomic<uint64_t>* acct = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(acct, o.acct);
    }
    void release() {
        if (p) (void)hipFree(p);
        if (acct) *acct -= cap * sizeof(T);
        p = nullptr;
        cap = 0;
        acct = nullptr;
    }
    int ensure(size_t count) {
        if (count <= cap && p) return KNNG_OK;
        release();
        size_t c = std::max<size_t>(count, 1);
        KNNG_HIP(hipMalloc(&p, c * sizeof(T)));
        cap = c;
        acct = tl_device_account;
        if (acct) *acct += cap * sizeof(T);
        return KNNG_OK;
    }
};

// Grow-only pinned host buffer.
template <class T>
struct PinBuf {
    T* p = nullptr;
    size_t cap = 0;
    PinBuf() = default;
    PinBuf(const PinBuf&) = delete;
    PinBuf& operator=(const PinBuf&) = delete;
    ~PinBuf() {
        if (p) (void)hipHostFree(p);
    }
    int ensure(size_t count) {
        if (count <= cap && p) return KNNG_OK;
        size_t c = std::max<size_t>({count, 2 * cap, 64});
        T* q = nullptr;
        KNNG_HIP(hipHostMalloc(reinterpret_cast<void**>(&q), c * sizeof(T)));
        if (p) (void)hipHostFree(p);
        p = q;
        cap = c;
        return KNNG_OK;
    }
};

// getSampleFromPercent size, SolutionKmeans.hpp:50-58: float percentage, float product, truncated.
uint32_t sample_count(uint32_t n) {
    float perc = (float)(std::pow(std::log((double)n) / std::log(30.0), 7) / (double)n);
    return (uint32_t)((float)n * perc);
}
// kmeansStartVecs sample size, SolutionKmeans.hpp:70, :105.
uint32_t start_count(uint32_t n) { return (uint32_t)std::pow(std::log10((double)n), 2.5); }

struct HostNode {
    uint32_t start, size;
    uint64_t key;
    uint32_t attempt;
};

bool valid_params(const knng_params& p, std::string& why) {
    if (p.k < 1 || p.k > kMaxK) { why = "k must be in [1, 128]"; return false; }
    if (p.search_k > kMaxK) { why = "search_k must be <= 128"; return false; }
    if (p.non_rank_lo && (p.non_rank_lo >= std::max(p.k, p.search_k) || std::max(p.k, p.search_k) > 100)) {
        why = "non_rank_lo must be below the list width, which must be <= 100 in the ordered form";
        return false;
    }
    if (p.leaf_max < 1 || p.leaf_max > kMaxLeafAny) { why = "leaf_max must be in [1, 65536]"; return false; }
    if (p.leaf_max_odd > kMaxLeafAny) { why = "leaf_max_odd must be <= 65536"; return false; }
    if (p.small_max > kMaxSmallNode) { why = "small_max must be <= 4096"; return false; }
    if (p.precision > KNNG_PREC_BF16) { why = "unknown precision"; return false; }
    if ((p.flags & KNNG_FLAG_LEAF_FP32_GRAM) && p.precision != KNNG_PREC_FP32) {
        why = "KNNG_FLAG_LEAF_FP32_GRAM is a form of the exact precision (KNNG_PREC_FP32)";
        return false;
    }
    return true;
}

}  // namespace

// ---- random-rotation KD splitter: host parts (SolutionRandomKD.hpp:258-428) ----------------
// The reference-order dot (LinearAlgebra.hpp:179-200) on the host: 8 lane accumulators with fused
// multiply-add over dims below floor(d/8)*8, lanes added in order into 0.0f, then the tail fused.
static float host_dot_ref(const float* a, const float* b, uint32_t d) {
    const uint32_t main = d & ~7u;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint32_t i = 0; i < main; i += 8)
        for (uint32_t j = 0; j < 8; ++j) acc[j] = std::fmaf(a[i + j], b[i + j], acc[j]);
    float s = 0.0f;
    for (uint32_t j = 0; j < 8; ++j) s = s + acc[j];
    for (uint32_t i = main; i < d; ++i) s = std::fmaf(a[i], b[i], s);
    return s;
}

// One N(0, 1) draw: Box-Muller on two counter-based draws of the stream `key` (in place of
// std::normal_distribution over the thread-local mt19937, LinearAlgebra.hpp:104).
static double kd_normal(uint64_t key, uint64_t i) {
    const uint64_t h1 = knng::mix64(key + 0x9E3779B97F4A7C15ull * (2 * i + 1));
    const uint64_t h2 = knng::mix64(key + 0x9E3779B97F4A7C15ull * (2 * i + 2));
    const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);  // (0, 1]
    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// rotateRandom's m directions for a round (SolutionRandomKD.hpp:258-264), as the reference
// computes them: randUniformUnitVec (LinearAlgebra.hpp:103-116: float draws, divided by the
// double square root of their float sum of squares, :79-92), then gramSchmidt (:211-222) whose
// normalize (:94-101) returns its argument unchanged: u_0 = v_0 and u_i = v_i - sum_j
// dot(u_j, v_i) u_j, not renormalized. Rows of d floats.
static std::vector<float> kd_directions(uint64_t seed, uint32_t round, uint32_t m, uint32_t d) {
    const uint64_t key = knng::mix64(knng::round_key(seed, round) ^ 0x6A09E667F3BCC909ull);
    std::vector<float> v((size_t)m * d), u((size_t)m * d);
    for (uint32_t c = 0; c < m; ++c) {
        float* vc = &v[(size_t)c * d];
        float ss = 0.0f;
        for (uint32_t j = 0; j < d; ++j) {
            vc[j] = (float)kd_normal(key, (uint64_t)c * d + j);
            const float sq = vc[j] * vc[j];
            ss = ss + sq;
        }
        const double nrm = std::sqrt((double)ss);
        for (uint32_t j = 0; j < d; ++j) vc[j] = (float)((double)vc[j] / nrm);
    }
    for (uint32_t i = 0; i < m; ++i) {
        float* ui = &u[(size_t)i * d];
        const float* vi = &v[(size_t)i * d];
        std::memcpy(ui, vi, d * sizeof(float));
        for (uint32_t j = 0; j < i; ++j) {
            const float* uj = &u[(size_t)j * d];
            const float c = host_dot_ref(uj, vi, d);
            for (uint32_t e = 0; e < d; ++e) {
                const float p = c * uj[e];
                ui[e] = ui[e] - p;
            }
        }
    }
    return u;
}

// pickIndexFull's choice (SolutionRandomKD.hpp:223-248): the dimensions ordered by variance,
// descending (ties: the larger index first, std::greater on (variance, index)); a float draw in
// [0, totalVar) with totalVar the float sum of the variances in index order; the first dimension
// whose float running sum reaches the draw (the last one if rounding leaves the sum short, where
// the reference reaches __builtin_unreachable).
static uint32_t kd_pick(const double* var, uint32_t m, uint64_t node_key) {
    float total = 0.0f;
    for (uint32_t j = 0; j < m; ++j) total = (float)((double)total + var[j]);
    std::vector<std::pair<double, uint32_t>> order(m);
    for (uint32_t j = 0; j < m; ++j) order[j] = {var[j], j};
    std::sort(order.begin(), order.end(), std::greater<std::pair<double, uint32_t>>());
    const float u = (float)(knng::mix64(node_key + 0x9E3779B97F4A7C15ull) >> 40) * (1.0f / 16777216.0f);
    const float sample = u * total;
    float so_far = 0.0f;
    for (auto& [vj, j] : order) {
        so_far = (float)((double)so_far + vj);
        if (so_far >= sample) return j;
    }
    return order.back().second;
}

struct Builder {
    knng_params prm{};
    hipStream_t st = nullptr;
    // A round's split runs on its own stream, so that it overlaps the previous round's leaf stage
    // (the split reads only the points; the host drives it level by level).
    hipStream_t sst = nullptr;
    hipEvent_t ev[8] = {};
    knng_stats stats{};

    // points (internal copy when the caller's layout is not usable directly)
    DevBuf<float> xbuf;
    DevBuf<uint16_t> x16;  // fp16 copy of the bound rows for the side test (split_ld16(d) halves each)
    bool x16_ready = false;
    const float* X = nullptr;
    uint32_t n = 0, d = 0, ld = 0;

    // split state
    DevBuf<uint32_t> perm, perm_next, perm_tmp, tile_lo, tile_base, pcnt;  // perm_next: the other round's
    uint32_t* split_perm = nullptr;  // the permutation the running split writes
    DevBuf<uint8_t> flags;
    DevBuf<NodeDesc> nodes;
    DevBuf<ChunkDesc> chunks;
    DevBuf<TileDesc> tiles;
    DevBuf<float> c1, c2, coef, off;
    DevBuf<double> part;
    PinBuf<NodeDesc> h_nodes;
    PinBuf<ChunkDesc> h_chunks;
    PinBuf<TileDesc> h_tiles;
    // KD splitter: projected rows, directions, per-node statistics and choices
    DevBuf<float> kd_rot, kd_dirs;
    DevBuf<double> kd_part, kd_mean, kd_var, kd_split;
    DevBuf<uint32_t> kd_dim;
    PinBuf<double> h_kd_stat;   // per node: kKdStride means then kKdStride variances
    PinBuf<double> h_kd_split;
    PinBuf<uint32_t> h_kd_dim;
    // leaf descriptors, one set per round parity; up_ev[s]: set s's upload has been read
    PinBuf<LeafDesc> h_leaves[2];
    PinBuf<BigBlock> h_big[2];
    hipEvent_t up_ev[2] = {};

    // top-k state
    DevBuf<uint32_t> knn_ids, snap;
    DevBuf<float> knn_d, bound;
    DevBuf<LeafDesc> leaves[2];
    DevBuf<BigBlock> bigblocks[2];
    // batched leaf stage: work-queue headers per round parity, and the captured launches
    DevBuf<LeafQueue> lq[2];
    PinBuf<LeafQueue> h_lq[2];
    hipGraphExec_t leaf_graph[2] = {};
    LeafArgs leaf_graph_args[2] = {};
    const BigBlock* leaf_graph_big[2] = {};
    uint64_t leaf_graph_replays = 0;
    // Leaf-stage configuration: the launch form and, in the exact form, the screen's 16-bit type
    // (fp16 or bf16 Gram, each with its own rigorous margin: the same lists). Set by the flags, or
    // by the autotuner on a build's first leaf stage (KNNG_FLAG_AUTOTUNE).
    bool leaf_batched = false;
    uint32_t leaf_half = 0;
    bool leaf_timing = false;  // a tuner pass: no statistics, no per-bin shape tuning
    DevBuf<uint64_t> counters;

    // exact k-NN scratch
    DevBuf<double> norms64, cand_d, ubuf;
    DevBuf<uint32_t> cand_id, qcount, ovf;

    // neighbours-of-neighbours scratch
    // candidate slots per point before the direct path (41 GB at 10M points per set); the points
    // that overflow them go through the direct kernel
    static constexpr uint32_t kNonCap = 512;
    static constexpr uint32_t kNonCapSmall = 16;       // KNNG_FLAG_SMALL_NON_CAP (tests)
    static constexpr uint32_t kNonSubpasses = 5;       // first pass of a build
    static constexpr uint32_t kNonSubpassesLater = 1;  // later passes: lists are close, buffers small
    uint32_t non_cap() const { return (prm.flags & KNNG_FLAG_SMALL_NON_CAP) ? kNonCapSmall : kNonCap; }
    DevBuf<uint32_t> ovf_flag;
    DevBuf<uint32_t> rcnt, rfill, rsl, ncnt, nbuf_id;
    DevBuf<uint32_t> snap_prev, rnew, rfill_old;
    DevBuf<unsigned char> rs_ws;  // the reverse-list sort's buffers when the candidate set is too small
    bool have_prev = false;  // snap_prev holds the previous pass's snapshot of the current lists
    uint32_t pass_no = 0;    // neighbours-of-neighbours passes so far in this build
    DevBuf<uint64_t> roff, scan_tmp;
    DevBuf<float> nbuf_d, norms32;
    DevBuf<uint32_t> norm_rng;  // smallest nonzero and largest squared norm (float bits)
    // the second candidate-buffer set: sub-passes alternate, so a sub-pass's merge (on sst) runs
    // beside the next sub-pass's screen
    DevBuf<uint32_t> ncnt2, nbuf_id2;
    DevBuf<float> nbuf_d2;
    bool non_two_sets = true;  // this pass uses the second set

    // generator scratch
    DevBuf<float> gen_a, gen_b, gen_c;

    // pinned staging for host <-> device copies: two buffers, so that the CPU fills (or drains)
    // one while the DMA engine moves the other
    static constexpr size_t kStageBytes = 32u << 20;
    PinBuf<unsigned char> stage[2];
    hipEvent_t stage_ev[2] = {};

    ~Builder() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : stage_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : up_ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
        if (sst) (void)hipStreamDestroy(sst);
        for (auto& g : leaf_graph)
            if (g) (void)hipGraphExecDestroy(g);
    }

    bool verbose = false;

    int init(const knng_params& p) {
        verbose = std::getenv("KNNG_VERBOSE") != nullptr;
        std::string why;
        if (!valid_params(p, why)) return fail(KNNG_ERR_INVALID, why);
        prm = p;
        KNNG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        KNNG_HIP(hipStreamCreateWithFlags(&sst, hipStreamNonBlocking));
        for (auto& e : ev) KNNG_HIP(hipEventCreate(&e));
        for (auto& e : stage_ev) KNNG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto& e : up_ev) KNNG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return KNNG_OK;
    }

    // Runs fn(lo, hi) over [0, count) on up to 4 host threads.
    template <class F>
    static void parallel_rows(size_t count, F fn) {
        const size_t nt = count >= 4096 ? 4 : 1;
        if (nt == 1) {
            fn((size_t)0, count);
            return;
        }
        std::vector<std::thread> th;
        for (size_t t = 1; t < nt; ++t)
            th.emplace_back(fn, count * t / nt, count * (t + 1) / nt);
        fn((size_t)0, count / nt);
        for (auto& x : th) x.join();
    }

    // Host rows (any stride) -> device rows of ld floats, zero-padded, through pinned staging.
    int upload_rows(float* dst, const float* src, size_t stride, uint32_t rows) {
        const size_t rb = (size_t)ld * sizeof(float);
        const size_t per = std::max<size_t>(1, kStageBytes / rb);
        for (int b = 0; b < 2; ++b) KNNG_TRY(stage[b].ensure(per * rb));
        bool used[2] = {false, false};
        for (size_t i0 = 0, blk = 0; i0 < rows; i0 += per, ++blk) {
            const int b = (int)(blk & 1);
            if (used[b]) KNNG_HIP(hipEventSynchronize(stage_ev[b]));
            const size_t m = std::min<size_t>(per, rows - i0);
            float* buf = reinterpret_cast<float*>(stage[b].p);
            const uint32_t dd = d, ldd = ld;
            parallel_rows(m, [=](size_t lo, size_t hi) {
                for (size_t r = lo; r < hi; ++r) {
                    std::memcpy(buf + r * ldd, src + (i0 + r) * stride, dd * sizeof(float));
                    for (uint32_t j = dd; j < ldd; ++j) buf[r * ldd + j] = 0.f;
                }
            });
            KNNG_HIP(hipMemcpyAsync(dst + i0 * ld, buf, m * rb, hipMemcpyHostToDevice, st));
            KNNG_HIP(hipEventRecord(stage_ev[b], st));
            used[b] = true;
        }
        KNNG_HIP(hipStreamSynchronize(st));
        return KNNG_OK;
    }

    // Device bytes -> pageable host memory through pinned staging.
    int download(void* dst, const void* src, size_t bytes) {
        for (int b = 0; b < 2; ++b) KNNG_TRY(stage[b].ensure(kStageBytes));
        const unsigned char* s8 = static_cast<const unsigned char*>(src);
        unsigned char* d8 = static_cast<unsigned char*>(dst);
        size_t pend_off[2] = {0, 0}, pend_len[2] = {0, 0};
        auto drain = [&](int b) -> int {
            if (!pend_len[b]) return KNNG_OK;
            KNNG_HIP(hipEventSynchronize(stage_ev[b]));
            const unsigned char* from = stage[b].p;
            unsigned char* to = d8 + pend_off[b];
            parallel_rows(pend_len[b] / 4096 + 1, [=](size_t lo, size_t hi) {
                const size_t a0 = std::min(lo * 4096, pend_len[b]), a1 = std::min(hi * 4096, pend_len[b]);
                if (a1 > a0) std::memcpy(to + a0, from + a0, a1 - a0);
            });
            pend_len[b] = 0;
            return KNNG_OK;
        };
        for (size_t o = 0, blk = 0; o < bytes; o += kStageBytes, ++blk) {
            const int b = (int)(blk & 1);
            KNNG_TRY(drain(b));
            const size_t m = std::min(kStageBytes, bytes - o);
            KNNG_HIP(hipMemcpyAsync(stage[b].p, s8 + o, m, hipMemcpyDeviceToHost, st));
            KNNG_HIP(hipEventRecord(stage_ev[b], st));
            pend_off[b] = o;
            pend_len[b] = m;
        }
        KNNG_TRY(drain(0));
        KNNG_TRY(drain(1));
        return KNNG_OK;
    }

    // list width while building; the output keeps each list's first prm.k
    uint32_t kl() const { return std::max(prm.k, prm.search_k); }

    // LeafArgs::rank_cap: in the exact screen the list width (the k-th smallest upper bound of a
    // bucket loses none of the k nearest), unless KNNG_FLAG_NO_BUCKET_LIMIT; in the approximate
    // screens the re-rank cap max(rerank, width), up to 160 (0, or a larger cap: no cap).
    uint32_t leaf_rank_cap() const {
        if (prm.precision == KNNG_PREC_FP32) return (prm.flags & KNNG_FLAG_NO_BUCKET_LIMIT) ? 0u : kl();
        const uint32_t m = prm.rerank ? std::max(prm.rerank, kl()) : 0u;
        return m <= 160 ? m : 0u;
    }

    uint32_t leaf_max_for(uint32_t round) const {
        return (round % 2 == 1 && prm.leaf_max_odd) ? prm.leaf_max_odd : prm.leaf_max;
    }

    // Make the points usable by the kernels: 16-byte aligned rows with a stride of 4k floats whose
    // first ceil(d/4) float4 hold nothing beyond dim d-1 but zeros. A device buffer is used in place
    // when it qualifies (d a multiple of 4: the caller's padding is never loaded); otherwise the
    // rows are copied into zero-padded rows of (d + 3) & ~3 floats.
    int bind_points(const float* pts, bool on_device, uint32_t nn, uint32_t dd, size_t stride) {
        if (dd < 1 || dd > kMaxDim) return fail(KNNG_ERR_UNSUPPORTED, "d must be in [1, 256]");
        if (stride < dd) return fail(KNNG_ERR_INVALID, "row_stride < d");
        if (nn && !pts) return fail(KNNG_ERR_INVALID, "points is null");
        n = nn;
        d = dd;
        x16_ready = false;
        const bool usable = on_device && dd % 4 == 0 && stride % 4 == 0 &&
                            (reinterpret_cast<uintptr_t>(pts) & 15) == 0;
        if (usable) {
            X = pts;
            ld = (uint32_t)stride;
            return KNNG_OK;
        }
        ld = (dd + 3) & ~3u;
        KNNG_TRY(xbuf.ensure((size_t)nn * ld));
        if (nn && on_device) {
            KNNG_HIP(hipMemsetAsync(xbuf.p, 0, (size_t)nn * ld * sizeof(float), st));
            KNNG_HIP(hipMemcpy2DAsync(xbuf.p, ld * sizeof(float), pts, stride * sizeof(float),
                                      dd * sizeof(float), nn, hipMemcpyDeviceToDevice, st));
        } else if (nn) {
            KNNG_TRY(upload_rows(xbuf.p, pts, stride, nn));
        }
        X = xbuf.p;
        return KNNG_OK;
    }

    // The fp16 rows of the side test, converted once per binding of the points (on stream s).
    int ensure_x16(hipStream_t s) {
        if (x16_ready || (prm.flags & KNNG_FLAG_SIDE_FP32)) return KNNG_OK;
        const uint32_t ld16 = split_ld16(d);
        KNNG_TRY(x16.ensure((size_t)n * ld16));
        launch_rows_to_half(X, ld, d, n, x16.p, ld16, s);
        KNNG_HIP(hipGetLastError());
        x16_ready = true;
        return KNNG_OK;
    }

    SplitLevelArgs level_args(uint32_t nn, uint32_t nc, uint32_t nt, bool lloyd) {
        SplitLevelArgs a{};
        a.X = X;
        a.ld = ld;
        a.d = d;
        const bool half = x16_ready && !(prm.flags & KNNG_FLAG_SIDE_FP32);
        a.X16 = half ? x16.p : nullptr;
        a.ld16 = half ? split_ld16(d) : 0u;
        a.perm = split_perm;
        a.perm_tmp = perm_tmp.p;
        a.nodes = nodes.p;
        a.nnodes = nn;
        a.chunks = chunks.p;
        a.nchunks = nc;
        a.tiles = tiles.p;
        a.ntiles = nt;
        a.c1 = c1.p;
        a.c2 = c2.p;
        a.coef = coef.p;
        a.off = off.p;
        a.part = part.p;
        a.pcnt = pcnt.p;
        a.flags = flags.p;
        a.tile_lo = tile_lo.p;
        a.tile_base = tile_base.p;
        a.iterations = prm.iterations;
        a.run_lloyd = lloyd;
        return a;
    }

    // Upload the tables of `batch` nodes (already in h_nodes), run the level, read back.
    int run_level(uint32_t nb, uint32_t nc, uint32_t nt, bool lloyd, hipStream_t s) {
        KNNG_TRY(nodes.ensure(nb));
        KNNG_TRY(chunks.ensure(nc));
        KNNG_TRY(tiles.ensure(nt));
        KNNG_TRY(c1.ensure((size_t)nb * ld));
        KNNG_TRY(c2.ensure((size_t)nb * ld));
        KNNG_TRY(coef.ensure((size_t)nb * ld));
        KNNG_TRY(off.ensure(nb));
        KNNG_TRY(part.ensure((size_t)std::max(nc, 1u) * 2 * ld));
        KNNG_TRY(pcnt.ensure((size_t)std::max(nc, 1u) * 2));
        KNNG_TRY(tile_lo.ensure(nt));
        KNNG_TRY(tile_base.ensure((size_t)nt + 1));
        KNNG_HIP(hipMemcpyAsync(nodes.p, h_nodes.p, nb * sizeof(NodeDesc), hipMemcpyHostToDevice, s));
        if (nc) KNNG_HIP(hipMemcpyAsync(chunks.p, h_chunks.p, nc * sizeof(ChunkDesc), hipMemcpyHostToDevice, s));
        KNNG_HIP(hipMemcpyAsync(tiles.p, h_tiles.p, nt * sizeof(TileDesc), hipMemcpyHostToDevice, s));
        SplitLevelArgs a = level_args(nb, nc, nt, lloyd);
        launch_split_prepare(a, s);
        KNNG_HIP(hipEventRecord(ev[2], s));
        launch_split_side(a, s);
        KNNG_HIP(hipEventRecord(ev[3], s));
        launch_split_partition(a, s);
        KNNG_HIP(hipGetLastError());
        KNNG_HIP(hipMemcpyAsync(h_nodes.p, nodes.p, nb * sizeof(NodeDesc), hipMemcpyDeviceToHost, s));
        KNNG_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        KNNG_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
        stats.side_ms += ms;
        stats.side_launches++;
        return KNNG_OK;
    }

    // Fill h_nodes/h_chunks/h_tiles for the given pending nodes.
    int plan(const std::vector<HostNode>& lvl, const std::vector<uint32_t>& pending, uint32_t& nc,
             uint32_t& nt) {
        const uint32_t nb = (uint32_t)pending.size();
        KNNG_TRY(h_nodes.ensure(nb));
        nc = 0;
        nt = 0;
        for (uint32_t b = 0; b < nb; ++b) {
            const HostNode& hn = lvl[pending[b]];
            const bool large = hn.size >= prm.small_max;
            const uint32_t m = large ? sample_count(hn.size) : 0;
            uint32_t s = start_count(hn.size);
            if (large) s = std::min(m, s);
            s = std::max(s, 2u);
            s = std::min(s, 1024u);
            const uint32_t nch = large ? (m + kLargeChunk - 1) / kLargeChunk : 1;
            const uint32_t ntl = (hn.size + kTileSize - 1) / kTileSize;
            NodeDesc& nd = h_nodes.p[b];
            nd.start = hn.start;
            nd.size = hn.size;
            nd.akey = attempt_key(hn.key, hn.attempt);
            nd.m = m;
            nd.s = s;
            nd.chunk0 = nc;
            nd.nchunks = nch;
            nd.tile0 = nt;
            nd.ntiles = ntl;
            nd.status = kNodeActive;
            nd.lo = 0;
            nc += nch;
            nt += ntl;
        }
        KNNG_TRY(h_chunks.ensure(nc));
        KNNG_TRY(h_tiles.ensure(nt));
        for (uint32_t b = 0; b < nb; ++b) {
            const NodeDesc& nd = h_nodes.p[b];
            for (uint32_t c = 0; c < nd.nchunks; ++c) {
                ChunkDesc& ch = h_chunks.p[nd.chunk0 + c];
                ch.node = b;
                if (nd.m) {
                    ch.begin = c * kLargeChunk;
                    ch.count = std::min(kLargeChunk, nd.m - ch.begin);
                } else {
                    ch.begin = 0;
                    ch.count = nd.size;
                }
            }
            for (uint32_t t = 0; t < nd.ntiles; ++t) {
                TileDesc& tl = h_tiles.p[nd.tile0 + t];
                tl.node = b;
                tl.begin = t * kTileSize;
                tl.count = std::min(kTileSize, nd.size - tl.begin);
            }
        }
        return KNNG_OK;
    }

    // One round of splitKmeansBinaryProcess from the identity permutation, into `dst`, on stream
    // `s` (the host waits for each level on it).
    int split_round(uint32_t round, std::vector<LeafDesc>& out_leaves, hipStream_t s,
                    DevBuf<uint32_t>& dst) {
        KNNG_TRY(dst.ensure(n));
        KNNG_TRY(perm_tmp.ensure(n));
        KNNG_TRY(flags.ensure(n));
        split_perm = dst.p;
        launch_iota(dst.p, n, s);
        perm_is_tree = true;
        if (prm.flags & KNNG_FLAG_KD_SPLIT) return kd_split_round(round, out_leaves, s, dst);
        out_leaves.clear();
        const uint32_t leaf_max = leaf_max_for(round);
        if (n <= leaf_max) {
            if (n) out_leaves.push_back({0, n});
            return KNNG_OK;
        }
        KNNG_TRY(ensure_x16(s));
        std::vector<HostNode> level{{0, n, round_key(prm.seed, round), 0}}, next;
        std::vector<uint32_t> lo, pending, again;
        while (!level.empty()) {
            lo.assign(level.size(), 0);
            pending.resize(level.size());
            for (uint32_t i = 0; i < level.size(); ++i) pending[i] = i;
            while (!pending.empty()) {
                uint32_t nc = 0, nt = 0;
                KNNG_TRY(plan(level, pending, nc, nt));
                KNNG_TRY(run_level((uint32_t)pending.size(), nc, nt, true, s));
                again.clear();
                for (uint32_t b = 0; b < pending.size(); ++b) {
                    HostNode& hn = level[pending[b]];
                    const NodeDesc& nd = h_nodes.p[b];
                    stats.side_rows += hn.size;
                    if (nd.status == kNodeActive) {
                        lo[pending[b]] = nd.lo;
                    } else {
                        stats.split_failures++;
                        if (++hn.attempt >= kMaxAttempts) {  // defined fallback: halve by position
                            stats.split_fallbacks++;
                            lo[pending[b]] = hn.size / 2;
                        } else {
                            again.push_back(pending[b]);
                        }
                    }
                }
                pending.swap(again);
            }
            next.clear();
            for (uint32_t i = 0; i < level.size(); ++i) {
                const HostNode& hn = level[i];
                HostNode ch[2] = {{hn.start, lo[i], child_key(hn.key, 0), 0},
                                  {hn.start + lo[i], hn.size - lo[i], child_key(hn.key, 1), 0}};
                for (auto& c : ch) {
                    if (c.size <= leaf_max) out_leaves.push_back({c.start, c.size});
                    else next.push_back(c);
                }
            }
            level.swap(next);
        }
        std::sort(out_leaves.begin(), out_leaves.end(),
                  [](const LeafDesc& a, const LeafDesc& b) { return a.start < b.start; });
        return KNNG_OK;
    }

    // One round of the random-rotation KD splitter (SolutionRandomKD.hpp:258-428) into `dst` (already
    // the identity): project onto m = min(30, d) directions, then split level by level: each node
    // above leaf_max takes the FP64 means and sample variances of its projections, a dimension
    // drawn by variance, and the mean as split value (x < mean first, stable). A node whose split
    // leaves a side empty (the reference recurses forever) is cut at its middle position.
    int kd_split_round(uint32_t round, std::vector<LeafDesc>& out_leaves, hipStream_t s,
                       DevBuf<uint32_t>& dst) {
        const uint32_t leaf_max = leaf_max_for(round);
        out_leaves.clear();
        if (n <= leaf_max) {
            if (n) out_leaves.push_back({0, n});
            return KNNG_OK;
        }
        const uint32_t m = std::min(kKdDims, d);
        if (kd_rotate_lds(ld) > 64 * 1024) return fail(KNNG_ERR_UNSUPPORTED, "KD splitter supports d <= 256");
        const std::vector<float> u = kd_directions(prm.seed, round, m, d);
        std::vector<float> R((size_t)m * ld, 0.f);
        for (uint32_t c = 0; c < m; ++c) std::memcpy(&R[(size_t)c * ld], &u[(size_t)c * d], d * sizeof(float));
        KNNG_TRY(kd_dirs.ensure(R.size()));
        KNNG_TRY(kd_rot.ensure((size_t)n * kKdStride));
        KNNG_HIP(hipMemcpyAsync(kd_dirs.p, R.data(), R.size() * sizeof(float), hipMemcpyHostToDevice, s));
        launch_kd_rotate(X, ld, d, kd_dirs.p, m, n, kd_rot.p, s);
        KNNG_HIP(hipStreamSynchronize(s));  // R (pageable) is released when this returns
        std::vector<HostNode> level{{0, n, round_key(prm.seed, round), 0}}, next;
        while (!level.empty()) {
            const uint32_t nb = (uint32_t)level.size();
            uint32_t nt = 0;
            KNNG_TRY(h_nodes.ensure(nb));
            for (uint32_t b = 0; b < nb; ++b) {
                NodeDesc& nd = h_nodes.p[b];
                nd = NodeDesc{};
                nd.start = level[b].start;
                nd.size = level[b].size;
                nd.tile0 = nt;
                nd.ntiles = (nd.size + kTileSize - 1) / kTileSize;
                nd.status = kNodeActive;
                nt += nd.ntiles;
            }
            KNNG_TRY(h_tiles.ensure(nt));
            for (uint32_t b = 0; b < nb; ++b) {
                const NodeDesc& nd = h_nodes.p[b];
                for (uint32_t t = 0; t < nd.ntiles; ++t)
                    h_tiles.p[nd.tile0 + t] = TileDesc{b, t * kTileSize, std::min(kTileSize, nd.size - t * kTileSize)};
            }
            KNNG_TRY(nodes.ensure(nb));
            KNNG_TRY(tiles.ensure(nt));
            KNNG_TRY(tile_lo.ensure(nt));
            KNNG_TRY(tile_base.ensure((size_t)nt + 1));
            KNNG_TRY(kd_part.ensure((size_t)nt * kKdStride));
            KNNG_TRY(kd_mean.ensure((size_t)nb * kKdStride));
            KNNG_TRY(kd_var.ensure((size_t)nb * kKdStride));
            KNNG_TRY(kd_split.ensure(nb));
            KNNG_TRY(kd_dim.ensure(nb));
            KNNG_TRY(h_kd_stat.ensure((size_t)nb * 2 * kKdStride));
            KNNG_TRY(h_kd_split.ensure(nb));
            KNNG_TRY(h_kd_dim.ensure(nb));
            KNNG_HIP(hipMemcpyAsync(nodes.p, h_nodes.p, nb * sizeof(NodeDesc), hipMemcpyHostToDevice, s));
            KNNG_HIP(hipMemcpyAsync(tiles.p, h_tiles.p, nt * sizeof(TileDesc), hipMemcpyHostToDevice, s));
            launch_kd_stats(kd_rot.p, dst.p, nodes.p, nb, tiles.p, nt, m, kd_part.p, kd_mean.p, kd_var.p, s);
            KNNG_HIP(hipMemcpyAsync(h_kd_stat.p, kd_mean.p, (size_t)nb * kKdStride * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
            KNNG_HIP(hipMemcpyAsync(h_kd_stat.p + (size_t)nb * kKdStride, kd_var.p,
                                    (size_t)nb * kKdStride * sizeof(double), hipMemcpyDeviceToHost, s));
            KNNG_HIP(hipStreamSynchronize(s));
            for (uint32_t b = 0; b < nb; ++b) {
                const double* mean = h_kd_stat.p + (size_t)b * kKdStride;
                const double* var = h_kd_stat.p + ((size_t)nb + b) * kKdStride;
                const uint32_t j = kd_pick(var, m, level[b].key);
                h_kd_dim.p[b] = j;
                h_kd_split.p[b] = mean[j];
            }
            KNNG_HIP(hipMemcpyAsync(kd_dim.p, h_kd_dim.p, nb * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            KNNG_HIP(hipMemcpyAsync(kd_split.p, h_kd_split.p, nb * sizeof(double), hipMemcpyHostToDevice, s));
            launch_kd_side(kd_rot.p, dst.p, nodes.p, tiles.p, nt, kd_dim.p, kd_split.p, flags.p, tile_lo.p, s);
            SplitLevelArgs a = level_args(nb, 0, nt, false);
            launch_split_partition(a, s);
            KNNG_HIP(hipGetLastError());
            KNNG_HIP(hipMemcpyAsync(h_nodes.p, nodes.p, nb * sizeof(NodeDesc), hipMemcpyDeviceToHost, s));
            KNNG_HIP(hipStreamSynchronize(s));
            next.clear();
            for (uint32_t b = 0; b < nb; ++b) {
                const HostNode& hn = level[b];
                const NodeDesc& nd = h_nodes.p[b];
                stats.side_rows += hn.size;
                uint32_t lo = nd.lo;
                if (nd.status != kNodeActive) {  // a side was empty: the order is unchanged
                    stats.split_failures++;
                    stats.split_fallbacks++;
                    lo = hn.size / 2;
                }
                HostNode ch[2] = {{hn.start, lo, child_key(hn.key, 0), 0},
                                  {hn.start + lo, hn.size - lo, child_key(hn.key, 1), 0}};
                for (auto& c : ch) {
                    if (c.size <= leaf_max) out_leaves.push_back({c.start, c.size});
                    else next.push_back(c);
                }
            }
            level.swap(next);
        }
        std::sort(out_leaves.begin(), out_leaves.end(),
                  [](const LeafDesc& a, const LeafDesc& b) { return a.start < b.start; });
        return KNNG_OK;
    }

    int ensure_state() {
        KNNG_TRY(knn_ids.ensure((size_t)n * kl()));
        KNNG_TRY(knn_d.ensure((size_t)n * kl()));
        KNNG_TRY(bound.ensure(n));
        KNNG_TRY(counters.ensure(4));
        return KNNG_OK;
    }

    // Leaves are binned by size: each bin of the LDS kernel gets LDS sized for its own largest
    // leaf (small leaves then fit several workgroups per CU), and leaves beyond it (or all of them
    // with KNNG_FLAG_BIG_LEAF) go to the big-bucket kernel as 64-row blocks.
    // Leaf stage of one round on `st`, reading the round's permutation `pm` and its descriptors
    // through set `slot`. With sync, the host waits and the stage's time is added to the stats;
    // without, the caller times it from the events e0/e1 recorded around its kernels.
    // A leaf stage's launches: the leaves of each LDS size bin (in bin order) and the row blocks
    // of the leaves that go to the big-bucket kernel.
    struct LeafPlan {
        std::vector<LeafDesc> bin[4];
        std::vector<BigBlock> big;
        uint64_t pairs = 0, nleaves = 0;
        size_t nl = 0;  // leaves over all bins
    };

    int plan_leaves(const std::vector<LeafDesc>& lv, LeafPlan& pl) const {
        if (d > 128) return fail(KNNG_ERR_UNSUPPORTED, "leaf stage supports d <= 128");
        static constexpr uint32_t kEdges[4] = {64, 128, 256, kMaxLeaf};
        const bool force_big = (prm.flags & (KNNG_FLAG_BIG_LEAF | KNNG_FLAG_LEAF_FP32_GRAM)) != 0;
        pl.nleaves = lv.size();
        for (auto& l : lv) {
            pl.pairs += (uint64_t)l.size * (l.size - 1) / 2;
            if (l.size < 2) continue;
            int bi = 0;
            while (bi < 4 && l.size > kEdges[bi]) ++bi;
            // a bin's LDS is sized by its largest leaf, so the test is on the leaf's own size
            if (force_big || (bi < 4 && leaf_lds_bytes(l.size, d, ld) > 160 * 1024)) bi = 4;
            if (bi < 4) {
                pl.bin[bi].push_back(l);
            } else {
                for (uint32_t r0 = 0; r0 < l.size; r0 += kLeafBigRows) pl.big.push_back({l.start, l.size, r0});
            }
        }
        if (!pl.big.empty() && leaf_big_lds_bytes(d, ld) > 160 * 1024)
            return fail(KNNG_ERR_UNSUPPORTED, "big-bucket leaf kernel does not fit LDS at this d");
        for (auto& v : pl.bin) pl.nl += v.size();
        return KNNG_OK;
    }

    // Launches a planned leaf stage on stream st. The plan's bins sit back to back at d_leaves and
    // its big-bucket blocks at d_big; ids/dists/bnd are the lists and their bounds; cnt the
    // statistics counters. Events e0/e1 bracket the launches (timed into stats when sync).
    int launch_leaf_plan(const LeafPlan& pl, const LeafDesc* d_leaves, const BigBlock* d_big,
                         const uint32_t* pm, uint32_t* ids, float* dists, float* bnd, uint64_t* cnt,
                         bool tune, bool sync, hipEvent_t e0, hipEvent_t e1) {
        LeafArgs a{};
        a.X = X;
        a.ld = ld;
        a.d = d;
        a.k = kl();
        a.perm = pm;
        a.knn_ids = ids;
        a.knn_d = dists;
        a.bound = bnd;
        a.screen = prm.precision == KNNG_PREC_FP32 ? kScreenExact : kScreenApprox;
        a.half_kind = screen_half();
        a.counters = leaf_timing ? nullptr : cnt;
        a.tune = tune && !leaf_timing ? 1u : 0u;
        a.rank_cap = leaf_rank_cap();
        KNNG_HIP(hipEventRecord(e0, st));
        size_t off = 0;
        for (auto& v : pl.bin) {
            if (!v.empty()) {
                uint32_t mx = 0;
                for (auto& l : v) mx = std::max(mx, l.size);
                a.leaves = d_leaves + off;
                a.nleaves = (uint32_t)v.size();
                a.max_leaf = mx;
                launch_leaf(a, st);
                stats.leaf_launches++;
            }
            off += v.size();
        }
        if (!pl.big.empty()) {
            a.leaves = nullptr;
            a.nleaves = 0;
            launch_leaf_big(a, d_big, (uint32_t)pl.big.size(), st);
            stats.leaf_launches++;
        }
        KNNG_HIP(hipGetLastError());
        KNNG_HIP(hipEventRecord(e1, st));
        if (sync) {
            KNNG_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            KNNG_HIP(hipEventElapsedTime(&ms, e0, e1));
            stats.leaf_ms += ms;
        }
        stats.leaves += pl.nleaves;
        stats.leaf_pairs += pl.pairs;
        return KNNG_OK;
    }

    // Batched leaf stage: one persistent kernel over one work list (big-bucket blocks, then the
    // LDS-resident buckets by decreasing size), optionally replayed from a captured graph.
    int leaf_stage_batched(const std::vector<LeafDesc>& lv, const uint32_t* pm, uint32_t slot,
                           bool sync, hipEvent_t e0, hipEvent_t e1) {
        if (d > 128) return fail(KNNG_ERR_UNSUPPORTED, "leaf stage supports d <= 128");
        const uint32_t lim = leaf_batched_limit(d, ld);
        if (leaf_big_lds_bytes(d, ld) > 160 * 1024 || leaf_batched_lds(d, ld) > 160 * 1024)
            return fail(KNNG_ERR_UNSUPPORTED, "batched leaf kernel does not fit LDS at this d");
        std::vector<LeafDesc> small;
        std::vector<BigBlock> big;
        uint64_t pairs = 0;
        small.reserve(lv.size());
        for (auto& l : lv) {
            pairs += (uint64_t)l.size * (l.size - 1) / 2;
            if (l.size < 2) continue;
            if (l.size <= lim && !(prm.flags & KNNG_FLAG_BIG_LEAF)) small.push_back(l);
            else for (uint32_t r0 = 0; r0 < l.size; r0 += kLeafBigRows) big.push_back({l.start, l.size, r0});
        }
        std::stable_sort(small.begin(), small.end(),
                         [](const LeafDesc& x, const LeafDesc& y) { return x.size > y.size; });
        KNNG_HIP(hipEventSynchronize(up_ev[slot]));
        KNNG_TRY(h_leaves[slot].ensure(std::max<size_t>(small.size(), 1)));
        KNNG_TRY(leaves[slot].ensure(std::max<size_t>(small.size(), 1)));
        KNNG_TRY(h_big[slot].ensure(std::max<size_t>(big.size(), 1)));
        KNNG_TRY(bigblocks[slot].ensure(std::max<size_t>(big.size(), 1)));
        KNNG_TRY(h_lq[slot].ensure(1));
        KNNG_TRY(lq[slot].ensure(1));
        if (!small.empty()) {
            std::memcpy(h_leaves[slot].p, small.data(), small.size() * sizeof(LeafDesc));
            KNNG_HIP(hipMemcpyAsync(leaves[slot].p, h_leaves[slot].p, small.size() * sizeof(LeafDesc),
                                    hipMemcpyHostToDevice, st));
        }
        if (!big.empty()) {
            std::memcpy(h_big[slot].p, big.data(), big.size() * sizeof(BigBlock));
            KNNG_HIP(hipMemcpyAsync(bigblocks[slot].p, h_big[slot].p, big.size() * sizeof(BigBlock),
                                    hipMemcpyHostToDevice, st));
        }
        *h_lq[slot].p = LeafQueue{0u, (uint32_t)big.size(), (uint32_t)small.size(), 0u};
        KNNG_HIP(hipMemcpyAsync(lq[slot].p, h_lq[slot].p, sizeof(LeafQueue), hipMemcpyHostToDevice, st));
        KNNG_HIP(hipEventRecord(up_ev[slot], st));
        LeafArgs a{};
        a.X = X;
        a.ld = ld;
        a.d = d;
        a.k = kl();
        a.perm = pm;
        a.knn_ids = knn_ids.p;
        a.knn_d = knn_d.p;
        a.bound = bound.p;
        a.screen = prm.precision == KNNG_PREC_FP32 ? kScreenExact : kScreenApprox;
        a.half_kind = screen_half();
        a.counters = leaf_timing ? nullptr : counters.p;
        a.tune = 0;
        a.rank_cap = leaf_rank_cap();
        a.leaves = leaves[slot].p;
        a.nleaves = (uint32_t)small.size();
        a.max_leaf = lim;
        KNNG_HIP(hipEventRecord(e0, st));
        if (prm.flags & KNNG_FLAG_LEAF_GRAPH) {
            // the launches depend on the round only through the uploaded lists and counts, so a
            // graph captured once per round parity is replayed while its arguments are unchanged
            a.nleaves = 0;  // the kernel reads the counts from the queue header
            const bool same = leaf_graph[slot] && leaf_graph_big[slot] == bigblocks[slot].p &&
                              std::memcmp(&leaf_graph_args[slot], &a, sizeof(LeafArgs)) == 0;
            if (!same) {
                if (leaf_graph[slot]) KNNG_HIP(hipGraphExecDestroy(leaf_graph[slot]));
                leaf_graph[slot] = nullptr;
                hipGraph_t g = nullptr;
                KNNG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                launch_leaf_batched(a, bigblocks[slot].p, lq[slot].p, st);
                KNNG_HIP(hipStreamEndCapture(st, &g));
                const hipError_t ie = hipGraphInstantiate(&leaf_graph[slot], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                KNNG_HIP(ie);
                leaf_graph_args[slot] = a;
                leaf_graph_big[slot] = bigblocks[slot].p;
            }
            KNNG_HIP(hipGraphLaunch(leaf_graph[slot], st));
            ++leaf_graph_replays;
        } else {
            launch_leaf_batched(a, bigblocks[slot].p, lq[slot].p, st);
        }
        KNNG_HIP(hipGetLastError());
        KNNG_HIP(hipEventRecord(e1, st));
        stats.leaf_launches += 2;
        if (sync) {
            KNNG_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            KNNG_HIP(hipEventElapsedTime(&ms, e0, e1));
            stats.leaf_ms += ms;
        }
        stats.leaves += lv.size();
        stats.leaf_pairs += pairs;
        return KNNG_OK;
    }

    // The 16-bit type of the leaf screen: the precision's in the approximate forms; in the exact
    // form fp16 unless the tuner kept bf16.
    uint32_t screen_half() const {
        if (prm.flags & KNNG_FLAG_LEAF_FP32_GRAM) return knng::kHalfF32;
        if (prm.precision == KNNG_PREC_BF16) return 1;
        if (prm.precision == KNNG_PREC_FP16) return 0;
        return leaf_half;
    }

    int leaf_stage(const std::vector<LeafDesc>& lv, const uint32_t* pm, uint32_t slot, bool sync,
                   hipEvent_t e0, hipEvent_t e1) {
        if (lv.empty()) {
            KNNG_HIP(hipEventRecord(e0, st));
            KNNG_HIP(hipEventRecord(e1, st));
            return KNNG_OK;
        }
        if (prm.flags & KNNG_FLAG_LEAF_FP32_GRAM)  // one form: every bucket on the big-bucket kernel
            return leaf_stage_binned(lv, pm, slot, sync, e0, e1);
        if (prm.flags & KNNG_FLAG_AUTOTUNE) {
            bool ran = false;
            KNNG_TRY(tune_leaf_stage(lv, pm, slot, sync, e0, e1, ran));
            if (ran) return KNNG_OK;
        }
        if (leaf_batched || (prm.flags & (KNNG_FLAG_LEAF_BATCHED | KNNG_FLAG_LEAF_GRAPH)))
            return leaf_stage_batched(lv, pm, slot, sync, e0, e1);
        return leaf_stage_binned(lv, pm, slot, sync, e0, e1);
    }

    // Autotuner over the leaf stage's forms: {one launch per size bin, one persistent kernel over
    // the work list} x {fp16, bf16 screen} (the screen type only in the exact form, where both
    // give the same lists). On the first leaf stage of a (device, d, k, precision, leaf sizes) key
    // in the process, the stage runs once for real in the default form, then once more in each
    // candidate form, timed: a repeated pass offers pairs that were already offered, so the lists
    // stay as they are, and every candidate sees the same lists. The fastest form is kept for the
    // process and reported on stderr.
    // ran: the stage itself was run here (the real pass, then the timed ones).
    int tune_leaf_stage(const std::vector<LeafDesc>& lv, const uint32_t* pm, uint32_t slot,
                        bool sync, hipEvent_t e0, hipEvent_t e1, bool& ran) {
        ran = false;
        int dev = 0;
        KNNG_HIP(hipGetDevice(&dev));
        const uint64_t key = ((uint64_t)(uint32_t)dev << 56) ^ ((uint64_t)d << 44) ^ ((uint64_t)kl() << 36) ^
                             ((uint64_t)prm.precision << 32) ^ ((uint64_t)prm.leaf_max << 16) ^
                             (uint64_t)prm.leaf_max_odd;
        {
            std::lock_guard<std::mutex> g(leaf_tune_mu());
            const auto it = leaf_tune_map().find(key);
            if (it != leaf_tune_map().end()) {
                leaf_batched = it->second.first;
                leaf_half = it->second.second;
                return KNNG_OK;
            }
        }
        // the real pass, in the default form
        leaf_batched = false;
        leaf_half = 0;
        KNNG_TRY(leaf_stage_binned(lv, pm, slot, sync, e0, e1));
        ran = true;
        const bool exact = prm.precision == KNNG_PREC_FP32;
        struct Cand { bool batched; uint32_t half; float ms; };
        std::vector<Cand> cands = {{false, 0, 0.f}, {true, 0, 0.f}};
        if (exact) {
            cands.push_back({false, 1, 0.f});
            cands.push_back({true, 1, 0.f});
        }
        const knng_stats saved = stats;
        leaf_timing = true;
        int rc = KNNG_OK;
        for (auto& c : cands) {
            leaf_batched = c.batched;
            leaf_half = c.half;
            rc = c.batched ? leaf_stage_batched(lv, pm, slot, true, ev[6], ev[7])
                           : leaf_stage_binned(lv, pm, slot, true, ev[6], ev[7]);
            if (rc != KNNG_OK) break;
            KNNG_HIP(hipEventElapsedTime(&c.ms, ev[6], ev[7]));
        }
        leaf_timing = false;
        stats = saved;
        if (rc != KNNG_OK) return rc;
        const Cand* best = &cands[0];
        for (auto& c : cands)
            if (c.ms < best->ms) best = &c;
        leaf_batched = best->batched;
        leaf_half = best->half;
        std::string line;
        for (auto& c : cands) {
            char buf[96];
            std::snprintf(buf, sizeof buf, "%s%s %s %.2f ms", line.empty() ? "" : ", ",
                          c.batched ? "batched" : "binned", c.half ? "bf16" : "fp16", c.ms);
            line += buf;
        }
        std::fprintf(stderr, "[knng] leaf stage tuned (d %u, k %u, leaves <= %u/%u): %s; kept %s %s\n",
                     d, kl(), prm.leaf_max, prm.leaf_max_odd ? prm.leaf_max_odd : prm.leaf_max,
                     line.c_str(), leaf_batched ? "batched" : "binned", leaf_half ? "bf16" : "fp16");
        std::lock_guard<std::mutex> g(leaf_tune_mu());
        leaf_tune_map().emplace(key, std::make_pair(leaf_batched, leaf_half));
        return KNNG_OK;
    }
    static std::mutex& leaf_tune_mu() {
        static std::mutex m;
        return m;
    }
    static std::unordered_map<uint64_t, std::pair<bool, uint32_t>>& leaf_tune_map() {
        static std::unordered_map<uint64_t, std::pair<bool, uint32_t>> m;
        return m;
    }

    int leaf_stage_binned(const std::vector<LeafDesc>& lv, const uint32_t* pm, uint32_t slot,
                          bool sync, hipEvent_t e0, hipEvent_t e1) {
        LeafPlan pl;
        KNNG_TRY(plan_leaves(lv, pl));
        // the pinned set's previous upload (two rounds ago) must have been read before it is refilled
        KNNG_HIP(hipEventSynchronize(up_ev[slot]));
        KNNG_TRY(h_leaves[slot].ensure(std::max<size_t>(pl.nl, 1)));
        KNNG_TRY(leaves[slot].ensure(std::max<size_t>(pl.nl, 1)));
        size_t off = 0;
        for (auto& v : pl.bin) {
            if (!v.empty()) std::memcpy(h_leaves[slot].p + off, v.data(), v.size() * sizeof(LeafDesc));
            off += v.size();
        }
        if (pl.nl)
            KNNG_HIP(hipMemcpyAsync(leaves[slot].p, h_leaves[slot].p, pl.nl * sizeof(LeafDesc),
                                    hipMemcpyHostToDevice, st));
        if (!pl.big.empty()) {
            KNNG_TRY(h_big[slot].ensure(pl.big.size()));
            std::memcpy(h_big[slot].p, pl.big.data(), pl.big.size() * sizeof(BigBlock));
            KNNG_TRY(bigblocks[slot].ensure(pl.big.size()));
            KNNG_HIP(hipMemcpyAsync(bigblocks[slot].p, h_big[slot].p, pl.big.size() * sizeof(BigBlock),
                                    hipMemcpyHostToDevice, st));
        }
        KNNG_HIP(hipEventRecord(up_ev[slot], st));
        return launch_leaf_plan(pl, leaves[slot].p, bigblocks[slot].p, pm, knn_ids.p, knn_d.p,
                                bound.p, counters.p, (prm.flags & KNNG_FLAG_AUTOTUNE) != 0, sync, e0,
                                e1);
    }

    // Permutation of the last split round; processing order of the reverse NoN kernel.
    bool perm_is_tree = false;

    int non_pass() {
        const uint32_t k = kl();
        const size_t nk = (size_t)n * k;
        if (n > kIdMask) return fail(KNNG_ERR_UNSUPPORTED, "neighbours of neighbours needs n < 2^30");
        KNNG_TRY(snap.ensure(nk));
        // later passes of a build screen only offers with a new hop (not in the ordered form: an
        // old entry can cross the rank threshold and be expanded for the first time)
        const bool incremental = have_prev && (prm.flags & KNNG_FLAG_INCREMENTAL_NON) && prm.non_rank_lo == 0;
        // 3 (default): fp16 Gram of the scaled rows with the exact margin, in every precision
        // form; 0: FP32 Gram; 1/2: 16-bit Gram of rows centred on q (opt-in, slower)
        const uint32_t screen16 = !non_screen16_supported(d) ? 0u
                                  : (prm.flags & KNNG_FLAG_NON_SCREEN16) ? (prm.precision == KNNG_PREC_BF16 ? 2u : 1u)
                                  : (prm.flags & KNNG_FLAG_NON_FP32_GRAM) ? 0u : 3u;
        const bool reverse = ld <= 128 && non_reverse_lds(k, ld, d, screen16) <= 160 * 1024 &&
                             !(prm.flags & KNNG_FLAG_DIRECT_NON);
        // Full passes: the snapshot kernel writes an (id, p) pair per entry that goes to a reverse
        // list, and a stable radix sort by id turns them into the lists (in p order); incremental
        // passes count and fill the lists with atomics, new entries first.
        const bool sorted_rs = reverse && !incremental;
        uint32_t *ka = nullptr, *kb = nullptr, *va = nullptr;
        void* stmp = nullptr;
        size_t stb = 0;
        const uint32_t sbits = 32u - (uint32_t)__builtin_clz(n | 1u);  // 2^bits - 1 >= n
        if (reverse) {
            KNNG_TRY(rcnt.ensure(n));
            KNNG_TRY(rnew.ensure(n));
            KNNG_TRY(rsl.ensure(nk));
            KNNG_TRY(nbuf_d.ensure((size_t)n * non_cap()));
            KNNG_TRY(nbuf_id.ensure((size_t)n * non_cap()));
        }
        if (sorted_rs) {
            // The sort's buffers lie in the first candidate set, idle until the pass's first
            // screen (every merge of the previous pass is done): keys a and b in nbuf_d, values a
            // and the temporary storage in nbuf_id. Buffers too small for them (the tests' 16-slot
            // candidate buffers) give way to a workspace of their own.
            stb = rsort_temp_bytes(nk, sbits);
            if (stb == 0) return fail(KNNG_ERR_HIP, "radix sort: no temporary storage size");
            const size_t cap_b = (size_t)n * non_cap() * 4;
            if (2 * nk * 4 <= cap_b && nk * 4 + stb + 256 <= cap_b) {
                ka = reinterpret_cast<uint32_t*>(nbuf_d.p);
                kb = ka + nk;
                va = nbuf_id.p;
            } else {
                KNNG_TRY(rs_ws.ensure(3 * nk * 4 + stb + 256));
                ka = reinterpret_cast<uint32_t*>(rs_ws.p);
                kb = ka + nk;
                va = kb + nk;
            }
            stmp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(va + nk) + 255) & ~(uintptr_t)255);
        }
        launch_non_snapshot(knn_ids.p, snap.p, n, k, incremental ? snap_prev.p : nullptr,
                            prm.non_rank_lo, reverse && incremental ? rcnt.p : nullptr,
                            reverse && incremental ? rnew.p : nullptr, ka, va, st);
        NonArgs da{};
        da.X = X;
        da.ld = ld;
        da.d = d;
        da.k = k;
        da.n = n;
        da.snap = snap.p;
        da.knn_ids = knn_ids.p;
        da.knn_d = knn_d.p;
        da.bound = bound.p;
        if (!reverse) {
            launch_non_direct(da, nullptr, n, st);
            KNNG_HIP(hipGetLastError());
            have_prev = false;
            return KNNG_OK;
        }
        KNNG_TRY(rfill.ensure(n));
        KNNG_TRY(rfill_old.ensure(n));
        KNNG_TRY(roff.ensure((size_t)n + 1));
        KNNG_TRY(scan_tmp.ensure(scan_scratch_elems(n)));
        KNNG_TRY(ncnt.ensure(n));
        // The second candidate set (the merge of one sub-pass beside the next screen; 13 ms per
        // 10M build) is taken only while it leaves half the free memory: at 10M points its 41 GB
        // fit, and a build near 40M points on one GPU still runs, on one set with each merge
        // before the next screen. KNNG_FLAG_NON_ONE_SET (tests) asks for one set.
        {
            const size_t set_bytes = (size_t)n * non_cap() * 8 + (size_t)n * 4;
            size_t free_b = 0, total_b = 0;
            KNNG_HIP(hipMemGetInfo(&free_b, &total_b));
            const bool have = nbuf_id2.cap >= (size_t)n * non_cap() && ncnt2.cap >= n;
            non_two_sets = !(prm.flags & KNNG_FLAG_NON_ONE_SET) && (have || set_bytes <= free_b / 2);
        }
        if (non_two_sets) {
            KNNG_TRY(ncnt2.ensure(n));
            KNNG_TRY(nbuf_d2.ensure((size_t)n * non_cap()));
            KNNG_TRY(nbuf_id2.ensure((size_t)n * non_cap()));
        }
        KNNG_TRY(ovf.ensure((size_t)n + 1));
        KNNG_TRY(ovf_flag.ensure(n));
        KNNG_TRY(perm.ensure(n));
        if (!perm_is_tree) launch_iota(perm.p, n, st);
        if (sorted_rs) {
            bool in_b = false;
            KNNG_HIP(rsort_pairs(stmp, stb, ka, kb, va, rsl.p, nk, sbits, st, &in_b));
            if (!in_b) KNNG_HIP(hipMemcpyAsync(rsl.p, va, nk * 4, hipMemcpyDeviceToDevice, st));
            launch_roff_search(in_b ? kb : ka, nk, n, roff.p, st);
        } else {
            launch_reverse_lists(snap.p, n, k, rcnt.p, rnew.p, roff.p, scan_tmp.p, rfill.p, rfill_old.p,
                                 rsl.p, st);
        }
        if (verbose) {
            std::vector<uint64_t> ro((size_t)n + 1);
            KNNG_HIP(hipMemcpyAsync(ro.data(), roff.p, ((size_t)n + 1) * 8, hipMemcpyDeviceToHost, st));
            KNNG_HIP(hipStreamSynchronize(st));
            std::vector<uint32_t> rc(n);
            for (uint32_t q = 0; q < n; ++q) rc[q] = (uint32_t)(ro[q + 1] - ro[q]);
            std::sort(rc.begin(), rc.end());
            uint64_t tot = 0, big = 0;
            for (uint32_t v : rc) { tot += v; if (v > 1024) big += v; }
            std::fprintf(stderr, "[knng] reverse lists: total %llu, p50 %u p99 %u p99.9 %u max %u, "
                                 "entries in lists > 1024: %llu\n",
                         (unsigned long long)tot, rc[n / 2], rc[(size_t)n * 99 / 100],
                         rc[(size_t)n * 999 / 1000], rc[n - 1], (unsigned long long)big);
        }
        KNNG_HIP(hipMemsetAsync(ovf.p, 0, ((size_t)n + 1) * 4, st));
        KNNG_HIP(hipMemsetAsync(ovf_flag.p, 0, (size_t)n * 4, st));
        NonRevArgs a{};
        a.X = X;
        a.ld = ld;
        a.d = d;
        a.k = k;
        a.n = n;
        a.snap = snap.p;
        a.roff = roff.p;
        a.rs = rsl.p;
        a.rnew = incremental ? rnew.p : nullptr;
        a.knn_ids = knn_ids.p;
        a.knn_d = knn_d.p;
        a.bound = bound.p;
        a.ncnt = ncnt.p;
        a.nbuf_d = nbuf_d.p;
        a.nbuf_id = nbuf_id.p;
        a.cap = non_cap();
        a.half_kind = prm.precision == KNNG_PREC_BF16 ? 1 : 0;
        a.screen16 = screen16;
        a.exact = prm.precision == KNNG_PREC_FP32 || screen16 == 3 ? 1u : 0u;
        a.counters = verbose ? counters.p : nullptr;  // same-address atomics: diagnostics only
        KNNG_TRY(norms32.ensure(n));
        launch_row_norms(X, ld, d, n, norms32.p, st);
        a.norms = norms32.p;
        // The fp16 screen may convert rows without their power-of-two scales when every nonzero
        // squared norm lies in [2^-12, 2^28]: no value then exceeds fp16's range, and the absolute
        // rounding of values in fp16's subnormal range stays inside the margin's slack.
        a.unscaled = 0;
        if (screen16 == 3) {
            KNNG_TRY(norm_rng.ensure(2));
            launch_norm_range(norms32.p, n, norm_rng.p, st);
            uint32_t rng[2] = {0u, 0u};
            KNNG_HIP(hipMemcpyAsync(rng, norm_rng.p, sizeof rng, hipMemcpyDeviceToHost, st));
            KNNG_HIP(hipStreamSynchronize(st));
            float lo = 0.f, hi = 0.f;
            std::memcpy(&lo, &rng[0], 4);
            std::memcpy(&hi, &rng[1], 4);
            a.unscaled = lo >= 0x1p-12f && hi <= 0x1p28f ? 1u : 0u;  // all-zero rows: lo = +inf
        }
        // Sub-passes over interleaved parts of the tree order (every subs-th q, so a point's q are
        // spread over all of them), each merged before the next: the bounds tighten as the pass
        // goes, which keeps the candidate buffers small.
        // a sub-pass merges every point's list once, so fewer of them save list traffic; the first
        // pass needs more, to keep the candidate buffers within their capacity (at 10M points,
        // with each merge beside the next screen: 5 + 1 sub-passes build in 1952 ms, 6 + 1 in
        // 1955, 6 + 2 in 1964, 8 + 2 in 1968, 6 + 3 in 1978, means of 2-5 builds)
        const uint32_t subs = std::max(1u, std::min(pass_no ? kNonSubpassesLater : kNonSubpasses, n / 1024 + 1));
        // Sub-pass sp screens into candidate set sp % 2 on st; its merge runs on sst once the screen
        // is done, beside the next screen, which may read a bound before or after the merge lowers
        // it (a larger bound only admits candidates the merge rejects). The next screen into the
        // same set waits for that set's merge; the pass ends when the last merge has.
        float t_rev = 0.f, t_merge = 0.f, ms = 0.f;
        std::vector<hipEvent_t> spev(4 * subs, nullptr);
        struct EvGuard {
            std::vector<hipEvent_t>& v;
            ~EvGuard() {
                for (auto e : v)
                    if (e) (void)hipEventDestroy(e);
            }
        } spguard{spev};
        for (auto& e : spev) KNNG_HIP(hipEventCreate(&e));
        for (uint32_t sp = 0; sp < subs; ++sp) {
            a.order = perm.p + sp;
            a.order_stride = subs;
            a.nq = (n - sp + subs - 1) / subs;
            const bool second = non_two_sets && (sp & 1);
            a.ncnt = second ? ncnt2.p : ncnt.p;
            a.nbuf_d = second ? nbuf_d2.p : nbuf_d.p;
            a.nbuf_id = second ? nbuf_id2.p : nbuf_id.p;
            // the screen that reuses a set waits for that set's merge
            const uint32_t nsets = non_two_sets ? 2u : 1u;
            if (sp >= nsets) KNNG_HIP(hipStreamWaitEvent(st, spev[4 * (sp - nsets) + 3], 0));
            KNNG_HIP(hipMemsetAsync(a.ncnt, 0, (size_t)n * 4, st));
            KNNG_HIP(hipEventRecord(spev[4 * sp], st));
            launch_non_reverse(a, st);
            KNNG_HIP(hipEventRecord(spev[4 * sp + 1], st));
            KNNG_HIP(hipStreamWaitEvent(sst, spev[4 * sp + 1], 0));
            KNNG_HIP(hipEventRecord(spev[4 * sp + 2], sst));
            launch_non_merge(a, ovf.p, ovf.p + 1, ovf_flag.p, sst);
            KNNG_HIP(hipEventRecord(spev[4 * sp + 3], sst));
            KNNG_HIP(hipGetLastError());
        }
        KNNG_HIP(hipStreamWaitEvent(st, spev[4 * (subs - 1) + 3], 0));
        KNNG_HIP(hipEventSynchronize(spev[4 * (subs - 1) + 3]));
        for (uint32_t sp = 0; sp < subs; ++sp) {
            KNNG_HIP(hipEventElapsedTime(&ms, spev[4 * sp], spev[4 * sp + 1]));
            t_rev += ms;
            KNNG_HIP(hipEventElapsedTime(&ms, spev[4 * sp + 2], spev[4 * sp + 3]));
            t_merge += ms;
        }
        uint32_t novf = 0;
        uint64_t entries = 0;
        KNNG_HIP(hipMemcpyAsync(&novf, ovf.p, 4, hipMemcpyDeviceToHost, st));
        KNNG_HIP(hipMemcpyAsync(&entries, roff.p + n, 8, hipMemcpyDeviceToHost, st));
        KNNG_HIP(hipStreamSynchronize(st));
        stats.non_overflow += novf;
        stats.non_screen_ms += t_rev;
        stats.non_screen_launches += subs;
        // per q: its k snapshot rows; per reverse entry: the row and the id snapshot of p
        stats.non_screen_bytes += (uint64_t)n * k * 4 * d + entries * (4ull * d + 4ull * k);
        KNNG_HIP(hipEventRecord(ev[2], st));
        if (novf) launch_non_direct(da, ovf.p + 1, novf, st);
        KNNG_HIP(hipEventRecord(ev[3], st));
        KNNG_HIP(hipGetLastError());
        if (verbose) {
            KNNG_HIP(hipEventSynchronize(ev[3]));
            KNNG_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
            uint64_t nnew = entries;
            if (incremental) {
                std::vector<uint32_t> rn(n);
                KNNG_HIP(hipMemcpy(rn.data(), rnew.p, (size_t)n * 4, hipMemcpyDeviceToHost));
                nnew = 0;
                for (uint32_t v : rn) nnew += v;
            }
            std::fprintf(stderr, "[knng] non pass (%s): %u sub-passes, reverse %.1f ms merge %.1f ms, "
                                 "%u overflowed points direct %.1f ms, new entries %.3f\n",
                         incremental ? "incremental" : "full", subs, t_rev, t_merge, novf, ms,
                         entries ? (double)nnew / (double)entries : 0.0);
        }
        if (prm.flags & KNNG_FLAG_INCREMENTAL_NON) {  // keep this snapshot for the next pass
            snap.swap(snap_prev);
            have_prev = true;
        }
        ++pass_no;
        return KNNG_OK;
    }

    // One scan + select over `nq` queries with `cap` candidate slots each; qcount keeps the counts.
    int exact_pass(const uint32_t* d_queries, uint32_t nq, uint32_t k, const double* ub,
                   uint32_t cap, uint32_t* out_ids, double* out_d) {
        KNNG_TRY(cand_d.ensure((size_t)nq * cap));
        KNNG_TRY(cand_id.ensure((size_t)nq * cap));
        KNNG_TRY(qcount.ensure(nq));
        KNNG_TRY(ovf.ensure(1));
        KNNG_HIP(hipMemsetAsync(qcount.p, 0, (size_t)nq * 4, st));
        KNNG_HIP(hipMemsetAsync(ovf.p, 0, 4, st));
        ExactScanArgs a{};
        a.X = X;
        a.ld = ld;
        a.d = d;
        a.n = n;
        a.norms = norms64.p;
        a.queries = d_queries;
        a.nq = nq;
        a.ub = ub;
        a.cap = cap;
        a.count = qcount.p;
        a.cand_d = cand_d.p;
        a.cand_id = cand_id.p;
        const uint32_t qblocks = (nq + 63) / 64;
        const uint32_t slices =
            std::max(1u, std::min<uint32_t>(2048u / std::max(qblocks, 1u) + 1, (n + 4095) / 4096));
        launch_exact_scan(a, slices, st);
        launch_exact_select(a, k, out_ids, out_d, ovf.p, st);
        KNNG_HIP(hipGetLastError());
        KNNG_HIP(hipStreamSynchronize(st));
        return KNNG_OK;
    }

    int build(uint32_t* out_ids, float* out_d) {
        stats = knng_stats{};
        perm_is_tree = false;
        KNNG_TRY(ensure_state());
        KNNG_HIP(hipMemsetAsync(counters.p, 0, 4 * sizeof(uint64_t), st));
        KNNG_HIP(hipEventRecord(ev[0], st));
        launch_init_knn(knn_ids.p, knn_d.p, bound.p, n, kl(), st);
        // Rounds are pipelined over two streams: split r+1 (sst, host-driven level by level) runs
        // while leaf stage r (st) does. Leaf r waits for split r; split r+1 writes the permutation
        // buffer that leaf r-1 read, so it waits for leaf r-1. Events per round: split start/end
        // (sst), leaf start/end (st).
        const uint32_t R = prm.rounds;
        std::vector<hipEvent_t> rev(4 * (size_t)R, nullptr);
        struct EvGuard {
            std::vector<hipEvent_t>& v;
            ~EvGuard() {
                for (auto e : v)
                    if (e) (void)hipEventDestroy(e);
            }
        } guard{rev};
        for (auto& e : rev) KNNG_HIP(hipEventCreate(&e));
        DevBuf<uint32_t>* pb[2] = {&perm, &perm_next};
        std::vector<LeafDesc> lv[2];
        KNNG_HIP(hipEventRecord(ev[2], st));  // the points are bound (uploads on st) before any split
        KNNG_HIP(hipStreamWaitEvent(sst, ev[2], 0));
        auto split = [&](uint32_t r) -> int {
            if (r >= 2) KNNG_HIP(hipStreamWaitEvent(sst, rev[4 * (r - 2) + 3], 0));
            KNNG_HIP(hipEventRecord(rev[4 * r], sst));
            KNNG_TRY(split_round(r, lv[r & 1], sst, *pb[r & 1]));
            KNNG_HIP(hipEventRecord(rev[4 * r + 1], sst));
            return KNNG_OK;
        };
        if (R) KNNG_TRY(split(0));
        for (uint32_t r = 0; r < R; ++r) {
            KNNG_HIP(hipStreamWaitEvent(st, rev[4 * r + 1], 0));
            KNNG_TRY(leaf_stage(lv[r & 1], pb[r & 1]->p, r & 1, false, rev[4 * r + 2], rev[4 * r + 3]));
            if (r + 1 < R) KNNG_TRY(split(r + 1));
            if (verbose) {
                uint64_t c[4];
                KNNG_HIP(hipStreamSynchronize(st));
                KNNG_HIP(hipMemcpy(c, counters.p, sizeof(c), hipMemcpyDeviceToHost));
                std::fprintf(stderr, "[knng] round %u: leaves %zu survivors %llu inserts %llu\n", r,
                             lv[r & 1].size(), (unsigned long long)c[0], (unsigned long long)c[2]);
            }
        }
        KNNG_HIP(hipStreamSynchronize(st));
        KNNG_HIP(hipStreamSynchronize(sst));
        for (uint32_t r = 0; r < R; ++r) {
            float ms = 0.f;
            KNNG_HIP(hipEventElapsedTime(&ms, rev[4 * r], rev[4 * r + 1]));
            stats.split_ms += ms;
            KNNG_HIP(hipEventElapsedTime(&ms, rev[4 * r + 2], rev[4 * r + 3]));
            stats.leaf_ms += ms;
        }
        // neighbours of neighbours take the last round's tree order from perm
        if (R && ((R - 1) & 1)) perm.swap(perm_next);
        perm_is_tree = R > 0;
        uint64_t leafc[4];
        KNNG_HIP(hipMemcpy(leafc, counters.p, sizeof(leafc), hipMemcpyDeviceToHost));
        stats.survivors = leafc[0];
        stats.inserts = leafc[2];
        KNNG_HIP(hipMemsetAsync(counters.p, 0, 4 * sizeof(uint64_t), st));
        KNNG_HIP(hipEventRecord(ev[6], st));
        have_prev = false;
        pass_no = 0;
        for (uint32_t p = 0; p < prm.non_passes; ++p) KNNG_TRY(non_pass());
        KNNG_HIP(hipEventRecord(ev[7], st));
        if (verbose && prm.non_passes) {
            uint64_t c[4];
            KNNG_HIP(hipMemcpy(c, counters.p, sizeof(c), hipMemcpyDeviceToHost));
            std::fprintf(stderr, "[knng] non: screened %llu appended %llu overflow %llu\n",
                         (unsigned long long)c[0], (unsigned long long)c[1],
                         (unsigned long long)stats.non_overflow);
        }
        launch_finalize(knn_ids.p, knn_d.p, n, kl(), prm.k, out_ids, out_d, st);
        KNNG_HIP(hipGetLastError());
        KNNG_HIP(hipEventRecord(ev[1], st));
        KNNG_HIP(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        KNNG_HIP(hipEventElapsedTime(&ms, ev[6], ev[7]));
        stats.non_ms = ms;
        KNNG_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        stats.total_ms = ms;
        return KNNG_OK;
    }
};

}  // namespace knng

using knng::Builder;
using knng::fail;

struct knng_handle {
    Builder b;
};

namespace {

int check_handle(const knng_handle* h) {
    return h ? KNNG_OK : fail(KNNG_ERR_INVALID, "null handle");
}

// Host state (n * list width) <-> device state.
int upload_state(Builder& b, const uint32_t* ids, const float* dists, bool empty) {
    const size_t tot = (size_t)b.n * b.kl();
    KNNG_TRY(b.ensure_state());
    if (empty) {
        knng::launch_init_knn(b.knn_ids.p, b.knn_d.p, b.bound.p, b.n, b.kl(), b.st);
        return KNNG_OK;
    }
    KNNG_HIP(hipMemcpyAsync(b.knn_ids.p, ids, tot * 4, hipMemcpyHostToDevice, b.st));
    KNNG_HIP(hipMemcpyAsync(b.knn_d.p, dists, tot * 4, hipMemcpyHostToDevice, b.st));
    // bound = distance in slot k-1
    std::vector<float> bnd(b.n);
    for (uint32_t p = 0; p < b.n; ++p) bnd[p] = dists[(size_t)p * b.kl() + b.kl() - 1];
    KNNG_HIP(hipMemcpyAsync(b.bound.p, bnd.data(), b.n * 4, hipMemcpyHostToDevice, b.st));
    KNNG_HIP(hipStreamSynchronize(b.st));
    return KNNG_OK;
}

// Large device-to-host copies for callers without a builder (knng_memcpy_d2h): two pinned 32 MB
// buffers on a stream of their own, each block copied on to the destination by host threads while
// the next block arrives (as Builder::download does). Serialised by a mutex; kept for the process.
struct HostStager {
    std::mutex mu;
    knng::PinBuf<unsigned char> stage[2];
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
};

HostStager& host_stager() {
    static HostStager s;
    return s;
}

int staged_d2h(void* dst, const void* src, size_t bytes) {
    constexpr size_t kBlk = Builder::kStageBytes;
    HostStager& s = host_stager();
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.st) {
        KNNG_HIP(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
        for (auto& e : s.ev) KNNG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    for (int b = 0; b < 2; ++b) KNNG_TRY(s.stage[b].ensure(kBlk));
    // the source may have been written on any stream: the blocking hipMemcpy this replaces waited
    // for the device as well
    KNNG_HIP(hipDeviceSynchronize());
    const unsigned char* s8 = static_cast<const unsigned char*>(src);
    unsigned char* d8 = static_cast<unsigned char*>(dst);
    size_t pend_off[2] = {0, 0}, pend_len[2] = {0, 0};
    auto drain = [&](int b) -> int {
        if (!pend_len[b]) return KNNG_OK;
        KNNG_HIP(hipEventSynchronize(s.ev[b]));
        const unsigned char* from = s.stage[b].p;
        unsigned char* to = d8 + pend_off[b];
        const size_t len = pend_len[b];
        Builder::parallel_rows(len / 4096 + 1, [=](size_t lo, size_t hi) {
            const size_t a0 = std::min(lo * 4096, len), a1 = std::min(hi * 4096, len);
            if (a1 > a0) std::memcpy(to + a0, from + a0, a1 - a0);
        });
        pend_len[b] = 0;
        return KNNG_OK;
    };
    for (size_t o = 0, blk = 0; o < bytes; o += kBlk, ++blk) {
        const int b = (int)(blk & 1);
        KNNG_TRY(drain(b));
        const size_t m = std::min(kBlk, bytes - o);
        KNNG_HIP(hipMemcpyAsync(s.stage[b].p, s8 + o, m, hipMemcpyDeviceToHost, s.st));
        KNNG_HIP(hipEventRecord(s.ev[b], s.st));
        pend_off[b] = o;
        pend_len[b] = m;
    }
    KNNG_TRY(drain(0));
    KNNG_TRY(drain(1));
    return KNNG_OK;
}

int download_state(Builder& b, uint32_t* ids, float* dists) {
    const size_t tot = (size_t)b.n * b.kl();
    KNNG_HIP(hipMemcpyAsync(ids, b.knn_ids.p, tot * 4, hipMemcpyDeviceToHost, b.st));
    KNNG_HIP(hipMemcpyAsync(dists, b.knn_d.p, tot * 4, hipMemcpyDeviceToHost, b.st));
    KNNG_HIP(hipStreamSynchronize(b.st));
    return KNNG_OK;
}

}  // namespace

extern "C" {

const char* knng_last_error(void) { return knng::g_err.c_str(); }

void knng_default_params(knng_params* p) {
    if (!p) return;
    *p = knng_params{};
    p->k = 100;
    p->leaf_max = 400;
    p->leaf_max_odd = 0;
    p->small_max = 3000;
    p->iterations = 1;
    p->rounds = 6;       // with 2 NoN passes: recall@100 0.95 on 10M x 100 embedding-like points
    p->non_passes = 2;
    p->precision = KNNG_PREC_FP32;
    p->rerank = 160;
    p->seed = 12345;
}

// ---- device scratch of the no-allocation leaf stage ------------------------------------------
// One array per GPU, loaded with the library's code object and held until the process ends. Host
// code never uses the name `scratch` as an address: each device's copy is found once with
// hipGetSymbolAddress (with that device current) and that address is passed to the kernels.
// Calls that overlap on one GPU share this memory and race.
static constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address, looked up on the first call for that device.
int device_scratch(unsigned char** out) {
    int dev = -1;
    KNNG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(KNNG_ERR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(KNNG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return KNNG_OK;
}
}  // namespace

int knng_create(const knng_params* params, knng_handle** out) {
    if (!params || !out) return fail(KNNG_ERR_INVALID, "null argument");
    *out = nullptr;
    auto* h = new knng_handle();
    int r = h->b.init(*params);
    unsigned char* sp = nullptr;
    if (r == KNNG_OK) r = device_scratch(&sp);  // this device's scratch address, set up once
    if (r != KNNG_OK) {
        delete h;
        return r;
    }
    *out = h;
    return KNNG_OK;
}

void knng_destroy(knng_handle* h) { delete h; }

int knng_set_params(knng_handle* h, const knng_params* params) {
    KNNG_TRY(check_handle(h));
    if (!params) return fail(KNNG_ERR_INVALID, "null params");
    std::string why;
    if (!knng::valid_params(*params, why)) return fail(KNNG_ERR_INVALID, why);
    h->b.prm = *params;
    return KNNG_OK;
}

int knng_get_stats(const knng_handle* h, knng_stats* out) {
    KNNG_TRY(check_handle(h));
    if (!out) return fail(KNNG_ERR_INVALID, "null stats");
    *out = h->b.stats;
    return KNNG_OK;
}

int knng_build_device(knng_handle* h, const float* d_points, uint32_t n, uint32_t d,
                      size_t row_stride, uint32_t* d_out_ids, float* d_out_dists) {
    KNNG_TRY(check_handle(h));
    if (n && !d_out_ids) return fail(KNNG_ERR_INVALID, "null output");
    Builder& b = h->b;
    KNNG_TRY(b.bind_points(d_points, true, n, d, row_stride));
    return b.build(d_out_ids, d_out_dists);
}

int knng_build(knng_handle* h, const float* points, uint32_t n, uint32_t d, size_t row_stride,
               uint32_t* out_ids, float* out_dists) {
    KNNG_TRY(check_handle(h));
    if (n && !out_ids) return fail(KNNG_ERR_INVALID, "null output");
    Builder& b = h->b;
    KNNG_TRY(b.bind_points(points, false, n, d, row_stride));
    const size_t tot = (size_t)n * b.prm.k;
    knng::DevBuf<uint32_t> oid;
    knng::DevBuf<float> od;
    KNNG_TRY(oid.ensure(tot));
    if (out_dists) KNNG_TRY(od.ensure(tot));
    KNNG_TRY(b.build(oid.p, out_dists ? od.p : nullptr));
    KNNG_TRY(b.download(out_ids, oid.p, tot * 4));
    if (out_dists) KNNG_TRY(b.download(out_dists, od.p, tot * 4));
    return KNNG_OK;
}

int knng_split_round(knng_handle* h, const float* points, uint32_t n, uint32_t d,
                     size_t row_stride, uint32_t round, uint32_t* perm_out, uint32_t* leaf_start,
                     uint32_t* leaf_size, uint32_t* n_leaves) {
    KNNG_TRY(check_handle(h));
    if (!perm_out || !leaf_start || !leaf_size || !n_leaves) return fail(KNNG_ERR_INVALID, "null output");
    Builder& b = h->b;
    KNNG_TRY(b.bind_points(points, false, n, d, row_stride));
    std::vector<knng::LeafDesc> lv;
    KNNG_TRY(b.split_round(round, lv, b.st, b.perm));
    if (n) KNNG_HIP(hipMemcpyAsync(perm_out, b.perm.p, (size_t)n * 4, hipMemcpyDeviceToHost, b.st));
    KNNG_HIP(hipStreamSynchronize(b.st));
    for (size_t i = 0; i < lv.size(); ++i) {
        leaf_start[i] = lv[i].start;
        leaf_size[i] = lv[i].size;
    }
    *n_leaves = (uint32_t)lv.size();
    return KNNG_OK;
}

int knng_partition_injected(knng_handle* h, const float* points, uint32_t n, uint32_t d,
                            size_t row_stride, uint32_t* perm, uint32_t n_nodes,
                            const uint32_t* node_start, const uint32_t* node_size,
                            const float* coefs, const float* offsets, uint32_t* lo_out) {
    KNNG_TRY(check_handle(h));
    if (!perm || !node_start || !node_size || !coefs || !offsets || !lo_out)
        return fail(KNNG_ERR_INVALID, "null argument");
    Builder& b = h->b;
    KNNG_TRY(b.bind_points(points, false, n, d, row_stride));
    for (uint32_t i = 0; i < n_nodes; ++i)
        if ((uint64_t)node_start[i] + node_size[i] > n || node_size[i] == 0)
            return fail(KNNG_ERR_INVALID, "node range outside [0, n)");
    KNNG_TRY(b.perm.ensure(n));
    KNNG_TRY(b.perm_tmp.ensure(n));
    KNNG_TRY(b.flags.ensure(n));
    KNNG_HIP(hipMemcpyAsync(b.perm.p, perm, (size_t)n * 4, hipMemcpyHostToDevice, b.st));
    KNNG_TRY(b.h_nodes.ensure(n_nodes));
    uint32_t nt = 0;
    for (uint32_t i = 0; i < n_nodes; ++i) {
        knng::NodeDesc& nd = b.h_nodes.p[i];
        nd = knng::NodeDesc{};
        nd.start = node_start[i];
        nd.size = node_size[i];
        nd.tile0 = nt;
        nd.ntiles = (nd.size + knng::kTileSize - 1) / knng::kTileSize;
        nd.status = knng::kNodeInjected;
        nt += nd.ntiles;
    }
    KNNG_TRY(b.h_tiles.ensure(nt));
    for (uint32_t i = 0; i < n_nodes; ++i) {
        const knng::NodeDesc& nd = b.h_nodes.p[i];
        for (uint32_t t = 0; t < nd.ntiles; ++t) {
            knng::TileDesc& tl = b.h_tiles.p[nd.tile0 + t];
            tl.node = i;
            tl.begin = t * knng::kTileSize;
            tl.count = std::min(knng::kTileSize, nd.size - tl.begin);
        }
    }
    KNNG_TRY(b.coef.ensure((size_t)n_nodes * b.ld));
    KNNG_TRY(b.off.ensure(n_nodes));
    std::vector<float> cf((size_t)n_nodes * b.ld, 0.f);
    for (uint32_t i = 0; i < n_nodes; ++i)
        std::memcpy(&cf[(size_t)i * b.ld], coefs + (size_t)i * d, d * sizeof(float));
    KNNG_HIP(hipMemcpyAsync(b.coef.p, cf.data(), cf.size() * 4, hipMemcpyHostToDevice, b.st));
    KNNG_HIP(hipMemcpyAsync(b.off.p, offsets, n_nodes * 4, hipMemcpyHostToDevice, b.st));
    b.split_perm = b.perm.p;
    KNNG_TRY(b.ensure_x16(b.st));
    KNNG_TRY(b.run_level(n_nodes, 0, nt, false, b.st));
    for (uint32_t i = 0; i < n_nodes; ++i) lo_out[i] = b.h_nodes.p[i].lo;
    KNNG_HIP(hipMemcpy(perm, b.perm.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    return KNNG_OK;
}

int knng_leaf_stage(knng_handle* h, const float* points, uint32_t n, uint32_t d,
                    size_t row_stride, const uint32_t* perm, const uint32_t* leaf_start,
                    const uint32_t* leaf_size, uint32_t n_leaves, uint32_t* ids, float* dists,
                    int start_empty) {
    KNNG_TRY(check_handle(h));
    if (!perm || !ids || !dists || (n_leaves && (!leaf_start || !leaf_size)))
        return fail(KNNG_ERR_INVALID, "null argument");
    Builder& b = h->b;
    KNNG_TRY(b.bind_points(points, false, n, d, row_stride));
    std::vector<knng::LeafDesc> lv(n_leaves);
    for (uint32_t i = 0; i < n_leaves; ++i) {
        if ((uint64_t)leaf_start[i] + leaf_size[i] > n) return fail(KNNG_ERR_INVALID, "leaf outside [0, n)");
        if (leaf_size[i] > knng::kMaxLeafAny) return fail(KNNG_ERR_UNSUPPORTED, "leaf larger than 65536");
        lv[i] = {leaf_start[i], leaf_size[i]};
    }
    KNNG_TRY(b.perm.ensure(n));
    KNNG_HIP(hipMemcpyAsync(b.perm.p, perm, (size_t)n * 4, hipMemcpyHostToDevice, b.st));
    KNNG_TRY(upload_state(b, ids, dists, start_empty != 0));
    KNNG_HIP(hipMemsetAsync(b.counters.p, 0, 4 * sizeof(uint64_t), b.st));
    KNNG_TRY(b.leaf_stage(lv, b.perm.p, 0, true, b.ev[4], b.ev[5]));
    uint64_t c[4] = {0, 0, 0, 0};
    KNNG_HIP(hipMemcpyAsync(c, b.counters.p, sizeof c, hipMemcpyDeviceToHost, b.st));
    KNNG_HIP(hipStreamSynchronize(b.st));
    b.stats.survivors = c[0];
    b.stats.inserts = c[2];
    return download_state(b, ids, dists);
}

int knng_leaf_stage_device_noalloc(knng_handle* h, const float* d_points, uint32_t n, uint32_t d,
                                   size_t row_stride, const uint32_t* d_perm,
                                   const uint32_t* leaf_start, const uint32_t* leaf_size,
                                   uint32_t n_leaves, uint32_t* d_ids, float* d_dists,
                                   int start_empty) {
    KNNG_TRY(check_handle(h));
    if (!d_perm || !d_ids || !d_dists || (n && !d_points) || (n_leaves && (!leaf_start || !leaf_size)))
        return fail(KNNG_ERR_INVALID, "null argument");
    if (d < 1 || d > 128 || d % 4 != 0 || row_stride < d || row_stride % 4 != 0 ||
        (reinterpret_cast<uintptr_t>(d_points) & 15) != 0)
        return fail(KNNG_ERR_UNSUPPORTED,
                    "the no-allocation leaf stage reads the rows in place: d a multiple of 4 up to "
                    "128, row_stride a multiple of 4, 16-byte aligned rows");
    Builder& b = h->b;
    std::vector<knng::LeafDesc> lv(n_leaves);
    for (uint32_t i = 0; i < n_leaves; ++i) {
        if ((uint64_t)leaf_start[i] + leaf_size[i] > n) return fail(KNNG_ERR_INVALID, "leaf outside [0, n)");
        if (leaf_size[i] > knng::kMaxLeafAny) return fail(KNNG_ERR_UNSUPPORTED, "leaf larger than 65536");
        lv[i] = {leaf_start[i], leaf_size[i]};
    }
    KNNG_TRY(b.bind_points(d_points, true, n, d, row_stride));  // in place: allocates nothing
    Builder::LeafPlan pl;
    KNNG_TRY(b.plan_leaves(lv, pl));
    // scratch layout: bounds (n floats), 4 counters, the binned leaves, the big-bucket blocks
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_cnt = up((size_t)n * sizeof(float));
    const size_t o_leaves = o_cnt + up(4 * sizeof(uint64_t));
    const size_t o_big = o_leaves + up(pl.nl * sizeof(knng::LeafDesc));
    const size_t need = o_big + up(pl.big.size() * sizeof(knng::BigBlock));
    if (need > kNoAllocScratchBytes)
        return fail(KNNG_ERR_UNSUPPORTED,
                    "the no-allocation leaf stage needs " + std::to_string(need) +
                        " bytes of scratch (4 per point plus the leaf tables); it holds " +
                        std::to_string(kNoAllocScratchBytes));
    unsigned char* sp = nullptr;
    KNNG_TRY(device_scratch(&sp));
    float* bnd = reinterpret_cast<float*>(sp);
    uint64_t* cnt = reinterpret_cast<uint64_t*>(sp + o_cnt);
    knng::LeafDesc* dl = reinterpret_cast<knng::LeafDesc*>(sp + o_leaves);
    knng::BigBlock* dbig = reinterpret_cast<knng::BigBlock*>(sp + o_big);
    const uint32_t k = b.kl();
    if (start_empty) {
        knng::launch_init_knn(d_ids, d_dists, bnd, n, k, b.st);
        knng::launch_zero_u64(cnt, 4, b.st);
    } else {
        knng::launch_bound_from_lists(d_dists, n, k, bnd, cnt, 4, b.st);
    }
    // descriptors from ordinary host memory; the stream is synchronised before they go out of scope
    std::vector<knng::LeafDesc> flat;
    flat.reserve(pl.nl);
    for (auto& v : pl.bin) flat.insert(flat.end(), v.begin(), v.end());
    if (!flat.empty())
        KNNG_HIP(hipMemcpyAsync(dl, flat.data(), flat.size() * sizeof(knng::LeafDesc),
                                hipMemcpyHostToDevice, b.st));
    if (!pl.big.empty())
        KNNG_HIP(hipMemcpyAsync(dbig, pl.big.data(), pl.big.size() * sizeof(knng::BigBlock),
                                hipMemcpyHostToDevice, b.st));
    const int rc = n_leaves ? b.launch_leaf_plan(pl, dl, dbig, d_perm, d_ids, d_dists, bnd, cnt,
                                                 false, true, b.ev[4], b.ev[5])
                            : KNNG_OK;
    KNNG_HIP(hipStreamSynchronize(b.st));
    return rc;
}

int knng_non(knng_handle* h, const float* points, uint32_t n, uint32_t d, size_t row_stride,
             uint32_t* ids, float* dists) {
    KNNG_TRY(check_handle(h));
    if (!ids || !dists) return fail(KNNG_ERR_INVALID, "null argument");
    Builder& b = h->b;
    KNNG_TRY(b.bind_points(points, false, n, d, row_stride));
    KNNG_TRY(upload_state(b, ids, dists, false));
    KNNG_HIP(hipMemsetAsync(b.counters.p, 0, 4 * sizeof(uint64_t), b.st));
    b.perm_is_tree = false;
    b.have_prev = false;  // one full pass per call
    b.pass_no = 0;
    KNNG_TRY(b.non_pass());
    return download_state(b, ids, dists);
}

int knng_exact_upper_from_lists(knng_handle* h, const float* d_points, uint32_t n, uint32_t d,
                                size_t row_stride, const uint32_t* d_queries, uint32_t nq,
                                const uint32_t* d_lists, uint32_t k, double* d_upper) {
    KNNG_TRY(check_handle(h));
    Builder& b = h->b;
    KNNG_TRY(b.bind_points(d_points, true, n, d, row_stride));
    KNNG_TRY(b.norms64.ensure(n));
    knng::launch_norms64(b.X, b.ld, b.d, n, b.norms64.p, b.st);
    knng::ExactScanArgs a{};
    a.X = b.X;
    a.ld = b.ld;
    a.d = b.d;
    a.n = n;
    a.norms = b.norms64.p;
    a.queries = d_queries;
    a.nq = nq;
    knng::launch_exact_ub(a, d_lists, k, d_upper, b.st);
    KNNG_HIP(hipGetLastError());
    KNNG_HIP(hipStreamSynchronize(b.st));
    return KNNG_OK;
}

int knng_exact_knn_device(knng_handle* h, const float* d_points, uint32_t n, uint32_t d,
                          size_t row_stride, const uint32_t* d_queries, uint32_t nq, uint32_t k,
                          const double* d_upper, uint32_t cap, uint32_t* d_out_ids,
                          double* d_out_dists) {
    KNNG_TRY(check_handle(h));
    if (!d_queries || !d_out_ids || !d_out_dists) return fail(KNNG_ERR_INVALID, "null argument");
    if (k < 1 || cap < k) return fail(KNNG_ERR_INVALID, "need 1 <= k <= cap");
    Builder& b = h->b;
    KNNG_TRY(b.bind_points(d_points, true, n, d, row_stride));
    if (knng::exact_scan_lds(d) > 160 * 1024) return fail(KNNG_ERR_UNSUPPORTED, "d too large");
    KNNG_TRY(b.norms64.ensure(n));
    KNNG_TRY(b.cand_d.ensure((size_t)nq * cap));
    KNNG_TRY(b.cand_id.ensure((size_t)nq * cap));
    KNNG_TRY(b.qcount.ensure(nq));
    KNNG_TRY(b.ovf.ensure(1));
    const double* ub = d_upper;
    if (!ub) {
        KNNG_TRY(b.ubuf.ensure(nq));
        std::vector<double> inf(nq, std::numeric_limits<double>::infinity());
        KNNG_HIP(hipMemcpyAsync(b.ubuf.p, inf.data(), nq * 8, hipMemcpyHostToDevice, b.st));
        ub = b.ubuf.p;
    }
    knng::launch_norms64(b.X, b.ld, b.d, n, b.norms64.p, b.st);
    KNNG_TRY(b.exact_pass(d_queries, nq, k, ub, cap, d_out_ids, d_out_dists));
    // queries whose bound admitted more than cap points: one more scan over just those, with
    // a buffer sized from their counts
    std::vector<uint32_t> cnt(nq);
    KNNG_HIP(hipMemcpy(cnt.data(), b.qcount.p, (size_t)nq * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> over;
    uint32_t cap2 = 0;
    for (uint32_t i = 0; i < nq; ++i)
        if (cnt[i] > cap) {
            over.push_back(i);
            cap2 = std::max(cap2, cnt[i]);
        }
    if (over.empty()) return KNNG_OK;
    if ((double)over.size() * cap2 * 12.0 > 8e9)
        return fail(KNNG_ERR_OVERFLOW, std::to_string(over.size()) +
                                           " queries need more candidate memory than allowed");
    const uint32_t no = (uint32_t)over.size();
    knng::DevBuf<uint32_t> d_where, d_q2, d_i2;
    knng::DevBuf<double> d_u2, d_d2;
    KNNG_TRY(d_where.ensure(no));
    KNNG_TRY(d_q2.ensure(no));
    KNNG_TRY(d_u2.ensure(no));
    KNNG_TRY(d_i2.ensure((size_t)no * k));
    KNNG_TRY(d_d2.ensure((size_t)no * k));
    KNNG_HIP(hipMemcpyAsync(d_where.p, over.data(), (size_t)no * 4, hipMemcpyHostToDevice, b.st));
    knng::launch_gather_u32(d_queries, d_where.p, no, d_q2.p, b.st);
    knng::launch_gather_f64(ub, d_where.p, no, d_u2.p, b.st);
    KNNG_TRY(b.exact_pass(d_q2.p, no, k, d_u2.p, cap2, d_i2.p, d_d2.p));
    knng::launch_exact_scatter(d_where.p, no, k, d_i2.p, d_d2.p, d_out_ids, d_out_dists, b.st);
    KNNG_HIP(hipGetLastError());
    KNNG_HIP(hipStreamSynchronize(b.st));
    return KNNG_OK;
}

double knng_recall(const uint32_t* result_ids, const uint32_t* truth_ids, uint32_t rows,
                   uint32_t k) {
    if (!rows || !k || !result_ids || !truth_ids) return 0.0;
    uint64_t hits = 0;
    std::vector<uint32_t> r(k);
    for (uint32_t i = 0; i < rows; ++i) {
        std::memcpy(r.data(), result_ids + (size_t)i * k, k * 4);
        std::sort(r.begin(), r.end());
        for (uint32_t j = 0; j < k; ++j)
            hits += std::binary_search(r.begin(), r.end(), truth_ids[(size_t)i * k + j]);
    }
    return (double)hits / ((double)rows * k);
}

int knng_generate_device(knng_handle* h, int kind, uint32_t n, uint32_t d, size_t row_stride,
                         uint64_t seed, float* d_points) {
    KNNG_TRY(check_handle(h));
    if (!d_points || d < 1 || d > 256 || row_stride < d) return fail(KNNG_ERR_INVALID, "bad shape");
    Builder& b = h->b;
    // host-side parameters from a splitmix64 stream of the seed
    uint64_t state = knng::mix64(seed ^ 0xA0761D6478BD642Full);
    auto next_u = [&]() {
        state += 0x9E3779B97F4A7C15ull;
        return (double)(knng::mix64(state) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto next_n = [&]() {
        const double u1 = std::max(next_u(), 1e-300), u2 = next_u();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    if (kind == KNNG_GEN_LOWRANK) {
        const uint32_t latent = 24;
        std::vector<float> W((size_t)latent * d);
        for (auto& w : W) w = (float)next_n();
        KNNG_TRY(b.gen_a.ensure(W.size()));
        KNNG_HIP(hipMemcpyAsync(b.gen_a.p, W.data(), W.size() * 4, hipMemcpyHostToDevice, b.st));
        knng::launch_gen_lowrank(d_points, n, d, (uint32_t)row_stride, latent, 0.05f, seed, b.gen_a.p, b.st);
    } else if (kind == KNNG_GEN_MIXTURE || kind == KNNG_GEN_ZIPF) {
        const bool zipf = kind == KNNG_GEN_ZIPF;
        const uint32_t nc = zipf ? 20000 : 500;
        std::vector<float> C((size_t)nc * d), cdf(nc), sg(nc);
        for (auto& c : C) c = (float)(-10.0 + 20.0 * next_u());
        double tot = 0.0;
        std::vector<double> w(nc);
        for (uint32_t c = 0; c < nc; ++c) {
            w[c] = zipf ? 1.0 / std::pow((double)(c + 1), 1.1) : 1.0;
            tot += w[c];
            sg[c] = zipf ? (float)(0.3 + 1.7 * next_u()) : 1.0f;
        }
        double run = 0.0;
        for (uint32_t c = 0; c < nc; ++c) {
            run += w[c];
            cdf[c] = (float)(run / tot);
        }
        cdf[nc - 1] = 1.0f;
        KNNG_TRY(b.gen_a.ensure(C.size()));
        KNNG_TRY(b.gen_b.ensure(nc));
        KNNG_TRY(b.gen_c.ensure(nc));
        KNNG_HIP(hipMemcpyAsync(b.gen_a.p, C.data(), C.size() * 4, hipMemcpyHostToDevice, b.st));
        KNNG_HIP(hipMemcpyAsync(b.gen_b.p, cdf.data(), nc * 4, hipMemcpyHostToDevice, b.st));
        KNNG_HIP(hipMemcpyAsync(b.gen_c.p, sg.data(), nc * 4, hipMemcpyHostToDevice, b.st));
        knng::launch_gen_mixture(d_points, n, d, (uint32_t)row_stride, b.gen_a.p, nc, b.gen_b.p,
                                 b.gen_c.p, seed, b.st);
    } else {
        return fail(KNNG_ERR_INVALID, "unknown generator kind");
    }
    KNNG_HIP(hipGetLastError());
    KNNG_HIP(hipStreamSynchronize(b.st));
    return KNNG_OK;
}

int knng_mem_get_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(KNNG_ERR_INVALID, "null argument");
    KNNG_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return KNNG_OK;
}

int knng_device_alloc(size_t bytes, void** out) {
    if (!out) return fail(KNNG_ERR_INVALID, "null out");
    KNNG_HIP(hipMalloc(out, std::max<size_t>(bytes, 1)));
    return KNNG_OK;
}
int knng_device_free(void* p) {
    KNNG_HIP(hipFree(p));
    return KNNG_OK;
}
int knng_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    KNNG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return KNNG_OK;
}
int knng_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (bytes >= (64u << 20)) return staged_d2h(dst, src, bytes);
    KNNG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return KNNG_OK;
}
int knng_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    KNNG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return KNNG_OK;
}
int knng_synchronize(void) {
    KNNG_HIP(hipDeviceSynchronize());
    return KNNG_OK;
}

}  // extern "C"

// ---- serving -------------------------------------------------------------------------------------
namespace knng {

struct ServerJob {
    knng_job job{};
    uint64_t footprint = 0;
    std::chrono::steady_clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = KNNG_OK;
    uint32_t worker = 0;
    uint64_t held_after = 0;
};

}  // namespace knng

struct knng_server {
    // Admission accounting, under mu. A worker's charge is the device memory its builder keeps
    // between jobs; while it runs a job, that plus the job's footprint (grow-only buffers free the
    // old allocation before taking a larger one, so a job's peak is at most the two together).
    // in_use is the sum of the charges and never exceeds budget.
    uint64_t budget = 0, in_use = 0, peak_in_use = 0, trims = 0;
    std::vector<uint64_t> charge;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<knng::ServerJob*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<knng::ServerJob>> jobs;
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
};

namespace {

// What a serving worker keeps between jobs: a builder (stream, device buffers) and the device
// outputs it downloads from.
struct WorkerState {
    Builder b;
    knng::DevBuf<uint32_t> oid;
    knng::DevBuf<float> od;
    int init_rc = KNNG_OK;
    WorkerState() {
        knng_params p;
        knng_default_params(&p);
        init_rc = b.init(p);
    }
};

// One worker. It takes the queue's head job when the charges of all workers plus the job's
// footprint fit the budget. When the head does not fit and the worker holds device memory, it
// frees that memory first (a fresh builder), so memory kept by idle workers never blocks a job.
void server_worker(knng_server* s, uint32_t w) {
    std::atomic<uint64_t> held{0};  // device bytes this thread's buffers hold
    knng::tl_device_account = &held;
    auto ws = std::make_unique<WorkerState>();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->in_use += held - s->charge[w];
        s->charge[w] = held;
        s->peak_in_use = std::max(s->peak_in_use, s->in_use);
    }
    for (;;) {
        knng::ServerJob* j = nullptr;
        bool trim = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->queue.empty()) return s->stop;
                return s->in_use + s->queue.front()->footprint <= s->budget || s->charge[w] > 0;
            });
            if (s->queue.empty()) return;  // stopping
            if (s->in_use + s->queue.front()->footprint <= s->budget) {
                j = s->queue.front();
                s->queue.pop_front();
                s->charge[w] += j->footprint;
                s->in_use += j->footprint;
                s->peak_in_use = std::max(s->peak_in_use, s->in_use);
                j->t_start = std::chrono::steady_clock::now();
                j->worker = w;
            } else {
                trim = true;  // the charge stays counted until the memory is actually freed
            }
        }
        if (trim) {
            ws.reset();
            ws = std::make_unique<WorkerState>();
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->in_use = s->in_use - s->charge[w] + held;
                s->charge[w] = held;
                s->trims++;
            }
            s->cv_work.notify_all();
            continue;
        }
        int rc = ws->init_rc;
        if (rc == KNNG_OK) {
            Builder& b = ws->b;
            const knng_job& jb = j->job;
            const size_t tot = (size_t)jb.n * jb.params.k;
            b.prm = jb.params;
            rc = b.bind_points(jb.points, jb.points_on_device != 0, jb.n, jb.d, jb.row_stride);
            if (rc == KNNG_OK) rc = ws->oid.ensure(std::max<size_t>(tot, 1));
            if (rc == KNNG_OK && jb.out_dists) rc = ws->od.ensure(std::max<size_t>(tot, 1));
            if (rc == KNNG_OK) rc = b.build(ws->oid.p, jb.out_dists ? ws->od.p : nullptr);
            if (rc == KNNG_OK) rc = b.download(jb.out_ids, ws->oid.p, tot * 4);
            if (rc == KNNG_OK && jb.out_dists) rc = b.download(jb.out_dists, ws->od.p, tot * 4);
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->in_use = s->in_use - s->charge[w] + held;
            s->charge[w] = held;
            s->peak_in_use = std::max(s->peak_in_use, s->in_use);
            j->status = rc;
            j->done = true;
            j->t_done = std::chrono::steady_clock::now();
            j->held_after = held;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

uint64_t knng_job_footprint(uint32_t n, uint32_t d, const knng_params* params) {
    if (!params) return 0;
    const uint64_t ld = (d + 3) & ~3u, k = std::max(params->k, params->search_k);
    const bool incremental = (params->flags & KNNG_FLAG_INCREMENTAL_NON) != 0;
    const bool kd = (params->flags & KNNG_FLAG_KD_SPLIT) != 0;
    // Per point: the points copy; top-k ids/dists; outputs; the snapshot and the reverse lists
    // (plus the previous snapshot in incremental passes; the full passes' sort of the reverse lists
    // works in the first candidate set, or beside it for 16-slot buffers); two sets of NoN
    // candidate buffers; three permutations, side flags, bounds, norms, seven per-point counters,
    // overflow list, reverse offsets; two sets of leaf descriptors (at most one leaf per point);
    // projected rows for the KD splitter.
    uint64_t per_point = 4 * ld + 8 * k + 8 * k + (incremental ? 12 : 8) * k +
                         16 * (uint64_t)Builder::kNonCap + 12 + 1 + 4 + 4 + 28 + 4 + 8 + 16;
    if (kd) per_point += 4 * (uint64_t)knng::kKdStride;
    else if (!(params->flags & KNNG_FLAG_SIDE_FP32)) per_point += 2 * (uint64_t)knng::split_ld16(d);  // fp16 rows
    // Split tables of one level: at most one node per leaf_max + 1 points, each with its
    // descriptor, three centre/hyperplane rows, an offset and two FP64 partial rows per chunk
    // (one chunk per node plus one per 256 sample draws), and one tile per 1024 positions.
    const uint64_t lm = std::max<uint32_t>(1, std::min(params->leaf_max, params->leaf_max_odd
                                                                            ? params->leaf_max_odd
                                                                            : params->leaf_max));
    const uint64_t nodes = (uint64_t)n / (lm + 1) + 1;
    const uint64_t node_bytes = sizeof(knng::NodeDesc) + 12 * ld + 4 + 2 * (16 * ld + 8 + sizeof(knng::ChunkDesc));
    const uint64_t tiles = (uint64_t)n / 1024 + nodes + 1;
    const uint64_t scan = ((uint64_t)n / 1024 + 2) * 8;
    return (uint64_t)n * per_point + nodes * node_bytes + tiles * (sizeof(knng::TileDesc) + 8) + scan +
           (64ull << 20);
}

int knng_server_create(uint32_t workers, uint64_t hbm_budget, knng_server** out) {
    if (!out || workers == 0 || workers > 64) return fail(KNNG_ERR_INVALID, "workers must be in [1, 64]");
    auto* s = new knng_server();
    s->budget = hbm_budget;
    s->charge.assign(workers, 0);
    for (uint32_t w = 0; w < workers; ++w) s->workers.emplace_back(server_worker, s, w);
    *out = s;
    return KNNG_OK;
}

int knng_server_submit(knng_server* s, const knng_job* job, uint64_t* job_id) {
    if (!s || !job || !job_id) return fail(KNNG_ERR_INVALID, "null argument");
    std::string why;
    if (!knng::valid_params(job->params, why)) return fail(KNNG_ERR_INVALID, why);
    if (job->n && (!job->points || !job->out_ids)) return fail(KNNG_ERR_INVALID, "null points or output");
    if (job->d < 1 || job->d > knng::kMaxDim || job->row_stride < job->d)
        return fail(KNNG_ERR_INVALID, "bad d or row_stride");
    auto j = std::make_unique<knng::ServerJob>();
    j->job = *job;
    j->footprint = knng_job_footprint(job->n, job->d, &job->params);
    if (j->footprint > s->budget) return fail(KNNG_ERR_NOMEM, "job footprint exceeds the server's HBM budget");
    j->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return fail(KNNG_ERR_INVALID, "server is stopping");
        *job_id = s->next_id++;
        s->queue.push_back(j.get());
        s->jobs.emplace(*job_id, std::move(j));
    }
    s->cv_work.notify_all();
    return KNNG_OK;
}

int knng_server_poll(knng_server* s, uint64_t job_id, int* done) {
    if (!s || !done) return fail(KNNG_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(KNNG_ERR_INVALID, "unknown job id");
    *done = it->second->done ? 1 : 0;
    return KNNG_OK;
}

int knng_server_wait(knng_server* s, uint64_t job_id, knng_job_result* out) {
    if (!s) return fail(KNNG_ERR_INVALID, "null server");
    std::unique_ptr<knng::ServerJob> j;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job_id);
        if (it == s->jobs.end()) return fail(KNNG_ERR_INVALID, "unknown job id");
        knng::ServerJob* jp = it->second.get();
        s->cv_done.wait(lk, [&] { return jp->done; });
        j = std::move(it->second);
        s->jobs.erase(it);
    }
    if (out) {
        out->status = j->status;
        out->queue_ms = ms_between(j->t_submit, j->t_start);
        out->run_ms = ms_between(j->t_start, j->t_done);
        out->latency_ms = ms_between(j->t_submit, j->t_done);
        out->footprint = j->footprint;
        out->worker = j->worker;
        out->start_ms = ms_between(s->t0, j->t_start);
        out->done_ms = ms_between(s->t0, j->t_done);
        out->worker_device_bytes = j->held_after;
    }
    return j->status == KNNG_OK ? KNNG_OK : fail(j->status, "job failed");
}

int knng_server_get_stats(knng_server* s, knng_server_stats* out) {
    if (!s || !out) return fail(KNNG_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->budget = s->budget;
    out->in_use = s->in_use;
    out->peak_in_use = s->peak_in_use;
    out->trims = s->trims;
    out->queued = s->queue.size();
    return KNNG_OK;
}

void knng_server_destroy(knng_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    delete s;
}
