// Descriptors and launchers shared between the host driver and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace knng {

constexpr uint32_t kTileSize = 1024;       // partition positions per block
constexpr uint32_t kMaxSmallNode = 4096;   // small-branch nodes are summed by one block
constexpr uint32_t kMaxLeaf = 512;         // largest leaf bucket handled by the LDS leaf kernel
constexpr uint32_t kMaxLeafAny = 65536;    // largest leaf bucket (beyond kMaxLeaf: k_leaf_big)
constexpr uint32_t kLeafBigRows = 64;      // rows of a bucket per k_leaf_big workgroup
constexpr uint32_t kMaxK = 128;            // neighbours per point

enum : uint32_t { kNodeActive = 0, kNodeFailed = 1, kNodeInjected = 2 };

// One node of the current split level (a contiguous range of the permutation).
struct NodeDesc {
    uint32_t start, size;
    uint64_t akey;      // attempt key of this node's random draws
    uint32_t m;         // sample draws for the Lloyd passes (large branch), 0 for the small branch
    uint32_t s;         // rows sampled for the farthest-pair start
    uint32_t chunk0, nchunks;
    uint32_t tile0, ntiles;
    uint32_t status;    // kNode*; the device marks failed attempts
    uint32_t lo;        // out: size of the lo child
};
static_assert(sizeof(NodeDesc) == 48, "NodeDesc layout");

struct ChunkDesc {
    uint32_t node, begin, count;
};
struct TileDesc {
    uint32_t node, begin, count;
};

struct SplitLevelArgs {
    const float* X;
    uint32_t ld, d;
    const uint16_t* X16;  // nullable: fp16 copy of the rows (launch_rows_to_half) for the side test
    uint32_t ld16;
    uint32_t* perm;
    uint32_t* perm_tmp;
    NodeDesc* nodes;
    uint32_t nnodes;
    const ChunkDesc* chunks;
    uint32_t nchunks;
    const TileDesc* tiles;
    uint32_t ntiles;
    float *c1, *c2, *coef, *off;
    double* part;
    uint32_t* pcnt;
    uint8_t* flags;
    uint32_t* tile_lo;
    uint32_t* tile_base;
    uint32_t iterations;
    bool run_lloyd;     // false for injected hyperplanes
};

// Random-rotation KD splitter (knng_kd.hip): projections onto kKdDims directions, rows of
// kKdStride floats; per-node FP64 means and sample variances; side flags for the shared partition.
constexpr uint32_t kKdDims = 30;    // newDims, SolutionRandomKD.hpp:387
constexpr uint32_t kKdStride = 32;  // floats per projected row (and per node statistics row)
size_t kd_rotate_lds(uint32_t ld);
void launch_kd_rotate(const float* X, uint32_t ld, uint32_t d, const float* R, uint32_t m,
                      uint32_t n, float* rot, hipStream_t st);
void launch_kd_stats(const float* rot, const uint32_t* perm, const NodeDesc* nodes, uint32_t nnodes,
                     const TileDesc* tiles, uint32_t ntiles, uint32_t m, double* part,
                     double* mean, double* var, hipStream_t st);
void launch_kd_side(const float* rot, const uint32_t* perm, const NodeDesc* nodes,
                    const TileDesc* tiles, uint32_t ntiles, const uint32_t* kdim,
                    const double* ksplit, uint8_t* flags, uint32_t* tile_lo, hipStream_t st);

void launch_iota(uint32_t* perm, uint32_t n, hipStream_t st);
void launch_split_prepare(const SplitLevelArgs& a, hipStream_t st);    // start pair + Lloyd
void launch_split_side(const SplitLevelArgs& a, hipStream_t st);       // full-range side test
void launch_split_partition(const SplitLevelArgs& a, hipStream_t st);  // stable scatter
// fp16 copy of the rows for the side test: ld16 = split_ld16(d) halves per row, zero past d
inline uint32_t split_ld16(uint32_t d) { return (d + 63) & ~63u; }
void launch_rows_to_half(const float* X, uint32_t ld, uint32_t d, uint32_t n, uint16_t* X16, uint32_t ld16,
                         hipStream_t st);
size_t leaf_lds_bytes(uint32_t max_leaf, uint32_t d, uint32_t ld);

// ---- leaf stage -------------------------------------------------------------------------------
struct LeafDesc {
    uint32_t start, size;
};

enum : uint32_t { kScreenExact = 0, kScreenApprox = 1 };

struct LeafArgs {
    const float* X;
    uint32_t ld, d, k;
    const uint32_t* perm;
    const LeafDesc* leaves;
    uint32_t nleaves;
    uint32_t* knn_ids;     // n*k, each row sorted by (dist, id), padded (FLT_MAX, kNoId)
    float* knn_d;
    float* bound;          // n: knn_d[row][k-1]
    uint32_t screen;       // kScreenExact: rigorous margin; kScreenApprox: plain approximate filter
    uint32_t half_kind;    // 0 = fp16, 1 = bf16 screening Gram, 2 (kHalfF32) = FP32 Gram (big kernel)
    uint32_t max_leaf;     // largest leaf size in this launch
    uint64_t* counters;    // [0] survivors screened, [1] exact distances, [2] inserts (nullable)
    uint32_t tune;         // 1: pick the workgroup shape by timing both (KNNG_FLAG_AUTOTUNE)
    // In-bucket limit m (0: none, else at most 160). A row's screen bound is lowered to the m-th
    // smallest of its in-bucket distance bounds: upper bounds of the exact distances in the exact
    // screen (m = k: no pair of the k nearest is lost, so the lists are the same), the approximate
    // distances otherwise (the re-rank cap: about the m nearest by the screen are re-ranked).
    uint32_t rank_cap;
};
void launch_leaf(const LeafArgs& a, hipStream_t st);

// Buckets beyond the LDS leaf kernel: one workgroup per kLeafBigRows-row block of a bucket.
struct BigBlock {
    uint32_t start, size, row0;  // bucket range in perm, first row of this block
};
size_t leaf_big_lds_bytes(uint32_t d, uint32_t ld);
void launch_leaf_big(const LeafArgs& a, const BigBlock* blocks, uint32_t nblocks, hipStream_t st);

// Batched form: one persistent kernel takes the round's work from one list, the big-bucket blocks
// first, then the LDS-resident buckets (at most leaf_batched_limit points, LeafArgs::max_leaf set
// to it) by decreasing size.
struct LeafQueue {
    uint32_t next;   // next item to take (reset by the launch)
    uint32_t nbig;   // items [0, nbig): BigBlock entries
    uint32_t nleaf;  // items [nbig, nbig + nleaf): LeafDesc entries (LeafArgs::leaves)
    uint32_t pad_;
};
uint32_t leaf_batched_limit(uint32_t d, uint32_t ld);
size_t leaf_batched_lds(uint32_t d, uint32_t ld);
// The same two launches every round (queue reset, persistent kernel on the chip's resident slots).
void launch_leaf_batched(const LeafArgs& a, const BigBlock* blocks, LeafQueue* q, hipStream_t st);

void launch_init_knn(uint32_t* ids, float* dists, float* bound, uint32_t n, uint32_t k,
                     hipStream_t st);
// bound[p] = dists[p * k + k - 1] for p < n, and counters[0, ncounters) = 0.
void launch_bound_from_lists(const float* dists, uint32_t n, uint32_t k, float* bound,
                             uint64_t* counters, uint32_t ncounters, hipStream_t st);
void launch_zero_u64(uint64_t* p, uint32_t count, hipStream_t st);
// First k of each width-kl list -> out (padding id 1 for empty slots), dists -> out_d (nullable).
void launch_finalize(const uint32_t* ids, const float* dists, uint32_t n, uint32_t kl, uint32_t k,
                     uint32_t* out_ids, float* out_d, hipStream_t st);

// ---- neighbours of neighbours ------------------------------------------------------------------
struct NonArgs {
    const float* X;
    uint32_t ld, d, k, n;
    const uint32_t* snap;  // n*k snapshot ids (unfilled slots read as id 0)
    uint32_t* knn_ids;
    float* knn_d;
    float* bound;
};
struct NonRevArgs {
    const float* X;
    uint32_t ld, d, k, n;
    const uint32_t* snap;   // n*k snapshot ids, each row sorted ascending
    const uint64_t* roff;   // n+1 offsets of the reverse lists
    const uint32_t* rs;     // reverse lists: RS[q] = { p : q in snap[p] }, new entries first
    const uint32_t* rnew;   // n: new entries at the head of RS[q] (nullable: all new)
    const uint32_t* order;  // q processing order (tree order of the last round), this launch's part
    uint32_t order_stride;  // this launch takes order[0], order[stride], order[2 stride], ...
    uint32_t nq;            // q in this launch
    uint32_t* knn_ids;
    float* knn_d;
    float* bound;           // read as the pre-pass bound by the screen, updated by the merge
    uint32_t* ncnt;         // n: candidates appended per point
    float* nbuf_d;          // n*cap
    uint32_t* nbuf_id;
    uint32_t cap;
    uint32_t half_kind;
    uint32_t screen16;      // 0: FP32 Gram of the rows; 1: fp16, 2: bf16 Gram of rows centred on q;
                            // 3: fp16 Gram of the rows, each scaled by a power of two from its norm
    uint32_t exact;         // 16-bit screens: 1 = rigorous error margin, 0 = plain approximate test
    uint32_t unscaled;      // screen 3: every nonzero squared norm lies in [2^-12, 2^28], so rows
                            // convert to fp16 without their power-of-two scales
    const float* norms;     // n FP32 squared norms
    uint64_t* counters;     // [0] screened pairs, [1] appended (nullable)
};
void launch_row_norms(const float* X, uint32_t ld, uint32_t d, uint32_t n, float* out,
                      hipStream_t st);
// out[0]: the smallest nonzero, out[1]: the largest of n non-negative floats, as float bits
void launch_norm_range(const float* norms, uint32_t n, uint32_t* out, hipStream_t st);

// Sorted snapshot; an entry absent from `prev` (nullable: all) carries kNewBit, an expanded one
// kExpBit. rank_lo = 0: unfilled slots -> id 0, every entry expanded; rank_lo > 0: unfilled
// slots -> the point's own id, filled entries at ranks >= rank_lo expanded.
// rcnt (nullable): also count the reverse lists (their sizes for launch_reverse_lists), and rnew
// (nullable) their new entries.
// pkey/pval (nullable, used without rcnt): n*k (id, p) pairs of the entries that go to a reverse
// list, 0xFFFFFFFF keys for the other slots; sorted by key (rsort_pairs) their values are the
// reverse lists, and launch_roff_search gives their offsets.
void launch_non_snapshot(const uint32_t* ids, uint32_t* snap, uint32_t n, uint32_t k,
                         const uint32_t* prev, uint32_t rank_lo, uint32_t* rcnt, uint32_t* rnew,
                         uint32_t* pkey, uint32_t* pval, hipStream_t st);
void launch_roff_search(const uint32_t* keys, uint64_t count, uint32_t n, uint64_t* roff, hipStream_t st);
// Stable radix sort of u32 (key, value) pairs by the keys' low `bits` bits (knng_rsort.hip): the
// pairs start in (keys_a, vals_a) and end in the a or the b buffers (*in_b).
size_t rsort_temp_bytes(size_t count, uint32_t bits);
hipError_t rsort_pairs(void* temp, size_t temp_bytes, uint32_t* keys_a, uint32_t* keys_b, uint32_t* vals_a,
                       uint32_t* vals_b, size_t count, uint32_t bits, hipStream_t st, bool* in_b);
// plist == nullptr: points 0..np-1.
void launch_non_direct(const NonArgs& a, const uint32_t* plist, uint32_t np, hipStream_t st);
size_t scan_scratch_elems(uint32_t n);
void launch_exclusive_scan(const uint32_t* in, uint32_t n, uint64_t* out, uint64_t* scratch,
                           hipStream_t st);
// rcnt (and rnew) hold the counts launch_non_snapshot made: offsets, then the lists (entries
// placed by atomics; rnew nullable: no new-first order)
void launch_reverse_lists(const uint32_t* snap, uint32_t n, uint32_t k, uint32_t* rcnt,
                          uint32_t* rnew, uint64_t* roff, uint64_t* scratch, uint32_t* rfill,
                          uint32_t* rfill_old, uint32_t* rs, hipStream_t st);
size_t non_reverse_lds(uint32_t k, uint32_t ld, uint32_t d, uint32_t screen16);
// The reverse kernel's 16-bit centred screens cover d with 4 ceil(d / 4) in {64, 100, 128}.
bool non_screen16_supported(uint32_t d);
void launch_non_reverse(const NonRevArgs& a, hipStream_t st);
void launch_non_merge(const NonRevArgs& a, uint32_t* ovf_cnt, uint32_t* ovf_list, uint32_t* ovf_flag,
                      hipStream_t st);

// ---- synthetic data -------------------------------------------------------------------------------
void launch_gen_lowrank(float* X, uint32_t n, uint32_t d, uint32_t ld, uint32_t latent,
                        float noise, uint64_t seed, const float* W, hipStream_t st);
void launch_gen_mixture(float* X, uint32_t n, uint32_t d, uint32_t ld, const float* centres,
                        uint32_t ncentres, const float* cdf, const float* sigma, uint64_t seed,
                        hipStream_t st);

}  // namespace knng
