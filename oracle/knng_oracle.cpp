// CPU restatement of the reference's approximate KNN-graph path (SolutionKmeans::constructResult).
//
// TEST INFRASTRUCTURE ONLY. This file is the parity oracle: tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it, as the checker. The product library never links it.
//
// Every function cites the reference lines it restates (paths relative to the reference root).
// Where the reference leaves an order or a case undefined (TBB combine order, -Ofast contraction,
// thread_local mt19937 draws, null dereference on degenerate samples), this file fixes one
// definition; the GPU library implements the same definition, and the tests hold both to it.
//
// Fixed definitions (shared with the GPU library, restated independently there):
//  * dist/dot: 8 lane accumulators, lane j takes dims j, j+8, ... below floor(d/8)*8 with fused
//    multiply-add, lanes summed 0..7 into 0.0f in order, then the tail dims with fused
//    multiply-add (LinearAlgebra.hpp:53-77, :179-200).
//  * Random draws: counter-based splitmix64 keyed by (seed, round, node path, attempt, draw index)
//    in place of thread_local mt19937(12345) (Constants.hpp:22) and uniform_int_distribution
//    (SolutionKmeans.hpp:60, :108).
//  * Centroid sums: FP64, sequential over the range for the small branch (as
//    SolutionKmeans.hpp:170-175); for the large branch sequential within chunks of 256 sample
//    draws and the chunk partials added in chunk order (a fixed instance of the tbb::combinable
//    reduction at SolutionKmeans.hpp:226-249).
//  * Partition: stable for both branches (std::stable_partition at :193-197; the large branch's
//    combinable concatenation order at :262-304 is scheduling-dependent, stable is one instance).
//  * Degenerate splits retry with a new attempt key (the `goto` at :177-179, :203-205, :251-253,
//    :291-293); after 8 failed attempts the node is split at its middle position.
//  * Random-rotation KD splitter (SolutionRandomKD.hpp): normal draws by Box-Muller on the counter
//    stream; node FP64 sums sequential within tiles of 1024 positions, tiles added in order; a split
//    that leaves a side empty (infinite recursion in the reference) cuts the node at its middle.

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kMaxAttempts = 8;
constexpr uint32_t kLargeChunk = 256;
constexpr uint32_t kNoId = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------------------------
// Parallel helper (deterministic work, any schedule).
// ---------------------------------------------------------------------------------------------
static void parallel_for(uint32_t n, uint32_t threads, const std::function<void(uint32_t)>& fn) {
    if (threads <= 1 || n <= 1) {
        for (uint32_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            fn(i);
        }
    };
    uint32_t t = std::min(threads, n);
    std::vector<std::thread> pool;
    pool.reserve(t - 1);
    for (uint32_t i = 1; i < t; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------------------------
// Arithmetic kernels: LinearAlgebra.hpp:53-77 (distance) and :179-200 (dot).
// ---------------------------------------------------------------------------------------------
static inline float dist_ref(const float* a, const float* b, uint32_t d) {
    const uint32_t main = d & ~7u;
    __m256 acc = _mm256_setzero_ps();
    for (uint32_t i = 0; i < main; i += 8) {
        __m256 x = _mm256_loadu_ps(a + i);
        __m256 y = _mm256_loadu_ps(b + i);
        __m256 df = _mm256_sub_ps(x, y);
        acc = _mm256_fmadd_ps(df, df, acc);
    }
    alignas(32) float lanes[8];
    _mm256_store_ps(lanes, acc);
    float ans = 0.0f;
    for (int j = 0; j < 8; ++j) ans = ans + lanes[j];
    for (uint32_t i = main; i < d; ++i) {
        float df = a[i] - b[i];
        ans = std::fmaf(df, df, ans);
    }
    return ans;
}

static inline float dot_ref(const float* a, const float* b, uint32_t d) {
    const uint32_t main = d & ~7u;
    __m256 acc = _mm256_setzero_ps();
    for (uint32_t i = 0; i < main; i += 8) {
        __m256 x = _mm256_loadu_ps(a + i);
        __m256 y = _mm256_loadu_ps(b + i);
        acc = _mm256_fmadd_ps(y, x, acc);
    }
    alignas(32) float lanes[8];
    _mm256_store_ps(lanes, acc);
    float ans = 0.0f;
    for (int j = 0; j < 8; ++j) ans = ans + lanes[j];
    for (uint32_t i = main; i < d; ++i) ans = std::fmaf(a[i], b[i], ans);
    return ans;
}

// ---------------------------------------------------------------------------------------------
// Counter-based random draws (replace rd / uniform_int_distribution, Constants.hpp:22,
// SolutionKmeans.hpp:55-66, :103-110).
// ---------------------------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t round_key(uint64_t seed, uint32_t round) {
    return mix64(seed + 0x9E3779B97F4A7C15ull * (uint64_t)(round + 1));
}
static inline uint64_t child_key(uint64_t key, uint32_t side) { return mix64(2 * key + 1 + side); }
static inline uint64_t attempt_key(uint64_t key, uint32_t attempt) {
    return mix64(key ^ (0xD1B54A32D192ED03ull * (uint64_t)(attempt + 1)));
}
static inline uint32_t draw_below(uint64_t akey, uint32_t i, uint32_t n) {
    uint64_t h = mix64(akey + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1));
    return (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

// calcSamplePercent + getSampleFromPercent size, SolutionKmeans.hpp:50-58: the percentage is a
// float, and `rangeSize * perc` is a float product truncated to uint32.
static uint32_t sample_count(uint32_t n) {
    float perc = (float)(std::pow(std::log((double)n) / std::log(30.0), 7) / (double)n);
    return (uint32_t)((float)n * perc);
}
// kmeansStartVecs sample size, SolutionKmeans.hpp:70 and :105.
static uint32_t start_count(uint32_t n) { return (uint32_t)std::pow(std::log10((double)n), 2.5); }

// ---------------------------------------------------------------------------------------------
// KnnSetScannableSimd restated (KnnSets.hpp:26-127).
// ---------------------------------------------------------------------------------------------
struct KnnSetRef {
    uint32_t k = 0, padded = 0;
    std::vector<float> dists;     // padded to a multiple of 8 with zeros (KnnSets.hpp:28)
    std::vector<uint32_t> ids;    // zero-initialised (KnnSets.hpp:29)
    uint32_t size = 0;
    uint32_t lb_idx = kNoId;

    void init(uint32_t kk) {
        k = kk;
        padded = (kk + 7) & ~7u;
        dists.assign(padded, 0.0f);
        ids.assign(kk, 0u);
        size = 0;
        lb_idx = kNoId;
    }
    bool contains(uint32_t node) const {  // KnnSets.hpp:32-37
        for (uint32_t i = 0; i < size; ++i)
            if (ids[i] == node) return true;
        return false;
    }
    // getMaxIdx, KnnSets.hpp:82-111: lane-wise strict > over blocks of 8, then a cross-lane
    // strict > starting from FLT_MIN. When every value is <= FLT_MIN the reference returns -1 and
    // reads dists[-1]; slot 0 is taken here.
    uint32_t get_max_idx() const {
        float mx[8];
        uint32_t mi[8];
        for (uint32_t l = 0; l < 8; ++l) {
            mx[l] = dists[l];
            mi[l] = l;
        }
        for (uint32_t b = 8; b < padded; b += 8)
            for (uint32_t l = 0; l < 8; ++l)
                if (dists[b + l] > mx[l]) {
                    mx[l] = dists[b + l];
                    mi[l] = b + l;
                }
        float best = FLT_MIN;
        uint32_t idx = kNoId;
        for (uint32_t l = 0; l < 8; ++l)
            if (mx[l] > best) {
                best = mx[l];
                idx = mi[l];
            }
        return idx == kNoId ? 0u : idx;
    }
    // addCandidateLessThan, KnnSets.hpp:39-57 (the caller has already checked dist < bound).
    bool add_less_than(float& bound, uint32_t cand, float dist) {
        if (size < k) {
            if (!contains(cand)) {
                ids[size] = cand;
                dists[size] = dist;
                ++size;
                if (size == k) {
                    lb_idx = get_max_idx();
                    bound = dists[lb_idx];
                }
                return true;
            }
        } else if (!contains(cand)) {  // containsFull, KnnSets.hpp:59-72
            dists[lb_idx] = dist;
            ids[lb_idx] = cand;
            lb_idx = get_max_idx();
            bound = dists[lb_idx];
            return true;
        }
        return false;
    }
    // finalize, KnnSets.hpp:113-126: sort (dist, id) pairs ascending, emit ids.
    void finalize(uint32_t* out_ids, float* out_d) const {
        std::vector<std::pair<float, uint32_t>> q;
        q.reserve(size);
        for (uint32_t i = 0; i < size; ++i) q.emplace_back(dists[i], ids[i]);
        std::sort(q.begin(), q.end());
        for (uint32_t i = 0; i < size; ++i) {
            out_ids[i] = q[i].second;
            if (out_d) out_d[i] = q[i].first;
        }
    }
};

struct Params {
    uint32_t k, leaf_max, leaf_max_odd, small_max, iterations, rounds, non_passes;
    uint64_t seed;
    uint32_t non_rank_lo;
    uint32_t splitter;  // 0: 2-means (SolutionKmeans), 1: random-rotation KD (SolutionRandomKD)
};

struct Points {
    const float* X;
    uint32_t n, d, ld;
    const float* row(uint32_t id) const { return X + (size_t)id * ld; }
};

// ---------------------------------------------------------------------------------------------
// Split: splitKmeansBinaryProcess, SolutionKmeans.hpp:142-311.
// ---------------------------------------------------------------------------------------------
struct Hyperplane {
    std::vector<float> coef;
    float offset;
};

// between = 0.5 * (c1 + c2); coefs = c1 - between; offset = dot(between, coefs)
// (SolutionKmeans.hpp:162-164, :189-191, :219-221, :263-265; LinearAlgebra.hpp:128-161).
static Hyperplane make_hyperplane(const std::vector<float>& c1, const std::vector<float>& c2,
                                  uint32_t d) {
    Hyperplane h;
    std::vector<float> between(d);
    h.coef.resize(d);
    for (uint32_t i = 0; i < d; ++i) {
        float s = c1[i] + c2[i];
        between[i] = 0.5f * s;
    }
    for (uint32_t i = 0; i < d; ++i) h.coef[i] = c1[i] - between[i];
    h.offset = dot_ref(between.data(), h.coef.data(), d);
    return h;
}

// `dot(coefs, pt) >= offset` (SolutionKmeans.hpp:171, :196, :232, :274).
static inline bool side_lo(const Hyperplane& h, const float* pt, uint32_t d) {
    return dot_ref(h.coef.data(), pt, d) >= h.offset;
}

struct SplitStats {
    uint64_t attempts_failed = 0;
    uint64_t fallbacks = 0;
};

// What one node's split decided, for the known-answer tests of the split: the attempt that
// succeeded, its start sample and farthest pair, the centres the last Lloyd step produced, the
// hyperplane of the final side test and the lo size. Rows are stored only for d <= kTraceDims.
constexpr uint32_t kTraceDims = 512;
constexpr uint32_t kTraceRows = 256;
struct SplitTrace {
    uint32_t attempt, s, bi, bj, m, nlo, fallback, pad_;
    uint32_t rows[kTraceRows];
    float c1[kTraceDims], c2[kTraceDims], coef[kTraceDims];
    float offset;
};

// kmeansStartVecs, SolutionKmeans.hpp:68-101 / :103-138: farthest pair among the s sampled rows,
// first maximum in (i, j) order, starting from FLT_MIN with strict >. Returns false when no pair
// exceeds FLT_MIN (the reference dereferences null there).
static bool farthest_pair(const Points& P, const std::vector<uint32_t>& rows, uint32_t& bi,
                          uint32_t& bj) {
    float best = FLT_MIN;
    bool found = false;
    const uint32_t s = (uint32_t)rows.size();
    for (uint32_t i = 0; i + 1 < s; ++i)
        for (uint32_t j = i + 1; j < s; ++j) {
            float dd = dist_ref(P.row(rows[i]), P.row(rows[j]), P.d);
            if (dd > best) {
                best = dd;
                bi = i;
                bj = j;
                found = true;
            }
        }
    return found;
}

// Splits perm[start, start+n). Returns the size of the lo child (points with dot >= offset,
// placed first). `threads` parallelises the full-range side test and the sample sums.
static uint32_t split_node(const Points& P, const Params& prm, uint32_t* perm, uint32_t start,
                           uint32_t n, uint64_t key, uint32_t threads, SplitStats& st,
                           const Hyperplane* injected, SplitTrace* tr = nullptr) {
    const uint32_t d = P.d;
    const bool large = n >= prm.small_max;
    std::vector<uint8_t> flags(n);
    for (uint32_t attempt = 0; attempt < kMaxAttempts; ++attempt) {
        Hyperplane h;
        if (injected) {
            h = *injected;
        } else {
            const uint64_t ak = attempt_key(key, attempt);
            const uint32_t m = large ? sample_count(n) : 0;
            uint32_t s = start_count(n);
            if (large) s = std::min(m, s);
            if (s < 2) s = 2;
            std::vector<uint32_t> rows(s);
            for (uint32_t i = 0; i < s; ++i) rows[i] = perm[start + draw_below(ak, i, n)];
            uint32_t bi = 0, bj = 0;
            if (!farthest_pair(P, rows, bi, bj)) {
                st.attempts_failed++;
                continue;
            }
            if (tr) {
                tr->s = s;
                tr->m = m;
                tr->bi = bi;
                tr->bj = bj;
                for (uint32_t i = 0; i < s && i < kTraceRows; ++i) tr->rows[i] = rows[i];
            }
            std::vector<float> c1(P.row(rows[bi]), P.row(rows[bi]) + d);
            std::vector<float> c2(P.row(rows[bj]), P.row(rows[bj]) + d);
            bool failed = false;
            for (uint32_t it = 0; it < prm.iterations && !failed; ++it) {
                Hyperplane hh = make_hyperplane(c1, c2, d);
                std::vector<double> s1(d, 0.0), s2(d, 0.0);
                uint64_t n1 = 0, n2 = 0;
                if (large) {
                    // Sample-only pass (SolutionKmeans.hpp:218-249), chunked FP64 sums.
                    const uint32_t chunks = (m + kLargeChunk - 1) / kLargeChunk;
                    std::vector<double> p1((size_t)chunks * d, 0.0), p2((size_t)chunks * d, 0.0);
                    std::vector<uint64_t> q1(chunks, 0), q2(chunks, 0);
                    parallel_for(chunks, threads, [&](uint32_t c) {
                        double* a1 = &p1[(size_t)c * d];
                        double* a2 = &p2[(size_t)c * d];
                        const uint32_t e = std::min(m, (c + 1) * kLargeChunk);
                        for (uint32_t i = c * kLargeChunk; i < e; ++i) {
                            const float* pt = P.row(perm[start + draw_below(ak, i, n)]);
                            if (side_lo(hh, pt, d)) {
                                q1[c]++;
                                for (uint32_t j = 0; j < d; ++j) a1[j] += (double)pt[j];
                            } else {
                                q2[c]++;
                                for (uint32_t j = 0; j < d; ++j) a2[j] += (double)pt[j];
                            }
                        }
                    });
                    for (uint32_t c = 0; c < chunks; ++c) {
                        n1 += q1[c];
                        n2 += q2[c];
                        for (uint32_t j = 0; j < d; ++j) {
                            s1[j] += p1[(size_t)c * d + j];
                            s2[j] += p2[(size_t)c * d + j];
                        }
                    }
                } else {
                    // Full-range pass (SolutionKmeans.hpp:166-175), sequential FP64 sums.
                    for (uint32_t i = start; i < start + n; ++i) {
                        const float* pt = P.row(perm[i]);
                        if (side_lo(hh, pt, d)) {
                            n1++;
                            for (uint32_t j = 0; j < d; ++j) s1[j] += (double)pt[j];
                        } else {
                            n2++;
                            for (uint32_t j = 0; j < d; ++j) s2[j] += (double)pt[j];
                        }
                    }
                }
                if (n1 == 0 || n2 == 0) {
                    failed = true;
                    break;
                }
                // SolutionKmeans.hpp:182-185: double quotient cast to float.
                for (uint32_t j = 0; j < d; ++j) {
                    c1[j] = (float)(s1[j] / (double)n1);
                    c2[j] = (float)(s2[j] / (double)n2);
                }
            }
            if (failed) {
                st.attempts_failed++;
                continue;
            }
            if (tr && d <= kTraceDims) {
                std::memcpy(tr->c1, c1.data(), d * sizeof(float));
                std::memcpy(tr->c2, c2.data(), d * sizeof(float));
            }
            h = make_hyperplane(c1, c2, d);
        }
        // Final side test over the full range (SolutionKmeans.hpp:189-197, :262-289).
        const uint32_t blocks = (n + 4095) / 4096;
        parallel_for(large ? blocks : 1, large ? threads : 1, [&](uint32_t b) {
            const uint32_t lo = large ? b * 4096 : 0;
            const uint32_t hi = large ? std::min(n, lo + 4096) : n;
            for (uint32_t i = lo; i < hi; ++i) flags[i] = side_lo(h, P.row(perm[start + i]), d);
        });
        uint32_t nlo = 0;
        for (uint32_t i = 0; i < n; ++i) nlo += flags[i];
        if (nlo == 0 || nlo == n) {
            if (injected) return nlo;
            st.attempts_failed++;
            continue;
        }
        if (tr) {
            tr->attempt = attempt;
            tr->nlo = nlo;
            tr->fallback = 0;
            tr->offset = h.offset;
            if (d <= kTraceDims) std::memcpy(tr->coef, h.coef.data(), d * sizeof(float));
        }
        std::vector<uint32_t> tmp(perm + start, perm + start + n);
        uint32_t a = 0, b = nlo;
        for (uint32_t i = 0; i < n; ++i) {
            if (flags[i]) perm[start + a++] = tmp[i];
            else perm[start + b++] = tmp[i];
        }
        return nlo;
    }
    st.fallbacks++;
    if (tr) {
        tr->attempt = kMaxAttempts;
        tr->nlo = n / 2;
        tr->fallback = 1;
    }
    return n / 2;
}

// ---------------------------------------------------------------------------------------------
// Random-rotation KD splitter, SolutionRandomKD.hpp:258-428.
// ---------------------------------------------------------------------------------------------
constexpr uint32_t kKdDims = 30;    // newDims, SolutionRandomKD.hpp:387
constexpr uint32_t kKdTile = 1024;  // FP64 sums: sequential within tiles, tiles added in order

// N(0, 1) by Box-Muller on two counter draws (in place of std::normal_distribution over rd,
// LinearAlgebra.hpp:104).
static double kd_normal(uint64_t key, uint64_t i) {
    const uint64_t a = mix64(key + 0x9E3779B97F4A7C15ull * (2 * i + 1));
    const uint64_t b = mix64(key + 0x9E3779B97F4A7C15ull * (2 * i + 2));
    const double u1 = ((double)(a >> 11) + 1.0) * 0x1p-53;
    const double u2 = (double)(b >> 11) * 0x1p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// rotateRandom's directions (SolutionRandomKD.hpp:258-264): randUniformUnitVec
// (LinearAlgebra.hpp:103-116; norm's float sum of squares, :79-85; normalizeInPlace's division by
// its double root, :87-92), then gramSchmidt (:211-222) as written: normalize (:94-101) returns its
// argument, so u_0 = v_0 and u_i = v_i - sum_{j<i} dot(u_j, v_i) * u_j (project, :204-206; sub and
// scalarMult, :146-161), with no renormalization.
static std::vector<float> kd_directions(uint64_t seed, uint32_t round, uint32_t m, uint32_t d) {
    const uint64_t key = mix64(round_key(seed, round) ^ 0x6A09E667F3BCC909ull);
    std::vector<float> v((size_t)m * d), u((size_t)m * d);
    for (uint32_t c = 0; c < m; ++c) {
        float sum_sq = 0.0f;
        for (uint32_t j = 0; j < d; ++j) {
            const float x = (float)kd_normal(key, (uint64_t)c * d + j);
            v[(size_t)c * d + j] = x;
            sum_sq += x * x;
        }
        const double norm = std::sqrt((double)sum_sq);
        for (uint32_t j = 0; j < d; ++j) v[(size_t)c * d + j] = (float)(v[(size_t)c * d + j] / norm);
    }
    for (uint32_t i = 0; i < m; ++i) {
        for (uint32_t e = 0; e < d; ++e) u[(size_t)i * d + e] = v[(size_t)i * d + e];
        for (uint32_t j = 0; j < i; ++j) {
            const float c = dot_ref(&u[(size_t)j * d], &v[(size_t)i * d], d);
            for (uint32_t e = 0; e < d; ++e) u[(size_t)i * d + e] -= c * u[(size_t)j * d + e];
        }
    }
    return u;
}

// pickIndexFull's draw (SolutionRandomKD.hpp:223-248) given the node's variances.
static uint32_t kd_pick_dim(const std::vector<double>& var, uint64_t node_key) {
    const uint32_t m = (uint32_t)var.size();
    float total_var = 0;
    for (uint32_t j = 0; j < m; ++j) total_var += var[j];
    std::vector<std::pair<double, uint32_t>> by_var;
    for (uint32_t j = 0; j < m; ++j) by_var.emplace_back(var[j], j);
    std::sort(by_var.begin(), by_var.end(), std::greater<>{});
    const float u01 = (float)(mix64(node_key + 0x9E3779B97F4A7C15ull) >> 40) * 0x1p-24f;
    const float sample_var = u01 * total_var;
    float var_so_far = 0;
    for (auto& [vj, j] : by_var) {
        var_so_far += vj;
        if (var_so_far >= sample_var) return j;
    }
    return by_var.back().second;  // the reference's __builtin_unreachable
}

// One node (split, SolutionRandomKD.hpp:282-373): FP64 means and sample variances of the
// projections over the range (getMean / getVariance, :104-151), a dimension by pickIndexFull, the
// mean as split value; x < mean goes first (:338-339), stable. Returns the first group's size.
static uint32_t kd_split_node(const std::vector<float>& rot, uint32_t m, uint32_t* perm,
                              uint32_t start, uint32_t n, uint64_t key, SplitStats& st) {
    std::vector<double> mean(m), var(m);
    for (uint32_t c = 0; c < m; ++c) {
        double total = 0.0;
        for (uint32_t t0 = 0; t0 < n; t0 += kKdTile) {
            double s = 0.0;
            for (uint32_t e = t0; e < std::min(n, t0 + kKdTile); ++e)
                s += (double)rot[(size_t)perm[start + e] * m + c];
            total += s;
        }
        mean[c] = total / (double)n;
    }
    for (uint32_t c = 0; c < m; ++c) {
        double total = 0.0;
        for (uint32_t t0 = 0; t0 < n; t0 += kKdTile) {
            double s = 0.0;
            for (uint32_t e = t0; e < std::min(n, t0 + kKdTile); ++e) {
                const double diff = (double)rot[(size_t)perm[start + e] * m + c] - mean[c];
                s += diff * diff;
            }
            total += s;
        }
        var[c] = total / (double)(n - 1);
    }
    const uint32_t dim = kd_pick_dim(var, key);
    const double split_value = mean[dim];
    std::vector<uint32_t> g1, g2;
    for (uint32_t e = 0; e < n; ++e) {
        const uint32_t id = perm[start + e];
        ((double)rot[(size_t)id * m + dim] < split_value ? g1 : g2).push_back(id);
    }
    if (g1.empty() || g2.empty()) {  // the reference would recurse on the same range forever
        st.attempts_failed++;
        st.fallbacks++;
        return n / 2;
    }
    std::memcpy(perm + start, g1.data(), g1.size() * sizeof(uint32_t));
    std::memcpy(perm + start + g1.size(), g2.data(), g2.size() * sizeof(uint32_t));
    return (uint32_t)g1.size();
}

struct Node {
    uint32_t start, n;
    uint64_t key;
};
struct Leaf {
    uint32_t start, n;
};

// One round of splitting from the identity permutation (SolutionKmeans.hpp:589-591).
// Level-synchronous: each level's nodes are independent, so the schedule does not change results.
static void split_round(const Points& P, const Params& prm, uint32_t round, uint32_t* perm,
                        std::vector<Leaf>& leaves, uint32_t threads, SplitStats& st) {
    const uint32_t n = P.n;
    const uint32_t leaf_max = (round % 2 == 1 && prm.leaf_max_odd) ? prm.leaf_max_odd : prm.leaf_max;
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    leaves.clear();
    std::vector<Node> level;
    if (n <= leaf_max) {
        if (n > 0) leaves.push_back({0, n});
        return;
    }
    level.push_back({0, n, round_key(prm.seed, round)});
    if (prm.splitter == 1) {  // SolutionRandomKD: project (rotateRandom, :258-274), then split
        const uint32_t m = std::min(kKdDims, P.d);
        const std::vector<float> u = kd_directions(prm.seed, round, m, P.d);
        std::vector<float> rot((size_t)n * m);
        parallel_for(n, threads, [&](uint32_t r) {
            for (uint32_t c = 0; c < m; ++c) rot[(size_t)r * m + c] = dot_ref(&u[(size_t)c * P.d], P.row(r), P.d);
        });
        while (!level.empty()) {
            std::vector<uint32_t> nlo(level.size());
            std::vector<SplitStats> sts(level.size());
            parallel_for((uint32_t)level.size(), threads, [&](uint32_t i) {
                nlo[i] = kd_split_node(rot, m, perm, level[i].start, level[i].n, level[i].key, sts[i]);
            });
            std::vector<Node> next;
            for (uint32_t i = 0; i < level.size(); ++i) {
                st.attempts_failed += sts[i].attempts_failed;
                st.fallbacks += sts[i].fallbacks;
                const Node& nd = level[i];
                Node ch[2] = {{nd.start, nlo[i], child_key(nd.key, 0)},
                              {nd.start + nlo[i], nd.n - nlo[i], child_key(nd.key, 1)}};
                for (auto& c : ch) {
                    if (c.n <= leaf_max) leaves.push_back({c.start, c.n});
                    else next.push_back(c);
                }
            }
            level.swap(next);
        }
        std::sort(leaves.begin(), leaves.end(), [](const Leaf& a, const Leaf& b) { return a.start < b.start; });
        return;
    }
    while (!level.empty()) {
        std::vector<uint32_t> nlo(level.size());
        std::vector<SplitStats> sts(level.size());
        // Large nodes one at a time with internal parallelism, then small nodes across threads.
        for (uint32_t i = 0; i < level.size(); ++i)
            if (level[i].n >= prm.small_max)
                nlo[i] = split_node(P, prm, perm, level[i].start, level[i].n, level[i].key, threads,
                                    sts[i], nullptr);
        parallel_for((uint32_t)level.size(), threads, [&](uint32_t i) {
            if (level[i].n < prm.small_max)
                nlo[i] = split_node(P, prm, perm, level[i].start, level[i].n, level[i].key, 1,
                                    sts[i], nullptr);
        });
        std::vector<Node> next;
        for (uint32_t i = 0; i < level.size(); ++i) {
            st.attempts_failed += sts[i].attempts_failed;
            st.fallbacks += sts[i].fallbacks;
            const Node& nd = level[i];
            Node ch[2] = {{nd.start, nlo[i], child_key(nd.key, 0)},
                          {nd.start + nlo[i], nd.n - nlo[i], child_key(nd.key, 1)}};
            for (auto& c : ch) {
                if (c.n <= leaf_max) leaves.push_back({c.start, c.n});
                else next.push_back(c);
            }
        }
        level.swap(next);
    }
    std::sort(leaves.begin(), leaves.end(), [](const Leaf& a, const Leaf& b) { return a.start < b.start; });
}

// addCandidatesLessThan, KnnSets.hpp:131-160.
static void leaf_all_pairs(const Points& P, const uint32_t* perm, const Leaf& lf,
                           std::vector<KnnSetRef>& knn, std::vector<float>& bounds) {
    const uint32_t d = P.d;
    std::vector<float> copy((size_t)lf.n * d);
    for (uint32_t i = 0; i < lf.n; ++i)
        std::memcpy(&copy[(size_t)i * d], P.row(perm[lf.start + i]), d * sizeof(float));
    for (uint32_t i = 0; i + 1 < lf.n; ++i) {
        const uint32_t id1 = perm[lf.start + i];
        for (uint32_t j = i + 1; j < lf.n; ++j) {
            float dist = dist_ref(&copy[(size_t)i * d], &copy[(size_t)j * d], d);
            const uint32_t id2 = perm[lf.start + j];
            if (dist < bounds[id1]) knn[id1].add_less_than(bounds[id1], id2, dist);
            if (dist < bounds[id2]) knn[id2].add_less_than(bounds[id2], id1, dist);
        }
    }
}

// topUpSingle, SolutionKmeans.hpp:480-555. The snapshot holds all k slots of each set; slots past
// `size` are still zero in the reference (KnnSets.hpp:29), so they read as id 0. Candidates are
// offered in ascending id order (the reference's robin_set order is unspecified).
static void non_pass_ordered(const Points& P, std::vector<KnnSetRef>& knn,
                             std::vector<float>& bounds, uint32_t k, uint32_t threads,
                             uint32_t rank_lo);

static void non_pass(const Points& P, std::vector<KnnSetRef>& knn, std::vector<float>& bounds,
                     uint32_t k, uint32_t threads, uint32_t rank_lo = 0) {
    if (rank_lo) {
        non_pass_ordered(P, knn, bounds, k, threads, rank_lo);
        return;
    }
    const uint32_t n = P.n;
    std::vector<uint32_t> snap((size_t)n * k);
    parallel_for(n, threads, [&](uint32_t id) {
        for (uint32_t s = 0; s < k; ++s) snap[(size_t)id * k + s] = knn[id].ids[s];
        std::sort(&snap[(size_t)id * k], &snap[(size_t)id * k] + k);
    });
    const uint32_t chunk = 256;
    parallel_for((n + chunk - 1) / chunk, threads, [&](uint32_t c) {
        std::vector<uint32_t> cand;
        cand.reserve((size_t)k * k);
        for (uint32_t id1 = c * chunk; id1 < std::min(n, (c + 1) * chunk); ++id1) {
            cand.clear();
            for (uint32_t a = 0; a < k; ++a) {
                const uint32_t id2 = snap[(size_t)id1 * k + a];
                for (uint32_t b = 0; b < k; ++b) cand.push_back(snap[(size_t)id2 * k + b]);
            }
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            for (uint32_t id3 : cand) {
                if (id3 == id1) continue;
                float dist = dist_ref(P.row(id3), P.row(id1), P.d);
                if (dist < bounds[id1]) knn[id1].add_less_than(bounds[id1], id3, dist);
            }
        }
    });
}

// topUpSingleOrdered, SolutionKmeans.hpp:393-477: each set's filled slots sorted by (dist, id)
// (:398-419); point id1 expands only its neighbours at ranks [rank_lo, size) (the reference uses
// 50, :437), each contributing the first min(size, 100) of its own sorted neighbours (:439-442);
// id1 itself is removed (:447), current neighbours are not. Offers in ascending id order.
static void non_pass_ordered(const Points& P, std::vector<KnnSetRef>& knn,
                             std::vector<float>& bounds, uint32_t k, uint32_t threads,
                             uint32_t rank_lo) {
    const uint32_t n = P.n;
    std::vector<std::vector<uint32_t>> sorted(n);
    parallel_for(n, threads, [&](uint32_t id) {
        std::vector<std::pair<float, uint32_t>> di;
        for (uint32_t s = 0; s < knn[id].size; ++s) di.emplace_back(knn[id].dists[s], knn[id].ids[s]);
        std::sort(di.begin(), di.end());
        for (auto& e : di) sorted[id].push_back(e.second);
    });
    const uint32_t chunk = 256;
    parallel_for((n + chunk - 1) / chunk, threads, [&](uint32_t c) {
        std::vector<uint32_t> cand;
        for (uint32_t id1 = c * chunk; id1 < std::min(n, (c + 1) * chunk); ++id1) {
            cand.clear();
            const auto& k1 = sorted[id1];
            for (uint32_t i2 = rank_lo; i2 < k1.size(); ++i2) {
                const auto& k2 = sorted[k1[i2]];
                for (uint32_t i3 = 0; i3 < k2.size() && i3 < 100; ++i3) cand.push_back(k2[i3]);
            }
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            for (uint32_t id3 : cand) {
                if (id3 == id1) continue;
                float dist = dist_ref(P.row(id3), P.row(id1), P.d);
                if (dist < bounds[id1]) knn[id1].add_less_than(bounds[id1], id3, dist);
            }
        }
    });
    (void)k;
}

static void load_state(std::vector<KnnSetRef>& knn, std::vector<float>& bounds, uint32_t n,
                       uint32_t k, const uint32_t* ids, const float* dists, const uint32_t* sizes) {
    knn.resize(n);
    bounds.assign(n, FLT_MAX);
    for (uint32_t p = 0; p < n; ++p) {
        knn[p].init(k);
        if (!ids) continue;
        const uint32_t sz = sizes[p];
        for (uint32_t s = 0; s < sz; ++s) {
            knn[p].ids[s] = ids[(size_t)p * k + s];
            knn[p].dists[s] = dists[(size_t)p * k + s];
        }
        knn[p].size = sz;
        if (sz == k) {
            knn[p].lb_idx = knn[p].get_max_idx();
            bounds[p] = knn[p].dists[knn[p].lb_idx];
        }
    }
}

// State export: each point's set in (dist, id) order, padded with (FLT_MAX, 0xFFFFFFFF).
static void store_state(const std::vector<KnnSetRef>& knn, uint32_t n, uint32_t k, uint32_t* ids,
                        float* dists, uint32_t* sizes) {
    for (uint32_t p = 0; p < n; ++p) {
        knn[p].finalize(ids + (size_t)p * k, dists + (size_t)p * k);
        for (uint32_t s = knn[p].size; s < k; ++s) {
            ids[(size_t)p * k + s] = kNoId;
            dists[(size_t)p * k + s] = FLT_MAX;
        }
        if (sizes) sizes[p] = knn[p].size;
    }
}

static uint32_t pick_threads(int t) {
    if (t > 0) return (uint32_t)t;
    uint32_t h = std::thread::hardware_concurrency();
    return h ? h : 1;
}

}  // namespace

extern "C" {

struct orc_params {
    uint32_t k;
    uint32_t leaf_max;      // largest leaf bucket, inclusive (reference maxGroupSize - 1)
    uint32_t leaf_max_odd;  // leaf_max on odd rounds, 0 = same (reference alternates 199/399)
    uint32_t small_max;     // nodes below this take the serial small branch (reference 3000)
    uint32_t iterations;    // Lloyd iterations per split (reference knnIterations)
    uint32_t rounds;
    uint32_t non_passes;
    uint32_t non_rank_lo;   // 0: topUpSingle; r > 0: topUpSingleOrdered expanding ranks >= r
    uint32_t splitter;      // 0: 2-means (SolutionKmeans); 1: random-rotation KD (SolutionRandomKD)
    uint32_t pad_;
    uint64_t seed;
};

struct orc_stats {
    double split_s, leaf_s, non_s, finalize_s;
    uint64_t attempts_failed, fallbacks, leaves;
};

float orc_dist(const float* a, const float* b, uint32_t d) { return dist_ref(a, b, d); }
// the KD splitter's directions for (seed, round): m rows of d floats
void orc_kd_directions(uint64_t seed, uint32_t round, uint32_t m, uint32_t d, float* out) {
    const std::vector<float> u = kd_directions(seed, round, m, d);
    std::memcpy(out, u.data(), u.size() * sizeof(float));
}
double orc_kd_normal(uint64_t key, uint64_t i) { return kd_normal(key, i); }
uint32_t orc_kd_pick_dim(const double* var, uint32_t m, uint64_t node_key) {
    return kd_pick_dim(std::vector<double>(var, var + m), node_key);
}
float orc_dot(const float* a, const float* b, uint32_t d) { return dot_ref(a, b, d); }
uint32_t orc_sample_count(uint32_t n) { return sample_count(n); }
uint32_t orc_start_count(uint32_t n) { return start_count(n); }
uint64_t orc_mix64(uint64_t z) { return mix64(z); }
uint32_t orc_draw_below(uint64_t akey, uint32_t i, uint32_t n) { return draw_below(akey, i, n); }
uint64_t orc_round_key(uint64_t seed, uint32_t round) { return round_key(seed, round); }
uint64_t orc_child_key(uint64_t key, uint32_t side) { return child_key(key, side); }
uint64_t orc_attempt_key(uint64_t key, uint32_t attempt) { return attempt_key(key, attempt); }

// --- KnnSetScannableSimd handle, for the reference's known-answer tests -----------------------
void* orc_knnset_new(uint32_t k) {
    auto* s = new KnnSetRef();
    s->init(k);
    return s;
}
void orc_knnset_free(void* h) { delete static_cast<KnnSetRef*>(h); }
// Caller-side filter `if (dist < bound)` (KnnSets.hpp:152-157) plus addCandidateLessThan.
// The bound lives beside the set in the reference (`bounds[id]`); it is kept in slot `bound`.
int orc_knnset_offer(void* h, float* bound, uint32_t id, float dist) {
    auto* s = static_cast<KnnSetRef*>(h);
    if (!(dist < *bound)) return 0;
    return s->add_less_than(*bound, id, dist) ? 1 : 0;
}
int orc_knnset_contains(void* h, uint32_t id) { return static_cast<KnnSetRef*>(h)->contains(id); }
uint32_t orc_knnset_size(void* h) { return static_cast<KnnSetRef*>(h)->size; }
uint32_t orc_knnset_max_idx(void* h) { return static_cast<KnnSetRef*>(h)->get_max_idx(); }
void orc_knnset_set_slot(void* h, uint32_t slot, uint32_t id, float dist) {
    auto* s = static_cast<KnnSetRef*>(h);
    s->ids[slot] = id;
    s->dists[slot] = dist;
}
void orc_knnset_finalize(void* h, uint32_t* out_ids, float* out_d) {
    static_cast<KnnSetRef*>(h)->finalize(out_ids, out_d);
}

// contains(vector, id), KnnSets.hpp:363-379.
int orc_contains(const uint32_t* ids, uint32_t n, uint32_t cand) {
    for (uint32_t i = 0; i < n; ++i)
        if (ids[i] == cand) return 1;
    return 0;
}

// --- Split ----------------------------------------------------------------------------------
// kmeansStartVecs' pair search (SolutionKmeans.hpp:76-99) over the rows X[rows[0..s)]. Returns 0
// when no pair lies beyond FLT_MIN.
int orc_farthest_pair(const float* X, uint32_t n, uint32_t d, uint32_t ld, const uint32_t* rows,
                      uint32_t s, uint32_t* bi, uint32_t* bj) {
    Points P{X, n, d, ld};
    std::vector<uint32_t> r(rows, rows + s);
    *bi = 0;
    *bj = 0;
    return farthest_pair(P, r, *bi, *bj) ? 1 : 0;
}

// The root node of round `round` split once (splitKmeansBinaryProcess, SolutionKmeans.hpp:142-311)
// from the identity permutation, recording what it decided. perm_out receives the partitioned
// permutation. The node's branch (small or large) follows prm->small_max as in split_round.
int orc_split_root_trace(const float* X, uint32_t n, uint32_t d, uint32_t ld,
                         const orc_params* prm, uint32_t round, uint32_t* perm_out,
                         SplitTrace* trace, int threads) {
    if (d > kTraceDims || n < 2) return -1;
    Points P{X, n, d, ld};
    Params p{prm->k, prm->leaf_max, prm->leaf_max_odd, prm->small_max, prm->iterations,
             prm->rounds, prm->non_passes, prm->seed, prm->non_rank_lo, prm->splitter};
    for (uint32_t i = 0; i < n; ++i) perm_out[i] = i;
    std::memset(trace, 0, sizeof(SplitTrace));
    SplitStats st;
    split_node(P, p, perm_out, 0, n, round_key(p.seed, round), pick_threads(threads), st, nullptr,
               trace);
    return 0;
}

int orc_split_round(const float* X, uint32_t n, uint32_t d, uint32_t ld, const orc_params* prm,
                    uint32_t round, uint32_t* perm_out, uint32_t* leaf_start, uint32_t* leaf_size,
                    uint32_t* n_leaves, int threads) {
    Points P{X, n, d, ld};
    Params p{prm->k, prm->leaf_max, prm->leaf_max_odd, prm->small_max, prm->iterations,
             prm->rounds, prm->non_passes, prm->seed, prm->non_rank_lo, prm->splitter};
    std::vector<Leaf> leaves;
    SplitStats st;
    split_round(P, p, round, perm_out, leaves, pick_threads(threads), st);
    for (size_t i = 0; i < leaves.size(); ++i) {
        leaf_start[i] = leaves[i].start;
        leaf_size[i] = leaves[i].n;
    }
    *n_leaves = (uint32_t)leaves.size();
    return 0;
}

// One stable side-test partition per node with a caller-supplied hyperplane (coef row per node,
// stride d) over perm (in place). Writes the lo size per node.
int orc_partition_injected(const float* X, uint32_t n, uint32_t d, uint32_t ld, uint32_t* perm,
                           uint32_t n_nodes, const uint32_t* node_start, const uint32_t* node_size,
                           const float* coefs, const float* offsets, uint32_t* lo_out) {
    Points P{X, n, d, ld};
    Params p{};
    p.small_max = 0xFFFFFFFFu;
    SplitStats st;
    for (uint32_t i = 0; i < n_nodes; ++i) {
        Hyperplane h;
        h.coef.assign(coefs + (size_t)i * d, coefs + (size_t)(i + 1) * d);
        h.offset = offsets[i];
        lo_out[i] = split_node(P, p, perm, node_start[i], node_size[i], 0, 1, st, &h);
    }
    return 0;
}

// Hyperplane from two centres (for tests of the algebra).
float orc_hyperplane(const float* c1, const float* c2, uint32_t d, float* coef_out) {
    std::vector<float> a(c1, c1 + d), b(c2, c2 + d);
    Hyperplane h = make_hyperplane(a, b, d);
    std::memcpy(coef_out, h.coef.data(), d * sizeof(float));
    return h.offset;
}

// --- Leaf stage and NoN on an explicit state --------------------------------------------------
// State arrays are n*k ids/dists plus n sizes; on input, ids == nullptr means empty sets.
int orc_leaf_stage(const float* X, uint32_t n, uint32_t d, uint32_t ld, uint32_t k,
                   const uint32_t* perm, const uint32_t* leaf_start, const uint32_t* leaf_size,
                   uint32_t n_leaves, const uint32_t* ids_in, const float* dists_in,
                   const uint32_t* sizes_in, uint32_t* ids_out, float* dists_out,
                   uint32_t* sizes_out, int threads) {
    Points P{X, n, d, ld};
    std::vector<KnnSetRef> knn;
    std::vector<float> bounds;
    load_state(knn, bounds, n, k, ids_in, dists_in, sizes_in);
    parallel_for(n_leaves, pick_threads(threads), [&](uint32_t i) {
        leaf_all_pairs(P, perm, Leaf{leaf_start[i], leaf_size[i]}, knn, bounds);
    });
    store_state(knn, n, k, ids_out, dists_out, sizes_out);
    return 0;
}

int orc_non(const float* X, uint32_t n, uint32_t d, uint32_t ld, uint32_t k,
            const uint32_t* ids_in, const float* dists_in, const uint32_t* sizes_in,
            uint32_t* ids_out, float* dists_out, uint32_t* sizes_out, int threads,
            uint32_t rank_lo) {
    Points P{X, n, d, ld};
    std::vector<KnnSetRef> knn;
    std::vector<float> bounds;
    load_state(knn, bounds, n, k, ids_in, dists_in, sizes_in);
    // The reference's snapshot reads slots in set order; here they come from the caller's order.
    non_pass(P, knn, bounds, k, pick_threads(threads), rank_lo);
    store_state(knn, n, k, ids_out, dists_out, sizes_out);
    return 0;
}

// --- Whole build: constructResult, SolutionKmeans.hpp:558-616 ---------------------------------
// out_ids: n*k, padded with id 1 (padResult, KnnSets.hpp:334-344). out_dists (nullable): the
// matching distances, FLT_MAX for padding.
int orc_build(const float* X, uint32_t n, uint32_t d, uint32_t ld, const orc_params* prm,
              uint32_t* out_ids, float* out_dists, uint32_t* out_sizes, int threads,
              orc_stats* stats) {
    Points P{X, n, d, ld};
    Params p{prm->k, prm->leaf_max, prm->leaf_max_odd, prm->small_max, prm->iterations,
             prm->rounds, prm->non_passes, prm->seed, prm->non_rank_lo, prm->splitter};
    const uint32_t T = pick_threads(threads);
    std::vector<KnnSetRef> knn(n);
    for (auto& s : knn) s.init(p.k);
    std::vector<float> bounds(n, FLT_MAX);
    std::vector<uint32_t> perm(n);
    std::vector<Leaf> leaves;
    SplitStats st;
    orc_stats local{};
    using clk = std::chrono::steady_clock;
    for (uint32_t r = 0; r < p.rounds; ++r) {
        auto t0 = clk::now();
        split_round(P, p, r, perm.data(), leaves, T, st);
        auto t1 = clk::now();
        parallel_for((uint32_t)leaves.size(), T,
                     [&](uint32_t i) { leaf_all_pairs(P, perm.data(), leaves[i], knn, bounds); });
        auto t2 = clk::now();
        local.split_s += std::chrono::duration<double>(t1 - t0).count();
        local.leaf_s += std::chrono::duration<double>(t2 - t1).count();
        local.leaves += leaves.size();
    }
    auto t3 = clk::now();
    for (uint32_t pass = 0; pass < p.non_passes; ++pass) non_pass(P, knn, bounds, p.k, T, p.non_rank_lo);
    auto t4 = clk::now();
    std::vector<float> dtmp;
    float* dout = out_dists;
    if (!dout) {
        dtmp.resize((size_t)n * p.k);
        dout = dtmp.data();
    }
    store_state(knn, n, p.k, out_ids, dout, out_sizes);
    for (size_t i = 0; i < (size_t)n * p.k; ++i)
        if (out_ids[i] == kNoId) out_ids[i] = 1;
    auto t5 = clk::now();
    local.non_s = std::chrono::duration<double>(t4 - t3).count();
    local.finalize_s = std::chrono::duration<double>(t5 - t4).count();
    local.attempts_failed = st.attempts_failed;
    local.fallbacks = st.fallbacks;
    if (stats) *stats = local;
    return 0;
}

// --- Exact FP64 brute force (the recall oracle) -----------------------------------------------
// For each query id: the k nearest other points by FP64 squared L2, ascending (dist, id).
int orc_exact_knn(const float* X, uint32_t n, uint32_t d, uint32_t ld, const uint32_t* queries,
                  uint32_t nq, uint32_t k, uint32_t* out_ids, double* out_dists, int threads) {
    Points P{X, n, d, ld};
    parallel_for(nq, pick_threads(threads), [&](uint32_t qi) {
        const uint32_t q = queries[qi];
        const float* a = P.row(q);
        std::vector<std::pair<double, uint32_t>> heap;  // max-heap of the best k
        heap.reserve(k + 1);
        for (uint32_t r = 0; r < n; ++r) {
            if (r == q) continue;
            const float* b = P.row(r);
            double s = 0.0;
            for (uint32_t j = 0; j < d; ++j) {
                double df = (double)a[j] - (double)b[j];
                s += df * df;
            }
            std::pair<double, uint32_t> e{s, r};
            if (heap.size() < k) {
                heap.push_back(e);
                std::push_heap(heap.begin(), heap.end());
            } else if (e < heap.front()) {
                std::pop_heap(heap.begin(), heap.end());
                heap.back() = e;
                std::push_heap(heap.begin(), heap.end());
            }
        }
        std::sort(heap.begin(), heap.end());
        for (uint32_t i = 0; i < k; ++i) {
            out_ids[(size_t)qi * k + i] = i < heap.size() ? heap[i].second : kNoId;
            out_dists[(size_t)qi * k + i] = i < heap.size() ? heap[i].first : INFINITY;
        }
    });
    return 0;
}

}  // extern "C"
