/*
 * C ABI of the MI355X (gfx950) approximate KNN-graph builder.
 *
 * The drop-in boundary is the reference's
 *   static void SolutionKmeans::constructResult(float points[][112], uint32_t numPoints,
 *                                               std::vector<std::vector<uint32_t>>& result)
 * (src/SolutionKmeans.hpp:558, called from src/main.cpp:59). knng_build() replaces it; the C++
 * header knng_solution.hpp wraps it back into that exact signature.
 *
 * All pointers are plain host or device pointers with sizes; no library-specific types cross the
 * boundary except the opaque handle and the parameter/statistics structs below. Every function
 * returns KNNG_OK (0) or a negative status; knng_last_error() describes the last failure on the
 * calling thread. A handle is not reentrant: calls on one handle must not overlap.
 */
#ifndef KNNG_H
#define KNNG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KNNG_OK 0
#define KNNG_ERR_INVALID (-1)     /* bad argument */
#define KNNG_ERR_HIP (-2)         /* HIP runtime error (message in knng_last_error) */
#define KNNG_ERR_UNSUPPORTED (-3) /* shape outside what the kernels handle */
#define KNNG_ERR_NOMEM (-4)       /* device allocation failed */
#define KNNG_ERR_OVERFLOW (-5)    /* a bounded buffer overflowed (exact k-NN candidates) */

#define KNNG_PREC_FP32 0 /* exact FP32 lists: fp16 screen with a rigorous margin, FP32 re-rank */
#define KNNG_PREC_FP16 1 /* fp16 MFMA screen, no margin, FP32 re-rank of the survivors */
#define KNNG_PREC_BF16 2 /* bf16 MFMA screen, no margin, FP32 re-rank of the survivors */

#define KNNG_FLAG_DIRECT_NON 1u /* neighbours-of-neighbours by the per-point direct kernel */
#define KNNG_FLAG_BIG_LEAF 2u   /* every leaf through the big-bucket kernel (the path of leaves > 512) */
#define KNNG_FLAG_INCREMENTAL_NON 4u /* later neighbours-of-neighbours passes of a build skip offers
                                        whose two hops both predate the previous pass (same lists;
                                        pays when few entries change between passes) */
#define KNNG_FLAG_KD_SPLIT 16u      /* split rounds with the random-rotation KD splitter
                                        (SolutionRandomKD) instead of 2-means */
#define KNNG_FLAG_SMALL_NON_CAP 8u   /* tests: 16 neighbours-of-neighbours candidate slots per point
                                        instead of 512, so the overflow hand-off to the per-point
                                        direct kernel runs */
#define KNNG_FLAG_NON_ONE_SET 4096u  /* tests: one set of neighbours-of-neighbours candidate buffers
                                        (each sub-pass's merge before the next screen), as a build
                                        takes when a second set would not leave half the free
                                        device memory; same lists */
#define KNNG_FLAG_NON_SCREEN16 64u   /* neighbours-of-neighbours screen on a 16-bit Gram (16x16x32
                                        MFMA) of rows centred on each q instead of the FP32 Gram
                                        (d of 61-64, 97-100 or 125-128; fp16 with the exact margin
                                        in the FP32 form, so the same lists; bf16/fp16 without a
                                        margin in those forms). Measured slower at 10M x 100: the
                                        exact re-rank then reads S(q)'s rows from global memory */
#define KNNG_FLAG_NON_FP32_GRAM 512u /* neighbours-of-neighbours screen on the FP32 Gram (16x16x4
                                        MFMA) of the rows. Without it, for d of 61-64, 97-100 or
                                        125-128, the screen is an fp16 Gram (16x16x32 MFMA) of each
                                        row scaled by its power of two, read from the FP32 rows the
                                        exact re-rank uses, with the rigorous fp16 margin: the same
                                        lists, fewer MFMA cycles */
#define KNNG_FLAG_LEAF_BATCHED 128u  /* leaf stage as one persistent kernel over one work list per
                                        round (big-bucket blocks first, then buckets by decreasing
                                        size) instead of one launch per size bin; same lists */
#define KNNG_FLAG_LEAF_GRAPH 256u    /* as KNNG_FLAG_LEAF_BATCHED, its launches captured once in a
                                        hipGraph and replayed every round */
#define KNNG_FLAG_NO_BUCKET_LIMIT 1024u /* exact form: leaf buckets screen each row against its list's
                                        bound only, without the in-bucket limit (the k-th smallest
                                        upper bound of the row's distances in the bucket); the same
                                        lists, for comparisons */
#define KNNG_FLAG_LEAF_FP32_GRAM 2048u /* exact form: leaf screen on the FP32 Gram (16x16x4 MFMA) of
                                        the FP32 rows with its rigorous margin, every bucket through
                                        the big-bucket kernel (FP32 rows do not fit the LDS-resident
                                        kernel); the same lists. With KNNG_FLAG_NON_FP32_GRAM: the
                                        plain FP32 GPU form, no 16-bit Gram anywhere */
#define KNNG_FLAG_SIDE_FP32 8192u    /* split side test on the FP32 rows only, without the fp16
                                        copy whose rigorous error bound decides most points (the
                                        rest re-read their FP32 rows); the same partitions */
#define KNNG_FLAG_AUTOTUNE 32u       /* leaf stage: the first launch of each size bin above 128
                                        points (per device, precision and width, once per process)
                                        also times both workgroup shapes, 8 and 16 waves, and the
                                        faster is kept; the lists end as without the flag */

/* Build parameters. The reference hard-codes k = dims = 100 (src/Constants.hpp:23-24),
 * maxGroupSize 200/400 alternating (src/SolutionKmeans.hpp:582-587), knnIterations = 1 (:591),
 * the 3000-point small-branch threshold (:156) and a wall-clock round budget (:562). */
typedef struct knng_params {
    uint32_t k;            /* neighbours per point, 1..128 */
    uint32_t leaf_max;     /* largest leaf bucket, inclusive (reference maxGroupSize - 1); <= 65536 */
    uint32_t leaf_max_odd; /* leaf_max on odd rounds; 0 = leaf_max */
    uint32_t small_max;    /* nodes below this size take the small branch, <= 4096 */
    uint32_t iterations;   /* Lloyd iterations per split */
    uint32_t rounds;       /* split rounds */
    uint32_t non_passes;   /* neighbours-of-neighbours passes after the rounds */
    uint32_t precision;    /* KNNG_PREC_* */
    uint32_t rerank;       /* approximate forms (KNNG_PREC_FP16/BF16): the leaf stage re-ranks, per
                              point and bucket, only the survivors whose approximate distance is at
                              most the m-th smallest of the point's approximate distances in the
                              bucket, m = max(rerank, k) (taken over each lane's smallest keys: a
                              little above the m-th smallest); 0 or above 160: every survivor. The
                              exact form ignores it. Default 160 */
    uint32_t flags;        /* KNNG_FLAG_* */
    uint32_t search_k;     /* list width while building, <= 128 (0 = k); the output keeps each
                              list's first k. Small k search poorly through neighbours of
                              neighbours, whose candidate count is width^2. */
    uint32_t non_rank_lo;  /* 0: topUpSingle (SolutionKmeans.hpp:480-555), every neighbour expanded;
                              r > 0: topUpSingleOrdered (:393-477), only neighbours at distance
                              ranks >= r (the reference uses 50); list width <= 100 */
    uint64_t seed;
} knng_params;

/* Per-build timings (milliseconds, HIP events on the build stream) and counters. */
typedef struct knng_stats {
    double split_ms;       /* all split levels of all rounds */
    double side_ms;        /* of which: the full-range side test kernels */
    double leaf_ms;        /* leaf Gram + top-k merge */
    double non_ms;         /* neighbours-of-neighbours */
    double total_ms;       /* whole build on device, input resident, output written */
    uint64_t side_launches;
    uint64_t side_rows;    /* rows read by the side test */
    uint64_t leaf_launches;
    uint64_t leaves;
    uint64_t leaf_pairs;   /* sum over leaves of b*(b-1)/2 */
    uint64_t survivors;    /* screened pairs re-ranked in FP32 */
    uint64_t inserts;      /* list insertions */
    uint64_t split_failures;
    uint64_t split_fallbacks;
    uint64_t non_overflow;  /* points whose NoN candidate buffer overflowed (redone directly) */
    double non_screen_ms;   /* of non_ms: the reverse screening kernel launches */
    uint64_t non_screen_launches;
    uint64_t non_screen_bytes; /* rows and snapshots the screening kernel gathers (algorithmic) */
} knng_stats;

typedef struct knng_handle knng_handle;

const char* knng_last_error(void);
void knng_default_params(knng_params* p);

int knng_create(const knng_params* params, knng_handle** out);
void knng_destroy(knng_handle* h);
int knng_set_params(knng_handle* h, const knng_params* params);
int knng_get_stats(const knng_handle* h, knng_stats* out);

/* Replaces SolutionKmeans::constructResult (src/SolutionKmeans.hpp:558-616). points: host rows of
 * row_stride floats, the first d used (the reference passes stride 112, d 100). out_ids: n*k host,
 * each row ascending by (distance, id), padded with id 1 (padResult, src/KnnSets.hpp:334-344).
 * out_dists (nullable): n*k FP32 distances, FLT_MAX for padding. */
int knng_build(knng_handle* h, const float* points, uint32_t n, uint32_t d, size_t row_stride,
               uint32_t* out_ids, float* out_dists);

/* Same with device-resident input and output (row_stride a multiple of 4, 16-byte aligned).
 * Returns after the build completed on the device. */
int knng_build_device(knng_handle* h, const float* d_points, uint32_t n, uint32_t d,
                      size_t row_stride, uint32_t* d_out_ids, float* d_out_dists);

/* ---- stage entry points (host buffers), each one reference routine ------------------------- */

/* One round of splitKmeansBinaryProcess from the identity permutation
 * (src/SolutionKmeans.hpp:142-311, :589-591). perm_out: n; leaf_start/leaf_size: capacity n,
 * leaves ascending by start; *n_leaves receives the count. */
int knng_split_round(knng_handle* h, const float* points, uint32_t n, uint32_t d,
                     size_t row_stride, uint32_t round, uint32_t* perm_out, uint32_t* leaf_start,
                     uint32_t* leaf_size, uint32_t* n_leaves);

/* Side test and stable partition of node ranges of perm with given hyperplanes
 * (`dot(coefs, pt) >= offset` first, src/SolutionKmeans.hpp:189-197). coefs: n_nodes*d. */
int knng_partition_injected(knng_handle* h, const float* points, uint32_t n, uint32_t d,
                            size_t row_stride, uint32_t* perm, uint32_t n_nodes,
                            const uint32_t* node_start, const uint32_t* node_size,
                            const float* coefs, const float* offsets, uint32_t* lo_out);

/* addCandidatesLessThan over every leaf (src/KnnSets.hpp:131-160) on a state of n*k ids/dists,
 * each row sorted by (dist, id) and padded (FLT_MAX, 0xFFFFFFFF). ids/dists are updated in place;
 * ids == NULL starts from empty sets (then ids/dists must still be writable n*k buffers). */
int knng_leaf_stage(knng_handle* h, const float* points, uint32_t n, uint32_t d,
                    size_t row_stride, const uint32_t* perm, const uint32_t* leaf_start,
                    const uint32_t* leaf_size, uint32_t n_leaves, uint32_t* ids, float* dists,
                    int start_empty);

/* No counterpart in the reference: knng_leaf_stage for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Points, permutation and lists are device memory the caller owns: d_points n rows of row_stride
 * floats used in place (d and row_stride multiples of 4, d <= 128, 16-byte aligned), d_perm n ids,
 * d_ids/d_dists n*w (w = max(k, search_k) of the handle's parameters) updated in place, or
 * initialised first when start_empty; leaf_start/leaf_size are host arrays. Its scratch (the
 * per-point bounds, the binned leaf tables and counters) is a fixed 64 MiB device array per GPU,
 * loaded with the library and held until the process ends; no device or pinned memory is
 * allocated or freed. Needs 4 bytes per point plus 8 per leaf and 12 per 64 rows of leaves above
 * 512 points (rounded up to 256-byte blocks); a call that would need more returns
 * KNNG_ERR_UNSUPPORTED. All calls on one GPU share the array: calls that overlap on one GPU race.
 * Returns after the stage completed. */
int knng_leaf_stage_device_noalloc(knng_handle* h, const float* d_points, uint32_t n, uint32_t d,
                                   size_t row_stride, const uint32_t* d_perm,
                                   const uint32_t* leaf_start, const uint32_t* leaf_size,
                                   uint32_t n_leaves, uint32_t* d_ids, float* d_dists,
                                   int start_empty);

/* topUpSingle (src/SolutionKmeans.hpp:480-555) on a state as above, in place. */
int knng_non(knng_handle* h, const float* points, uint32_t n, uint32_t d, size_t row_stride,
             uint32_t* ids, float* dists);

/* ---- exact FP64 k-NN (recall measurement; the baseline grade.cpp compares against) -------- */

/* Exact k nearest (self excluded) of each query id by FP64 squared L2, ascending (dist, id).
 * Device pointers. upper (nullable, nq): a known upper bound on each query's k-th distance (only
 * points within it are kept); cap: candidate slots per query. Returns KNNG_ERR_OVERFLOW when a
 * query had more than cap candidates within its bound. */
int knng_exact_knn_device(knng_handle* h, const float* d_points, uint32_t n, uint32_t d,
                          size_t row_stride, const uint32_t* d_queries, uint32_t nq, uint32_t k,
                          const double* d_upper, uint32_t cap, uint32_t* d_out_ids,
                          double* d_out_dists);

/* FP64 upper bounds for knng_exact_knn_device from candidate lists (nq*k device ids): the FP64
 * distance of the k-th distinct valid entry, +inf if fewer. */
int knng_exact_upper_from_lists(knng_handle* h, const float* d_points, uint32_t n, uint32_t d,
                                size_t row_stride, const uint32_t* d_queries, uint32_t nq,
                                const uint32_t* d_lists, uint32_t k, double* d_upper);

/* Recall as grade.cpp:61-95 computes it: mean over rows of |truth_row ∩ result_row|, divided by
 * k. Host arrays of rows*k ids. */
double knng_recall(const uint32_t* result_ids, const uint32_t* truth_ids, uint32_t rows,
                   uint32_t k);

/* ---- synthetic data and device memory (for callers without their own GPU runtime) --------- */

#define KNNG_GEN_LOWRANK 0  /* z W + eps, 24 latent dims, sigma_i = 1/sqrt(1+i), unit norm */
#define KNNG_GEN_MIXTURE 1  /* 500 centres uniform in [-10,10]^d, isotropic noise 1.0 */
#define KNNG_GEN_ZIPF 2     /* 20000 centres, Zipf(1.1) weights, sigma ~ U[0.3, 2.0] */

/* Fills n rows of row_stride floats at d_points (device). */
int knng_generate_device(knng_handle* h, int kind, uint32_t n, uint32_t d, size_t row_stride,
                         uint64_t seed, float* d_points);

int knng_device_alloc(size_t bytes, void** out);
int knng_device_free(void* p);
int knng_memcpy_h2d(void* dst, const void* src, size_t bytes);
int knng_memcpy_d2h(void* dst, const void* src, size_t bytes);
int knng_memcpy_d2d(void* dst, const void* src, size_t bytes);
int knng_synchronize(void);
/* Free and total device memory of the current device (hipMemGetInfo). */
int knng_mem_get_info(size_t* free_bytes, size_t* total_bytes);

/* ---- serving: continuous batching of build jobs ---------------------------------------------
 * No counterpart in the reference (a one-shot CLI, src/main.cpp:39-71). A server owns `workers`
 * builders, each with its own HIP stream and device buffers reused across jobs. Jobs are admitted
 * in arrival order as soon as a worker is free and the device memory accounted to all workers
 * stays within hbm_budget bytes: what each idle worker's builder keeps from earlier jobs, plus,
 * for each running job, its footprint (knng_job_footprint) on top of what its worker kept. When
 * the queue's head does not fit, idle workers free what they keep. Admitted jobs run concurrently
 * on the GPU. A build is deterministic, so a job's result does not depend on what else runs. */
typedef struct knng_server knng_server;

typedef struct {
    const float* points;        /* n rows of row_stride floats (host, or device if points_on_device) */
    uint32_t n, d;
    size_t row_stride;
    uint32_t points_on_device;  /* 1: device memory, read in place until the job completes */
    knng_params params;
    uint32_t* out_ids;          /* host, n * params.k, rows padded with id 1 */
    float* out_dists;           /* host, n * params.k, nullable */
} knng_job;

typedef struct {
    int status;                 /* KNNG_OK or the job's error code */
    double queue_ms;            /* submit -> start */
    double run_ms;              /* start -> results in the caller's buffers */
    double latency_ms;          /* submit -> results in the caller's buffers */
    uint64_t footprint;         /* bytes the admission accounted for */
    uint32_t worker;
    double start_ms;            /* when a worker took the job, ms since the server was created */
    double done_ms;             /* when its results were in the caller's buffers, same clock */
    uint64_t worker_device_bytes; /* device memory the worker kept after the job */
} knng_job_result;

typedef struct {
    uint64_t budget;            /* hbm_budget of knng_server_create */
    uint64_t in_use;            /* device bytes accounted now: what idle workers keep, plus
                                   footprints of running jobs */
    uint64_t peak_in_use;       /* largest in_use so far; never above budget */
    uint64_t trims;             /* times an idle worker freed its memory for a job that did not fit */
    uint64_t queued;            /* jobs waiting for admission */
} knng_server_stats;

/* Device bytes a job needs: points, top-k state, NoN scratch, outputs. */
uint64_t knng_job_footprint(uint32_t n, uint32_t d, const knng_params* params);
int knng_server_create(uint32_t workers, uint64_t hbm_budget, knng_server** out);
/* Queues a job; buffers must stay valid until knng_server_wait returns for it. A job whose
 * footprint exceeds the budget is rejected with KNNG_ERR_NOMEM. */
int knng_server_submit(knng_server* s, const knng_job* job, uint64_t* job_id);
int knng_server_poll(knng_server* s, uint64_t job_id, int* done);
/* Blocks until the job is done, fills `out` (nullable) and forgets the job. */
int knng_server_wait(knng_server* s, uint64_t job_id, knng_job_result* out);
/* Admission accounting of the server (a snapshot). */
int knng_server_get_stats(knng_server* s, knng_server_stats* out);
/* Runs every queued job to completion, then stops the workers. */
void knng_server_destroy(knng_server* s);

#ifdef __cplusplus
}
#endif

#endif /* KNNG_H */
