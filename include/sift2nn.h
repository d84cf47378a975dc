/*
 * sift2nn.h — C ABI of the MI355X (gfx950) 2-NN SIFT matching library.
 *
 * Two layers:
 *   1. Reference-named entry points with the reference's argument order, so a
 *      caller of Scoooooo/matching_SIFT relinks against this library.  The
 *      cuBLAS handle becomes an opaque sift_handle_t; CUDA float4 becomes the
 *      layout-identical sift_float4.  Each declaration cites the reference
 *      declaration it replaces.
 *   2. The native API (sift_*): explicit compute kind, device outputs, return
 *      codes, optional asynchrony.
 *
 * Conventions (all entry points):
 *   - Pointers named *_dev are device pointers; FLOAT_HOST inputs are host
 *     memory (pinned or pageable).
 *   - Return 0 on success, a negative SIFT_ERR_* code on failure; the message
 *     is available from sift_last_error().  void entry points (kept void for
 *     source compatibility) record the code in sift_last_status().
 *   - A handle owns one HIP stream and grow-only scratch.  One handle per host
 *     thread.  Reference-named calls are blocking; native calls block unless
 *     named *_async.
 *   - Distances are squared L2.  For descriptors holding integers in [0,255]
 *     every compute kind except FP8 returns bit-exact indices and distances
 *     (ties go to the lowest base index).
 *   - Other data (e.g. unit-norm floats, SIFT / 512) with the float kinds
 *     (BF16, F16, F32): the MFMA pass ranks coarsely (values rounded to the
 *     kind), the 16 best coarse candidates of each query are re-ranked with
 *     the reference host_lenght's sequential FP32 sum of (q_i - r_i)^2 over
 *     the rows as given, and those are the results.  Distances are then
 *     bit-equal to host_brute's for the returned rows, and the indices are
 *     host_brute's unless a true 2-NN row is not among the candidates: this
 *     needs another row of the same slab (or 16 others) within the coarse
 *     rounding of its distance (2^-6 (|q| + |r|)^2 for BF16, 2^-9 for F16,
 *     2^-16 for F32).  I8 and FP8 return SIFT_ERR_NOT_EXACT on such data.
 */
#ifndef SIFT2NN_H
#define SIFT2NN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sift_context* sift_handle_t;

/* One descriptor: 128 floats (include/knn_brute.h:5 of the reference). */
typedef float des_t_f[128];

/* Layout-identical to CUDA/HIP float4 as used by the reference's sorted[]
 * output: x = d1, y = d2, z = (float)idx1, w = (float)idx2. */
typedef struct {
    float x, y, z, w;
} sift_float4;

/* Input memory/type codes.  0..2 match the reference's type_mem enum
 * (src/knn_brute.cu:15-19, src/lsh_h.cu:710-714); 3 and 4 are additions. */
enum {
    SIFT_FLOAT_HOST = 0,   /* float[n][128] in host memory                   */
    SIFT_FLOAT_DEVICE = 1, /* float[n][128] in device memory                 */
    SIFT_HALF_DEVICE = 2,  /* IEEE fp16 [n][128] in device memory            */
    SIFT_BF16_DEVICE = 3,  /* bfloat16 [n][128] in device memory             */
    SIFT_U8_DEVICE = 4     /* uint8 [n][128] in device memory (raw SIFT)     */
};

/* Arithmetic used by the fused distance + top-2 kernel. */
enum {
    SIFT_COMPUTE_AUTO = 0, /* integer data in [0,255] -> I8, else FP32       */
    SIFT_COMPUTE_BF16 = 1, /* BF16 MFMA, FP32 accumulate                     */
    SIFT_COMPUTE_F16 = 2,  /* FP16 MFMA, FP32 accumulate                     */
    SIFT_COMPUTE_F32 = 3,  /* FP32 MFMA                                      */
    SIFT_COMPUTE_I8 = 4,   /* int8 MFMA (x - 128, 127 - x), int32 accumulate */
    SIFT_COMPUTE_FP8 = 5   /* e4m3 coarse pass + exact re-rank, see below    */
};
/* SIFT_COMPUTE_FP8 (sift_knn2*, brute-force indexes, pairs plans): a coarse
 * pass on per-row-scaled e4m3 keeps 16 candidate rows per query, and an exact
 * int32 re-rank of those returns the two nearest among them with exact
 * distances.  Needs integer data in [0,255] (SIFT_ERR_NOT_EXACT otherwise).
 * Approximate: the true 2-NN can be missed (exact when every value is in
 * [0,15]); no counterpart in the reference. */

/* Hash kinds for the LSH index (lsh_type in the reference, src/lsh_h.cu:716-718). */
enum {
    SIFT_LSH_HYPERPLANE = 0,     /* Gaussian hyperplanes (reference DOT_INT_BUCKET) */
    SIFT_LSH_LEARNED = 1,        /* learned 128->256->b MLP hash                    */
    SIFT_LSH_HYPERPLANE_PM1 = 2  /* +-1 hyperplanes as in src/lsh_h.cu:51-84        */
};

enum {
    SIFT_OK = 0,
    SIFT_ERR_INVALID = -1,     /* bad argument                               */
    SIFT_ERR_OOM = -2,         /* device allocation failed                   */
    SIFT_ERR_HIP = -3,         /* HIP runtime error                          */
    SIFT_ERR_UNSUPPORTED = -4, /* valid but unsupported combination          */
    SIFT_ERR_NOT_EXACT = -5    /* data outside the exact domain of the kind  */
};

/* ---------------------------------------------------------------- handle -- */

/* Replaces cublasCreate/cublasDestroy of the reference's caller (src/test.cu:51-52). */
int sift_create(sift_handle_t* handle);
int sift_destroy(sift_handle_t handle);
const char* sift_last_error(sift_handle_t handle); /* NULL handle: thread-local */
int sift_last_status(void);                        /* status of the last call on this thread */
int sift_synchronize(sift_handle_t handle);
void* sift_get_stream(sift_handle_t handle);       /* hipStream_t */
int sift_device_name(char* buf, int len);

/* Device-memory plumbing for callers without their own HIP binding. */
int sift_malloc(void** ptr, size_t bytes);
int sift_free(void* ptr);
int sift_host_alloc(void** ptr, size_t bytes); /* pinned, device-mapped */
int sift_host_free(void* ptr);
int sift_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int sift_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int sift_memset(void* dst_dev, int value, size_t bytes);

/* Per-kernel timing with HIP events on the handle's stream (0 = off). */
int sift_profile_enable(sift_handle_t handle, int on);
int sift_profile_read(sift_handle_t handle, const char* kernel, double* total_ms, long* launches);
int sift_profile_reset(sift_handle_t handle);
/* The brute-force launch shape (query blocks per wave, waves per block, tiles
 * per stage) that calls of this compute kind (not AUTO) use on a base of r_n
 * rows: the one the handle's first such call timed, or the kind's default
 * before that.  No counterpart in the reference: for benchmark and tuning
 * reports. */
int sift_knn2_shape(sift_handle_t handle, int compute, uint32_t r_n, int* qw, int* waves, int* tps);

/* ------------------------------------------------ reference entry points -- */

/* Replaces int cublas_2nn_sift(...) (include/knn_brute.h:22, src/knn_brute.cu:24).
 * Exact 2-NN of q against r, then the ratio test sqrt(d1)/sqrt(d2) < threshold:
 * matches[i] = idx1 or UINT32_MAX.  type: SIFT_FLOAT_HOST/DEVICE/HALF/BF16/U8.
 * matches: device pointer (or host memory for SIFT_FLOAT_HOST).
 * stream_n, use and warps are accepted for source compatibility; the fused
 * kernel needs no distance buffer, so they do not change the result. */
int cublas_2nn_sift(void* q_points, void* r_points, int type, uint32_t q_n, uint32_t r_n,
                    uint32_t* matches, float threshold, sift_handle_t handle, int stream_n,
                    size_t use, int warps);

/* Replaces int cublas_2nn_sift_batch(...) (include/knn_brute.h:23, src/knn_brute.cu:213).
 * q_points typed by `type`; R is fp16 [r_n][128] on the device.  Q and dist are
 * caller scratch in the reference; here they are unused and may be NULL.
 * stream (hipStream_t or NULL) is the stream to run on. */
int cublas_2nn_sift_batch(void* q_points, void* Q, int type, void* R, uint32_t q_n, uint32_t r_n,
                          void* dist, uint32_t* matches, float threshold, sift_handle_t handle,
                          void* stream, int warps);

/* Replaces void device_brute(...) (include/knn_brute.h:14, src/knn_brute.cu:619).
 * Device float inputs, device float4 output {d1, d2, idx1, idx2} (sorted: a
 * float4 or sift_float4 array); no handle (a per-device default handle is
 * used).  Exact FP32 semantics of host_brute (see the conventions above).
 * Result: sift_last_status(). */
void device_brute(const des_t_f* q_points, const des_t_f* r_points, int q_n, int r_n, void* sorted);

/* Replaces void cublas_2nn_f(...) (include/knn_brute.h:18, src/knn_brute.cu:523). */
void cublas_2nn_f(const des_t_f* q_points, const des_t_f* r_points, int q_n, int r_n, void* sorted,
                  sift_handle_t handle);

/* Replaces void cublas_2nn_brute_f(...) (include/knn_brute.h:17, src/knn_brute.cu:567).
 * dist is the reference's caller-owned distance buffer; unused, may be NULL. */
void cublas_2nn_brute_f(const des_t_f* q_points, const des_t_f* r_points, int q_n, int r_n, void* sorted,
                        float* dist, sift_handle_t handle);

/* Replaces int lsh_gpu(...) (include/lsh_h.h:69, src/lsh_h.cu:722): build an
 * LSH index over r with l tables of nbits bits (the float mode when q or r is
 * not all integers in [0,255]), query it with q, write
 * matches[i] = idx1 of the re-ranked candidates when the ratio test passes,
 * UINT32_MAX otherwise (also when a query has no candidate).  lsh_type selects
 * SIFT_LSH_HYPERPLANE (0); stream_n is accepted and unused.  Hyperplanes come
 * from the handle's seed (sift_set_seed). */
int lsh_gpu(void* q_points, void* r_points, int type, uint32_t q_n, uint32_t r_n,
            uint32_t* matches, float threshold, sift_handle_t handle, int stream_n, int l,
            int lsh_type, int nbits);

/* ------------------------------------------------------------ native API -- */

int sift_set_seed(sift_handle_t handle, uint64_t seed);

/* Exact 2-NN.  q, r typed by `type` (device pointers unless SIFT_FLOAT_HOST).
 * Outputs (device pointers, each may be NULL):
 *   idx  uint32[q_n][2]  nearest, second nearest (UINT32_MAX when r_n < 2)
 *   dist float[q_n][2]   squared L2 (FLT_MAX when absent)
 *   sorted sift_float4[q_n]  reference float4 layout
 *   matches uint32[q_n]  idx1 if sqrt(d1)/sqrt(d2) < ratio else UINT32_MAX
 * compute: SIFT_COMPUTE_*.  I8 and FP8 require integer data in [0,255]
 * (checked: SIFT_ERR_NOT_EXACT otherwise); BF16, F16 and F32 take any data
 * (exact re-rank of coarse candidates, see the conventions above).  The
 * async variant returns without checking: errors of its launches (input
 * outside an integer kind's domain, a failed wait) are reported by the next
 * blocking call on the handle or by sift_synchronize. */
int sift_knn2(sift_handle_t handle, const void* q, const void* r, int type, uint32_t q_n,
              uint32_t r_n, int compute, uint32_t* idx, float* dist, sift_float4* sorted,
              uint32_t* matches, float ratio);
int sift_knn2_async(sift_handle_t handle, const void* q, const void* r, int type, uint32_t q_n,
                    uint32_t r_n, int compute, uint32_t* idx, float* dist, sift_float4* sorted,
                    uint32_t* matches, float ratio);

/* Brute-force index: the base packed once in the MFMA operand layout, queried
 * many times (removes the per-call base conversion of the reference).  Only
 * F32 indexes take a base that is not all integers in [0,255] (the others
 * return SIFT_ERR_NOT_EXACT at build: the packed rows are all they keep);
 * queries may be any data for BF16, F16 and F32 indexes. */
typedef struct sift_bf_index* sift_bf_index_t;
int sift_bf_index_build(sift_handle_t handle, const void* r, int type, uint32_t r_n, int compute,
                        sift_bf_index_t* index);
int sift_bf_index_query(sift_handle_t handle, sift_bf_index_t index, const void* q, int type,
                        uint32_t q_n, uint32_t* idx, float* dist, sift_float4* sorted,
                        uint32_t* matches, float ratio, int blocking);
int sift_bf_index_destroy(sift_bf_index_t index);
/* sift_bf_index_query without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Its scratch (packed queries, the
 * fused kernel's partial top-2 and seed bounds) lives in one device array of
 * SIFT_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; calls
 * whose scratch would not fit (e.g. more than about 16k queries against 1M
 * rows) return SIFT_ERR_INVALID.  Device query rows only (type 1..4); FP8
 * indexes return SIFT_ERR_UNSUPPORTED.  Every call on a GPU shares that
 * array: calls that overlap on one GPU (other handles, other threads) race
 * and must be serialised by the caller.  blocking = 0 enqueues on the
 * handle's stream and returns (capturable: no memset, allocation or sync);
 * input errors are then reported by the next blocking call. */
#define SIFT_NOALLOC_SCRATCH_BYTES (64u << 20)
int sift_bf_index_query_noalloc(sift_handle_t handle, sift_bf_index_t index, const void* q, int type,
                                uint32_t q_n, uint32_t* idx, float* dist, sift_float4* sorted,
                                uint32_t* matches, float ratio, int blocking);

/* LSH index.  Replaces the per-call table build of lsh_gpu
 * (src/lsh_h.cu:903-914) and lsh_thread_sort_reduce (include/lsh_h.h:66,
 * src/lsh_h.cu:565): L tables of nbits-bit sign codes (bit b set when
 * h_b . x < 0), all tables' (code, row) pairs radix-sorted on the device.
 * A base of integers in [0,255] hashes x itself and re-ranks candidates in
 * exact int32 (queries must then be integers in [0,255] too:
 * SIFT_ERR_NOT_EXACT otherwise).  Any other base (unit-norm floats, SIFT /
 * 512, ...) builds the float mode: x is hashed as the integers
 * clamp(rint(x * 2^k), -255, 255), k the largest with max|base| * 2^k <= 255,
 * so codes stay exact, and candidates are re-ranked with the reference
 * host_lenght's sequential FP32 distance over the rows as given; queries may
 * be any data.  nbits in [1,32], L in [1,256], nbits + ceil(log2 L) <= 32. */
typedef struct sift_lsh_index* sift_lsh_index_t;
int sift_lsh_build(sift_handle_t handle, const void* r, int type, uint32_t r_n, int nbits, int L,
                   int hash_kind, uint64_t seed, sift_lsh_index_t* index);
/* Exact int32 re-rank of the union of the query's buckets over all tables.
 * Outputs as sift_knn2 plus ncand[q_n] = candidates examined (with repeats
 * across tables); any output may be NULL. */
int sift_lsh_query(sift_handle_t handle, sift_lsh_index_t index, const void* q, int type,
                   uint32_t q_n, uint32_t* idx, float* dist, uint32_t* ncand, uint32_t* matches,
                   float ratio, int blocking);
int sift_lsh_destroy(sift_lsh_index_t index);
/* Query codes, device uint32[L][q_n] (blocking). */
int sift_lsh_codes(sift_handle_t handle, sift_lsh_index_t index, const void* q, int type,
                   uint32_t q_n, uint32_t* codes);
int sift_lsh_info(sift_lsh_index_t index, uint32_t* r_n, int* nbits, int* L);
/* 1 and the scale exponent k for a float-mode index, 0 and 0 otherwise. */
int sift_lsh_hash_input(sift_lsh_index_t index, int* float_mode, int* scale_exp);
/* Host copies: hyperplanes float[L][nbits][128]; sorted keys (t << nbits | code)
 * and base rows, uint32[L*r_n] each. */
int sift_lsh_get_hyperplanes(sift_lsh_index_t index, float* planes);
int sift_lsh_get_buckets(sift_lsh_index_t index, uint32_t* keys, uint32_t* rows);

/* ---------------------------------------------------------------------------
 * Batched image pairs: many (query image, base image) pairs in one persistent
 * launch.  Per pair the result is what cublas_2nn_sift (include/knn_brute.h:22,
 * src/knn_brute.cu:24) computes for one image pair in src/test.cu, with true
 * squared L2 and lowest-index ties: 2-NN indices into the pair's base image,
 * squared distances, and the ratio test.
 *
 * Descriptors are CSR: all images of one side concatenated row-wise in one
 * device array of `type` (1..4), image i = rows [off[i], off[i+1]).  pairs is
 * uint32[n_pairs][2] = (query image, base image), host memory.  Outputs are
 * device arrays over the pairs' queries in pair order (pair p's queries start
 * at the sum of the query-image sizes of pairs 0..p-1):
 *   idx uint32[Q][2], dist float[Q][2], matches uint32[Q] (idx1 or UINT32_MAX);
 * each may be NULL.  An empty base image gives UINT32_MAX / FLT_MAX.
 * ------------------------------------------------------------------------- */
/* ---------------------------------------------------------------------------
 * Learned hash (no counterpart in the reference): an MLP 128 -> 256 (ReLU) -> b,
 * b <= 32, whose output signs are the hash bits (bit k set when z_k < 0, the
 * hyperplane convention of src/lsh_h.cu:302-309).  BF16 weights on the MFMAs,
 * FP32 accumulation; trained on the GPU with sign binarisation and a
 * straight-through gradient on positive pairs (x, exact 1st / 2nd nearest base
 * row of x, labelled by the brute-force kernel).
 * Parameters (FP32, SIFT_MLP_NPARAM floats): W1[256][128] (layer 1 multiplies
 * by W1 / 64, keeping stored weights near unit scale for inputs in [0,255]),
 * b1[256], W2[32][256] (rows >= b unused), b2[32].
 * ------------------------------------------------------------------------- */
typedef struct sift_mlp* sift_mlp_t;
#define SIFT_MLP_NPARAM (256 * 128 + 256 + 32 * 256 + 32)
typedef struct {
    int epochs;      /* passes over the labelled pairs */
    int batch;       /* pairs per step */
    float lr;        /* Adam step size */
    float alpha;     /* weight of the bit-balance term */
    float beta;      /* weight of the bit-decorrelation term */
    float gamma;     /* weight of the |z| -> 1 term */
    float max_ratio; /* positives: (x, 1st NN) where sqrt(d1 / d2) < max_ratio (a distinctive
                        match); 0 = every (x, 1st NN) and (x, 2nd NN) */
    uint64_t seed;   /* pair shuffling */
} sift_mlp_train_opts;
int sift_mlp_create(sift_handle_t h, int nbits, uint64_t seed, sift_mlp_t* mlp);
int sift_mlp_destroy(sift_mlp_t mlp);
int sift_mlp_get_params(sift_mlp_t mlp, float* params);       /* host, SIFT_MLP_NPARAM */
int sift_mlp_set_params(sift_mlp_t mlp, const float* params);
/* Codes (uint32[n], device) and optionally the outputs z (float[n][32], device). */
int sift_mlp_codes(sift_handle_t h, sift_mlp_t mlp, const void* x, int type, uint32_t n,
                   uint32_t* codes, float* z);
/* One batch of B pairs (x_i, y_i), device rows of `type` (1..4): the loss
 * weights {alpha, beta, gamma}, gradients (host, SIFT_MLP_NPARAM), loss (host
 * float[2]: total, Hamming fraction of the pairs) and z (host float[2B][32]). */
int sift_mlp_grad(sift_handle_t h, sift_mlp_t mlp, const void* x, const void* y, int type,
                  uint32_t B, const float* loss_w, float* grads, float* loss, float* z);
/* Label x (device rows) by its exact 2-NN in base with the brute-force kernel,
 * then train (the first call also centres every output bit on the training
 * rows); epoch_loss (host, opts->epochs floats) may be NULL. */
int sift_mlp_train(sift_handle_t h, sift_mlp_t mlp, const void* x, uint32_t n, const void* base,
                   uint32_t r_n, int type, const sift_mlp_train_opts* opts, float* epoch_loss);
/* Multi-probe (no counterpart in the reference): every query also looks up,
 * per table, the buckets of its code with each of the `probes` least
 * confident bits flipped (smallest |h.x| or |z| first; 0 = off, at most
 * min(nbits, 16)).  A setting of the index used by every later query. */
int sift_lsh_set_probes(sift_lsh_index_t index, int probes);
/* Mean bucket size seen by a point drawn like the base rows (no counterpart
 * in the reference): sum over the buckets of every table of size^2, divided
 * by L * n, i.e. the expected candidates of one (table, probe) lookup.  The
 * query picks its re-rank schedule from it; callers can size probes and
 * tables with it. */
int sift_lsh_bucket_stat(sift_lsh_index_t index, double* cand_per_lookup);
/* The codes a query looks up, device [L][probes + 1][q_n]: per table the
 * base code, then the probes in order. */
int sift_lsh_probe_codes(sift_handle_t h, sift_lsh_index_t index, const void* q_points, int type,
                         uint32_t q_n, uint32_t* codes);
/* Persist an LSH index (hyperplane or learned: the hash, the sorted tables and
 * the gather copy) and load it on any handle without rebuilding (the reference
 * rebuilds per call, src/lsh_h.cu:903-914).  A loaded learned index owns
 * copies of its networks.  No counterpart in the reference. */
int sift_lsh_save(sift_lsh_index_t index, const char* path);
int sift_lsh_load(sift_handle_t h, const char* path, sift_lsh_index_t* index);
/* Check an index file on the host without loading it (sift_lsh_load runs the
 * same checks): header, size, row ids below the base size, every table's keys
 * sorted under its own prefix, stored norms equal to their rows, hyperplane
 * values integers in [-127, 127].  SIFT_ERR_INVALID with the reason in
 * sift_last_error(NULL) otherwise. */
int sift_lsh_check_file(const char* path);
/* LSH index whose table t hashes with network mlps[t] (all of the same b);
 * the networks must outlive the index.  Query it with sift_lsh_query. */
int sift_lsh_build_learned(sift_handle_t h, const void* r, int type, uint32_t r_n,
                           const sift_mlp_t* mlps, int L, sift_lsh_index_t* index);
/* sift_lsh_query, then exact brute force (bf: the same base, packed) for the
 * queries the LSH answer does not settle: fewer than min_cand candidates, or,
 * when accept_ratio > 0, a best candidate failing sqrt(d1)/sqrt(d2) <
 * accept_ratio against the next candidate (0 = count rule only; at most 1).
 * n_fallback (host) = how many.  No counterpart in the reference. */
int sift_lsh_query_fallback(sift_handle_t h, sift_lsh_index_t index, sift_bf_index_t bf,
                            const void* q_points, int type, uint32_t q_n, uint32_t min_cand,
                            float accept_ratio, uint32_t* idx, float* dist, uint32_t* ncand,
                            uint32_t* matches, float ratio, uint32_t* n_fallback);

typedef struct sift_pairs_plan* sift_pairs_plan_t;
#define SIFT_PAIRS_GRAPH 1u    /* capture the execute sequence in a HIP graph (replayed while
                                  the pointers and ratio stay the same) */
#define SIFT_PAIRS_AUTOTUNE 2u /* on the first execute, time the kind's tile shapes on the
                                  real data and keep the fastest */
/* compute: BF16, F16, F32 or I8 (AUTO is not accepted: a plan never reads the data at
 * build time).  BF16, F16 and I8 need integer descriptors in [0,255] (SIFT data) and
 * are then exact; F32 takes any data.  The plan owns its device buffers; execute
 * allocates nothing. */
int sift_pairs_plan_create(sift_handle_t h, int type, int compute, uint32_t n_qimg,
                           const uint32_t* q_off, uint32_t n_rimg, const uint32_t* r_off,
                           uint32_t n_pairs, const uint32_t* pairs, unsigned flags,
                           sift_pairs_plan_t* plan);
int sift_pairs_plan_execute(sift_handle_t h, sift_pairs_plan_t plan, const void* q_desc,
                            const void* r_desc, uint32_t* idx, float* dist, uint32_t* matches,
                            float ratio, int blocking);
/* Fix the tile shape (qw query blocks per wave, waves per block, tiles per
 * stage; one of the kind's listed shapes); drops a captured graph. */
int sift_pairs_plan_set_shape(sift_pairs_plan_t plan, int qw, int waves, int tps);
/* out[9]: [0] total queries Q, [1] work items, [2] grid blocks, [3..5] shape
 * (qw, waves, tps), [6] graph captured (0/1), [7] compute kind, [8] 1 if any
 * executed input held a value that is not an integer in [0,255] (sticky;
 * results of BF16/F16/I8 plans are then not exact). */
int sift_pairs_plan_info(sift_pairs_plan_t plan, int64_t* out9);
int sift_pairs_plan_destroy(sift_pairs_plan_t plan);
/* One-shot: plan (AUTO allowed: I8 for integer data in [0,255], else F32),
 * execute, synchronize, destroy.  type 0 (host float) is staged.  Explicit
 * BF16/F16/I8 with non-integer data returns SIFT_ERR_NOT_EXACT. */
int sift_match_pairs(sift_handle_t h, const void* q_desc, const void* r_desc, int type,
                     uint32_t n_qimg, const uint32_t* q_off, uint32_t n_rimg,
                     const uint32_t* r_off, uint32_t n_pairs, const uint32_t* pairs,
                     int compute, uint32_t* idx, float* dist, uint32_t* matches, float ratio);

/* ---- SIFT vector files and recall (reference src/test.cu:76-155, :224-301) ----
 * Records are an int32 dimension d followed by d elements: float32 (.fvecs),
 * int32 (.ivecs: ground-truth neighbour ids) or uint8 (.bvecs).  Host only;
 * errors go to sift_last_error(NULL). */
enum { SIFT_VECS_F32 = 0, SIFT_VECS_I32 = 1, SIFT_VECS_U8 = 2 };
/* Number of records and their dimension (checked: whole records). */
int sift_vecs_info(const char* path, int elem, uint32_t* n, uint32_t* dim);
/* Records [first, first + n) into out, n x cols row-major (cols 0 = d; the
 * first cols elements of each record, as the reference reads the first
 * `recall` of 100 ground-truth ids).  F32/U8 rows become float32 times
 * scale (the reference's real-data mode divides by 512); I32 rows stay int32. */
int sift_vecs_read(const char* path, int elem, uint64_t first, uint32_t n, uint32_t cols, float scale,
                   void* out);
/* hits[k] = number of queries whose found[q] is among the first depths[k]
 * ids of gt[q][0..gt_k) (UINT32_MAX never counts), 1 <= depths[k] <= gt_k:
 * the reference's recall loop, as counts of found rather than missed. */
int sift_recall_count(const uint32_t* found, uint32_t q_n, const int32_t* gt, uint32_t gt_k,
                      const uint32_t* depths, int n_depths, uint32_t* hits);

#ifdef __cplusplus
}
#endif

#endif /* SIFT2NN_H */
