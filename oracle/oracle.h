/*
 * oracle.h — CPU restatement of the reference's 2-NN semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product library never links or calls it.
 */
#ifndef SIFT_ORACLE_H
#define SIFT_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* host_lenght (src/knn_brute.cu:873-885) of n (query, base row) pairs:
 * out[i] = sequential FP32 sum of (q[qi[i]] - r[ri[i]])^2 over 128 dims. */
void oracle_host_lenght(const float* q, const float* r, const int64_t* qi, const int64_t* ri, int n,
                        float* out);
/* host_brute (src/knn_brute.cu:801-835) with host_lenght (:873-885). */
void oracle_host_brute(const float* q, const float* r, int q_n, int r_n, float* sorted4);
/* Same, queries split over `threads` OpenMP threads; identical results. */
void oracle_host_brute_mt(const float* q, const float* r, int q_n, int r_n, float* sorted4,
                          int threads);
/* host_brute with a planted mistake (used only to show the checks catch it):
 * 1 = '<=' in the first-slot test (later index wins ties, like src/knn_brute.cu:739)
 * 2 = second slot not shifted when the first slot is replaced
 * 3 = distance accumulates a-b instead of (a-b)^2 on the last dimension
 * 4 = loop starts at base row 1 (drops row 0) */
void oracle_host_brute_mutant(const float* q, const float* r, int q_n, int r_n, float* sorted4,
                              int mutation);
/* host_sort (src/knn_brute.cu:839-870): top-2 of each row of a distance matrix. */
void oracle_host_sort(const float* dist, int size, int array_size, float* sorted4);
/* Integer restatement for descriptors holding integers: int64 squared L2,
 * lexicographic (distance, index) top-2.  idx2/dist2 = UINT32_MAX if absent. */
void oracle_knn2_int(const float* q, const float* r, int q_n, int r_n, uint32_t* idx2,
                     uint32_t* dist2, int threads);
/* Ratio test: matches[i] = idx1 if sqrtf(d1)/sqrtf(d2) < thr else UINT32_MAX
 * (src/knn_brute.cu:430-442 without the unit-norm +2). */
void oracle_ratio(const float* sorted4, int q_n, float thr, uint32_t* matches);

/* ---- LSH (random hyperplanes) ---- */
/* codes[i] bit b = (h_b . x_i < 0), exact integer arithmetic
 * (src/lsh_h.cu:265-317 bit order).  H: nbits x 128 integer-valued floats. */
void oracle_lsh_codes(const float* x, int n, const float* H, int nbits, uint32_t* codes);
/* Stable sort of (code, index) pairs; perm[k] = index of k-th smallest. */
void oracle_lsh_sort(const uint32_t* codes, int n, uint32_t* perm);
/* Float mode of the LSH index (data that is not all integers in [0,255]):
 * hash input clamp(rint(x * 2^k), -255, 255); k = the largest in [-100, 100]
 * with max|x| * 2^k <= 255; re-rank by host_lenght. */
void oracle_lsh_quantize(const float* x, size_t count, int k, float* out);
int oracle_lsh_scale_exp(const float* x, size_t count);
void oracle_lsh_query_f32(const float* q, const float* r, int q_n, int r_n, int L, const uint32_t* codes_q,
                          const uint32_t* codes_r, uint32_t* idx2, float* dist2, uint32_t* ncand, int threads);
/* Exact 2-NN of each query restricted to the union over tables t of the base
 * rows whose code in table t equals the query's code.  codes_q[t*q_n+i],
 * codes_r[t*r_n+j].  ncand[i] = size of the candidate union. */
void oracle_lsh_query(const float* q, const float* r, int q_n, int r_n, int L,
                      const uint32_t* codes_q, const uint32_t* codes_r, uint32_t* idx2,
                      uint32_t* dist2, uint32_t* ncand, int threads);

/* ---- learned hash MLP 128 -> H (ReLU) -> B ---- */
/* y = W2 relu(W1 x + b1) + b2 in double precision. x: n x 128, W1: H x 128,
 * W2: B x H.  hid: n x H (post-ReLU), out: n x B. */
void oracle_mlp_forward(const float* x, int n, const float* W1, const float* b1, int H,
                        const float* W2, const float* b2, int B, double* hid, double* out);

/* The library's learned hash (H = 256, outputs padded to 32), parameters in
 * its flat layout: W1[256][128] (layer 1 multiplies by W1 / 64), b1[256],
 * W2[32][256], b2[32].
 * quant = 1 reproduces the GPU arithmetic's roundings: W1, W2 and the ReLU
 * outputs rounded to bfloat16 (nearest even); accumulation in double.
 * quant = 0: no rounding (a smooth function, for finite differences).
 * hid: n x 256 (post-ReLU, as fed to layer 2), z: n x 32 (bits >= nbits = 0). */
void oracle_hash_forward(const float* params, int nbits, const float* x, int n, int quant,
                         double* hid, double* z);
/* Loss of a batch of B pairs (rows i and B + i of s / z, N = 2B rows):
 * sum (s_x - s_y)^2 / (4 B b) + alpha sum_k mean_k(s)^2 / b
 * + beta sum_{k != l} (S^T S / N)_kl^2 / b^2 + gamma sum (|z| - 1)^2 / (N b),
 * with s and z independent inputs (N x 32, bits >= nbits ignored). */
double oracle_hash_loss(const double* s, const double* z, int B, int nbits, double alpha,
                        double beta, double gamma);
/* At s = sign(z) (sign(0) = +1): dls = dL/ds, and dz = dls (straight-through:
 * the sign passes the gradient unchanged) plus the derivative of the gamma
 * term; returns the loss. */
double oracle_hash_loss_grad(const double* z, int B, int nbits, double alpha, double beta,
                             double gamma, double* dls, double* dz);
/* Gradient of sum_ik dz_ik * z_ik over the parameters (flat layout) given the
 * forward's hid; quant as in oracle_hash_forward (rounded weights in the
 * chain rule). */
void oracle_hash_backward(const float* params, int nbits, const float* x, int n, int quant,
                          const double* hid, const double* dz, double* grads);

#ifdef __cplusplus
}
#endif
#endif
