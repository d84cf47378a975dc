/*
 * oracle.c — CPU restatement of the reference's 2-NN path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and the timed CPU
 * baseline of bench.py.  Never linked into or called by the product library.
 *
 * Built with -O2 -fno-fast-math -ffp-contract=off so the float loop keeps the
 * reference's evaluation order: dist += (a-b)*(a-b), one rounding per op.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_MAXFLOAT FLT_MAX /* MAXFLOAT in the reference (math.h) */

/* host_lenght, src/knn_brute.cu:873-885: sequential FP32 sum over 128 dims. */
static float host_lenght(const float* x, const float* y) {
    float dist = 0.0f;
    for (int i = 0; i < 128; i++) {
        float a = x[i];
        float b = y[i];
        float c = a - b;
        dist += c * c;
    }
    return dist;
}

/* One query of host_brute, src/knn_brute.cu:806-833: strict '<' on both
 * slots, so among equal distances the lowest base index is kept. */
static void host_brute_one(const float* qi, const float* r, int r_n, float* out4) {
    float x = ORACLE_MAXFLOAT, y = ORACLE_MAXFLOAT, z = ORACLE_MAXFLOAT, w = ORACLE_MAXFLOAT;
    for (int ii = 0; ii < r_n; ii++) {
        float temp = host_lenght(qi, r + (size_t)ii * 128);
        if (temp < x) {
            y = x;
            x = temp;
            w = z;
            z = (float)ii;
        } else if (temp < y) {
            y = temp;
            w = (float)ii;
        }
    }
    out4[0] = x;
    out4[1] = y;
    out4[2] = z;
    out4[3] = w;
}

void oracle_host_lenght(const float* q, const float* r, const int64_t* qi, const int64_t* ri, int n,
                        float* out) {
    for (int i = 0; i < n; i++) out[i] = host_lenght(q + (size_t)qi[i] * 128, r + (size_t)ri[i] * 128);
}

void oracle_host_brute(const float* q, const float* r, int q_n, int r_n, float* sorted4) {
    for (int i = 0; i < q_n; i++) host_brute_one(q + (size_t)i * 128, r, r_n, sorted4 + (size_t)i * 4);
}

void oracle_host_brute_mt(const float* q, const float* r, int q_n, int r_n, float* sorted4,
                          int threads) {
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
#endif
    for (int i = 0; i < q_n; i++) host_brute_one(q + (size_t)i * 128, r, r_n, sorted4 + (size_t)i * 4);
}

void oracle_host_brute_mutant(const float* q, const float* r, int q_n, int r_n, float* sorted4,
                              int mutation) {
    for (int i = 0; i < q_n; i++) {
        const float* qi = q + (size_t)i * 128;
        float x = ORACLE_MAXFLOAT, y = ORACLE_MAXFLOAT, z = ORACLE_MAXFLOAT, w = ORACLE_MAXFLOAT;
        for (int ii = (mutation == 4 ? 1 : 0); ii < r_n; ii++) {
            const float* rj = r + (size_t)ii * 128;
            float temp = 0.0f;
            for (int d = 0; d < 128; d++) {
                float c = qi[d] - rj[d];
                temp += (mutation == 3 && d == 127) ? c : c * c;
            }
            int first = (mutation == 1) ? (temp <= x) : (temp < x);
            if (first) {
                if (mutation != 2) {
                    y = x;
                    w = z;
                }
                x = temp;
                z = (float)ii;
            } else if (temp < y) {
                y = temp;
                w = (float)ii;
            }
        }
        float* o = sorted4 + (size_t)i * 4;
        o[0] = x;
        o[1] = y;
        o[2] = z;
        o[3] = w;
    }
}

/* host_sort, src/knn_brute.cu:839-870. */
void oracle_host_sort(const float* dist, int size, int array_size, float* sorted4) {
    for (int i = 0; i < array_size; i++) {
        float x = ORACLE_MAXFLOAT, y = ORACLE_MAXFLOAT, z = ORACLE_MAXFLOAT, w = ORACLE_MAXFLOAT;
        const float* row = dist + (size_t)i * size;
        for (int ii = 0; ii < size; ii++) {
            if (row[ii] < x) {
                y = x;
                x = row[ii];
                w = z;
                z = (float)ii;
            } else if (row[ii] < y) {
                y = row[ii];
                w = (float)ii;
            }
        }
        float* o = sorted4 + (size_t)i * 4;
        o[0] = x;
        o[1] = y;
        o[2] = z;
        o[3] = w;
    }
}

/* Lexicographic (distance, index) top-2 insert. */
static inline void top2_insert(uint64_t d, uint32_t j, uint64_t* d1, uint32_t* i1, uint64_t* d2,
                               uint32_t* i2) {
    if (d < *d1 || (d == *d1 && j < *i1)) {
        *d2 = *d1;
        *i2 = *i1;
        *d1 = d;
        *i1 = j;
    } else if (d < *d2 || (d == *d2 && j < *i2)) {
        *d2 = d;
        *i2 = j;
    }
}

void oracle_knn2_int(const float* q, const float* r, int q_n, int r_n, uint32_t* idx2,
                     uint32_t* dist2, int threads) {
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
#endif
    for (int i = 0; i < q_n; i++) {
        int32_t qi[128];
        for (int d = 0; d < 128; d++) qi[d] = (int32_t)q[(size_t)i * 128 + d];
        uint64_t d1 = UINT64_MAX, d2 = UINT64_MAX;
        uint32_t i1 = UINT32_MAX, i2 = UINT32_MAX;
        for (int j = 0; j < r_n; j++) {
            const float* rj = r + (size_t)j * 128;
            int64_t s = 0;
            for (int d = 0; d < 128; d++) {
                int64_t c = (int64_t)qi[d] - (int64_t)rj[d];
                s += c * c;
            }
            top2_insert((uint64_t)s, (uint32_t)j, &d1, &i1, &d2, &i2);
        }
        idx2[2 * (size_t)i] = i1;
        idx2[2 * (size_t)i + 1] = i2;
        dist2[2 * (size_t)i] = d1 == UINT64_MAX ? UINT32_MAX : (uint32_t)d1;
        dist2[2 * (size_t)i + 1] = d2 == UINT64_MAX ? UINT32_MAX : (uint32_t)d2;
    }
}

void oracle_ratio(const float* sorted4, int q_n, float thr, uint32_t* matches) {
    for (int i = 0; i < q_n; i++) {
        const float* s = sorted4 + (size_t)i * 4;
        float a = sqrtf(s[0]);
        float b = sqrtf(s[1]);
        matches[i] = (a / b < thr) ? (uint32_t)s[2] : UINT32_MAX;
    }
}

/* ---------------------------------------------------------------- LSH -- */

/* src/lsh_h.cu:302-309: bit b is set when the projection on hyperplane b is
 * negative.  Integer-valued x and H make the dot product exact in int64. */
void oracle_lsh_codes(const float* x, int n, const float* H, int nbits, uint32_t* codes) {
    for (int i = 0; i < n; i++) {
        uint32_t code = 0;
        for (int b = 0; b < nbits; b++) {
            int64_t s = 0;
            for (int d = 0; d < 128; d++)
                s += (int64_t)x[(size_t)i * 128 + d] * (int64_t)H[(size_t)b * 128 + d];
            if (s < 0) code |= 1u << b;
        }
        codes[i] = code;
    }
}

/* Hash input of the float mode: clamp(rint(x * 2^k), -255, 255), exact in
 * float (a power-of-two scale, then round-half-even as the GPU's rintf). */
void oracle_lsh_quantize(const float* x, size_t count, int k, float* out) {
    const float sc = ldexpf(1.0f, k);
    for (size_t i = 0; i < count; i++) {
        float v = nearbyintf(x[i] * sc);
        out[i] = v < -255.0f ? -255.0f : (v > 255.0f ? 255.0f : v);
    }
}

/* The largest k in [-100, 100] with max|x| * 2^k <= 255 (0 for all zeros). */
int oracle_lsh_scale_exp(const float* x, size_t count) {
    double mx = 0.0;
    for (size_t i = 0; i < count; i++) mx = fabs((double)x[i]) > mx ? fabs((double)x[i]) : mx;
    if (!(mx > 0.0)) return 0;
    int k = -100;
    while (k < 100 && ldexp(mx, k + 1) <= 255.0) k++;
    return k;
}

static const uint32_t* g_sort_codes;
static int cmp_code_idx(const void* a, const void* b) {
    uint32_t ia = *(const uint32_t*)a, ib = *(const uint32_t*)b;
    uint32_t ca = g_sort_codes[ia], cb = g_sort_codes[ib];
    if (ca != cb) return ca < cb ? -1 : 1;
    return ia < ib ? -1 : (ia > ib);
}

void oracle_lsh_sort(const uint32_t* codes, int n, uint32_t* perm) {
    for (int i = 0; i < n; i++) perm[i] = (uint32_t)i;
    g_sort_codes = codes;
    qsort(perm, (size_t)n, sizeof(uint32_t), cmp_code_idx);
}

void oracle_lsh_query(const float* q, const float* r, int q_n, int r_n, int L,
                      const uint32_t* codes_q, const uint32_t* codes_r, uint32_t* idx2,
                      uint32_t* dist2, uint32_t* ncand, int threads) {
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
#endif
    for (int i = 0; i < q_n; i++) {
        uint64_t d1 = UINT64_MAX, d2 = UINT64_MAX;
        uint32_t i1 = UINT32_MAX, i2 = UINT32_MAX, nc = 0;
        for (int j = 0; j < r_n; j++) {
            int hit = 0;
            for (int t = 0; t < L && !hit; t++)
                hit = codes_q[(size_t)t * q_n + i] == codes_r[(size_t)t * r_n + j];
            if (!hit) continue;
            nc++;
            int64_t s = 0;
            for (int d = 0; d < 128; d++) {
                int64_t c = (int64_t)q[(size_t)i * 128 + d] - (int64_t)r[(size_t)j * 128 + d];
                s += c * c;
            }
            top2_insert((uint64_t)s, (uint32_t)j, &d1, &i1, &d2, &i2);
        }
        idx2[2 * (size_t)i] = i1;
        idx2[2 * (size_t)i + 1] = i2;
        dist2[2 * (size_t)i] = d1 == UINT64_MAX ? UINT32_MAX : (uint32_t)d1;
        dist2[2 * (size_t)i + 1] = d2 == UINT64_MAX ? UINT32_MAX : (uint32_t)d2;
        ncand[i] = nc;
    }
}

/* The float mode's re-rank: the same candidate rule, distances by
 * host_lenght (sequential FP32), top-2 by (distance, index). */
void oracle_lsh_query_f32(const float* q, const float* r, int q_n, int r_n, int L, const uint32_t* codes_q,
                          const uint32_t* codes_r, uint32_t* idx2, float* dist2, uint32_t* ncand, int threads) {
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
#endif
    for (int i = 0; i < q_n; i++) {
        float d1 = FLT_MAX, d2 = FLT_MAX;
        uint32_t i1 = UINT32_MAX, i2 = UINT32_MAX, nc = 0;
        for (int j = 0; j < r_n; j++) {
            int hit = 0;
            for (int t = 0; t < L && !hit; t++)
                hit = codes_q[(size_t)t * q_n + i] == codes_r[(size_t)t * r_n + j];
            if (!hit) continue;
            nc++;
            const float d = host_lenght(q + (size_t)i * 128, r + (size_t)j * 128);
            if (d < d1) {  /* rows come in increasing order: strict '<' keeps the lowest index */
                d2 = d1;
                i2 = i1;
                d1 = d;
                i1 = (uint32_t)j;
            } else if (d < d2) {
                d2 = d;
                i2 = (uint32_t)j;
            }
        }
        idx2[2 * (size_t)i] = i1;
        idx2[2 * (size_t)i + 1] = i2;
        dist2[2 * (size_t)i] = d1;
        dist2[2 * (size_t)i + 1] = d2;
        ncand[i] = nc;
    }
}

/* ------------------------------------------------------- learned hash -- */

void oracle_mlp_forward(const float* x, int n, const float* W1, const float* b1, int H,
                        const float* W2, const float* b2, int B, double* hid, double* out) {
    for (int i = 0; i < n; i++) {
        for (int h = 0; h < H; h++) {
            double s = b1[h];
            for (int d = 0; d < 128; d++) s += (double)W1[(size_t)h * 128 + d] * x[(size_t)i * 128 + d];
            hid[(size_t)i * H + h] = s > 0.0 ? s : 0.0;
        }
        for (int b = 0; b < B; b++) {
            double s = b2[b];
            for (int h = 0; h < H; h++) s += (double)W2[(size_t)b * H + h] * hid[(size_t)i * H + h];
            out[(size_t)i * B + b] = s;
        }
    }
}

/* ------------------------------------------------------ learned hash -- */

enum { HH = 256, HB = 32, P_W1 = 0, P_B1 = HH * 128, P_W2 = P_B1 + HH, P_B2 = P_W2 + HB * HH };

/* bfloat16 round to nearest even (finite inputs), as the library's packers. */
static float bf16_round(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return f;
}

static double pw(const float* p, int i, int quant) { return quant ? (double)bf16_round(p[i]) : (double)p[i]; }
/* layer 1 uses the stored weights times 1/64 (exact: a power of two) */
static double pw1(const float* p, int i, int quant) {
    const float w = p[i] * (1.f / 64.f);
    return quant ? (double)bf16_round(w) : (double)w;
}

void oracle_hash_forward(const float* params, int nbits, const float* x, int n, int quant,
                         double* hid, double* z) {
    for (int i = 0; i < n; i++) {
        for (int u = 0; u < HH; u++) {
            double s = params[P_B1 + u];
            for (int d = 0; d < 128; d++) {
                const double xv = quant ? (double)bf16_round(x[(size_t)i * 128 + d]) : (double)x[(size_t)i * 128 + d];
                s += pw1(params, P_W1 + u * 128 + d, quant) * xv;
            }
            double hv = s > 0.0 ? s : 0.0;
            if (quant) hv = (double)bf16_round((float)hv);
            hid[(size_t)i * HH + u] = hv;
        }
        for (int k = 0; k < HB; k++) {
            double s = 0.0;
            if (k < nbits) {
                s = params[P_B2 + k];
                for (int u = 0; u < HH; u++) s += pw(params, P_W2 + k * HH + u, quant) * hid[(size_t)i * HH + u];
            }
            z[(size_t)i * HB + k] = s;
        }
    }
}

double oracle_hash_loss(const double* s, const double* z, int B, int nbits, double alpha,
                        double beta, double gamma) {
    const int N = 2 * B;
    const double b = nbits;
    double sim = 0, bal = 0, dec = 0, q = 0;
    for (int i = 0; i < B; i++)
        for (int k = 0; k < nbits; k++) {
            const double d = s[(size_t)i * HB + k] - s[(size_t)(i + B) * HB + k];
            sim += d * d / 4.0;
        }
    for (int k = 0; k < nbits; k++) {
        double m = 0;
        for (int i = 0; i < N; i++) m += s[(size_t)i * HB + k];
        m /= N;
        bal += m * m;
        for (int l = 0; l < nbits; l++) {
            if (l == k) continue;
            double c = 0;
            for (int i = 0; i < N; i++) c += s[(size_t)i * HB + k] * s[(size_t)i * HB + l];
            c /= N;
            dec += c * c;
        }
    }
    for (int i = 0; i < N; i++)
        for (int k = 0; k < nbits; k++) {
            const double a = fabs(z[(size_t)i * HB + k]) - 1.0;
            q += a * a;
        }
    return sim / (B * b) + alpha * bal / b + beta * dec / (b * b) + gamma * q / (N * b);
}

double oracle_hash_loss_grad(const double* z, int B, int nbits, double alpha, double beta,
                             double gamma, double* dls, double* dz) {
    const int N = 2 * B;
    const double b = nbits;
    double* s = (double*)malloc(sizeof(double) * (size_t)N * HB);
    double* mean = (double*)calloc(HB, sizeof(double));
    double* cr = (double*)calloc(HB * HB, sizeof(double));
    for (int i = 0; i < N; i++)
        for (int k = 0; k < HB; k++) s[(size_t)i * HB + k] = z[(size_t)i * HB + k] >= 0.0 ? 1.0 : -1.0;
    for (int k = 0; k < nbits; k++) {
        for (int i = 0; i < N; i++) mean[k] += s[(size_t)i * HB + k];
        mean[k] /= N;
        for (int l = 0; l < nbits; l++) {
            double c = 0;
            for (int i = 0; i < N; i++) c += s[(size_t)i * HB + k] * s[(size_t)i * HB + l];
            cr[k * HB + l] = c / N;
        }
    }
    for (int i = 0; i < N; i++) {
        const int p = i < B ? i + B : i - B;
        for (int k = 0; k < HB; k++) {
            const size_t e = (size_t)i * HB + k;
            if (k >= nbits) {
                dls[e] = 0.0;
                dz[e] = 0.0;
                continue;
            }
            double g = (s[e] - s[(size_t)p * HB + k]) / (2.0 * B * b);
            g += alpha * 2.0 * mean[k] / (b * N);
            double dc = 0;
            for (int l = 0; l < nbits; l++)
                if (l != k) dc += cr[k * HB + l] * s[(size_t)i * HB + l];
            g += beta * 4.0 * dc / (b * b * N);
            dls[e] = g;
            const double zv = z[e];
            dz[e] = g + gamma * 2.0 * (fabs(zv) - 1.0) * s[e] / (N * b);
        }
    }
    const double L = oracle_hash_loss(s, z, B, nbits, alpha, beta, gamma);
    free(s);
    free(mean);
    free(cr);
    return L;
}

void oracle_hash_backward(const float* params, int nbits, const float* x, int n, int quant,
                          const double* hid, const double* dz, double* grads) {
    memset(grads, 0, sizeof(double) * (size_t)(P_B2 + HB));
    double* dh = (double*)malloc(sizeof(double) * HH);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < nbits; k++) {
            const double g = dz[(size_t)i * HB + k];
            grads[P_B2 + k] += g;
            for (int u = 0; u < HH; u++) grads[P_W2 + k * HH + u] += g * hid[(size_t)i * HH + u];
        }
        for (int u = 0; u < HH; u++) {
            double s = 0;
            for (int k = 0; k < nbits; k++) s += dz[(size_t)i * HB + k] * pw(params, P_W2 + k * HH + u, quant);
            dh[u] = hid[(size_t)i * HH + u] > 0.0 ? s : 0.0;
        }
        for (int u = 0; u < HH; u++) {
            if (dh[u] == 0.0) continue;
            grads[P_B1 + u] += dh[u];
            for (int d = 0; d < 128; d++) {
                const double xv = quant ? (double)bf16_round(x[(size_t)i * 128 + d]) : (double)x[(size_t)i * 128 + d];
                grads[P_W1 + u * 128 + d] += dh[u] * xv / 64.0;
            }
        }
    }
    free(dh);
}
