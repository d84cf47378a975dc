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
