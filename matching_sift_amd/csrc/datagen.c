/*
 * datagen.c — seeded synthetic SIFT-like descriptors for the benchmark
 * workloads (host C, no GPU).
 *
 * Distribution of one descriptor (128 floats holding integers in [0,255]):
 *   each element: with probability 0.3 it is 0, otherwise
 *   rint(Exponential(mean 20)) clipped to 255.
 * A "noisy copy" of a descriptor adds rint(Normal(0, sd)) per element and
 * clips to [0,255].
 *
 * Random source: xoshiro256** seeded through splitmix64, uniforms in [0,1)
 * from the top 53 bits.  Exponential by inversion (-mean*log1p(-u)), normal by
 * Box-Muller (cosine branch only).  Rounding is rint() in the default
 * round-half-even mode.  Every element consumes exactly two uniforms, every
 * query one selector uniform plus one source-index uniform, so the streams are
 * reproducible element by element.
 *
 * The reference generates unit-norm rand() vectors instead (src/helper.cu:40-99);
 * the integer distribution here follows the benchmark configurations.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t* r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t* r) {
    uint64_t* s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static inline double rng_uniform(rng_t* r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

static inline float clip255(double v) {
    if (v < 0.0) return 0.0f;
    if (v > 255.0) return 255.0f;
    return (float)v;
}

static void gen_one(rng_t* r, float* out, double mean, double zero_p) {
    for (int d = 0; d < 128; d++) {
        double u_zero = rng_uniform(r);
        double u_val = rng_uniform(r);
        double v = rint(-mean * log1p(-u_val));
        out[d] = (u_zero < zero_p) ? 0.0f : clip255(v);
    }
}

static double normal(rng_t* r) {
    double u1 = rng_uniform(r);
    double u2 = rng_uniform(r);
    return sqrt(-2.0 * log1p(-u1)) * cos(6.283185307179586476925 * u2);
}

static void noisy_copy(rng_t* r, const float* src, float* out, double sd) {
    for (int d = 0; d < 128; d++) out[d] = clip255(rint((double)src[d] + sd * normal(r)));
}

/* n descriptors of the base distribution. */
void sift_gen_base(uint64_t seed, int64_t n, float* out) {
    rng_t r;
    rng_seed(&r, seed);
    for (int64_t i = 0; i < n; i++) gen_one(&r, out + i * 128, 20.0, 0.3);
}

/* m queries: with probability frac_copy a noisy copy (noise sd) of a uniformly
 * chosen base row, otherwise a fresh descriptor.  src[i] = copied base row or
 * -1 (src may be NULL). */
void sift_gen_queries(uint64_t seed, const float* base, int64_t n_base, int64_t m,
                      double frac_copy, double noise_sd, float* out, int64_t* src) {
    rng_t r;
    rng_seed(&r, seed ^ 0x5175657279ull);
    for (int64_t i = 0; i < m; i++) {
        double sel = rng_uniform(&r);
        double pick = rng_uniform(&r);
        if (n_base > 0 && sel < frac_copy) {
            int64_t j = (int64_t)(pick * (double)n_base);
            if (j >= n_base) j = n_base - 1;
            noisy_copy(&r, base + j * 128, out + i * 128, noise_sd);
            if (src) src[i] = j;
        } else {
            gen_one(&r, out + i * 128, 20.0, 0.3);
            if (src) src[i] = -1;
        }
    }
}

/* Keypoint counts of image pairs: log-normal with the given median and sigma,
 * rounded and clipped to [lo, hi].  counts[2*p] = image A, counts[2*p+1] = B. */
void sift_gen_pair_sizes(uint64_t seed, int64_t pairs, double median, double sigma, int64_t lo,
                         int64_t hi, int64_t* counts) {
    rng_t r;
    rng_seed(&r, seed ^ 0x5061697273ull);
    for (int64_t i = 0; i < 2 * pairs; i++) {
        double v = rint(median * exp(sigma * normal(&r)));
        if (v < (double)lo) v = (double)lo;
        if (v > (double)hi) v = (double)hi;
        counts[i] = (int64_t)v;
    }
}

/* Uniform-norm float vectors like the reference's make_rand_vector
 * (src/helper.cu:40-58): uniform entries in [0,1) normalised to unit L2. */
void sift_gen_unit(uint64_t seed, int64_t n, float* out) {
    rng_t r;
    rng_seed(&r, seed ^ 0x556e6974ull);
    for (int64_t i = 0; i < n; i++) {
        double v[128], s = 0.0;
        for (int d = 0; d < 128; d++) {
            v[d] = rng_uniform(&r);
            s += v[d] * v[d];
        }
        s = sqrt(s);
        for (int d = 0; d < 128; d++) out[i * 128 + d] = (float)(v[d] / s);
    }
}

/* Gaussian values (mean 0, sd) rounded to integers and clipped to [-lim, lim]. */
void sift_gen_gauss_int(uint64_t seed, int64_t n, double sd, double lim, float* out) {
    rng_t r;
    rng_seed(&r, seed ^ 0x4761757373ull);
    for (int64_t i = 0; i < n; i++) {
        double v = rint(sd * normal(&r));
        if (v < -lim) v = -lim;
        if (v > lim) v = lim;
        out[i] = (float)v;
    }
}

/* Raw uniforms (for host-side sampling in tests and benches). */
void sift_gen_uniform(uint64_t seed, int64_t n, double* out) {
    rng_t r;
    rng_seed(&r, seed ^ 0x556e69ull);
    for (int64_t i = 0; i < n; i++) out[i] = rng_uniform(&r);
}
