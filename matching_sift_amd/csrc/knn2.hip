// knn2.hip — exact 2-NN host orchestration: seeding pass, fused-kernel
// launch (fused.hip), slab merge with the output formats, and the
// brute-force entry points (reference src/knn_brute.cu).
//
// The reference materialises a q x r distance matrix with cuBLAS and re-reads
// it in a top-2 kernel (src/knn_brute.cu:244 then :265; :592 then :606;
// :644 then :648); here the distance never reaches HBM and the only
// intermediate is a per-(slab, query) top-2 partial of 16 bytes.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "fused.h"
#include "rows.h"

namespace sift {
namespace {

struct MergeArgs {
    const uint4* part;
    const float* q_norm;  // natural order (int32 bits for KIND_I8)
    uint32_t q_n, r_n, q_stride;
    int n_slabs;
    uint32_t* idx;
    float* dist;
    sift_float4* sorted;
    uint32_t* matches;
    float ratio;
    int* bound_out;       // seeding mode: write the shared 2nd-best bound instead
    const int* coarse;    // float kinds: nonzero -> coarse keys (see Knn2Out::coarse)
    float slack;          // coarse keys: relative rounding bound of the kind
    ExactSrc src;         // coarse keys: the rows of the exact re-rank
};

// Coarse keys (float kinds on data that is not all integers in [0,255]):
// with each value rounded to the kind with relative error at most u (2^-8
// for BF16, 2^-11 for FP16, 0 for FP32), a key |b|^2 - 2 q.b of the MFMA pass
// differs from the exact one by at most (2u + u^2) (|b|^2 + 2 |q| |b|) plus
// the FP32 accumulation error of its 130 terms (130 * 2^-24 of their
// magnitudes), i.e. by at most rel * (|q| + |b|)^2 with rel these values
// (the bounds with a margin of about 2x).
constexpr float COARSE_REL_BF16 = 1.f / 64.f;
constexpr float COARSE_REL_F16 = 1.f / 512.f;
constexpr float COARSE_REL_F32 = 1.f / 65536.f;
inline float coarse_rel(int kind) {
    return kind == KIND_BF16 ? COARSE_REL_BF16 : (kind == KIND_F16 ? COARSE_REL_F16 : COARSE_REL_F32);
}

// Candidates of the exact re-rank of coarse keys, per query.
constexpr int EXACT_CAND = 16;

struct RerankArgs {
    const int* coarse;
    const uint4* part;
    uint32_t q_n, q_stride;
    int n_slabs;
    ExactSrc src;
    uint32_t* idx;
    float* dist;
    sift_float4* sorted;
    uint32_t* matches;
    float ratio;
};

// Exact re-rank of coarse keys for query q, one wave (the final merge runs it
// when the data were not all integers in [0,255]): the EXACT_CAND smallest
// coarse (key, row) over every slab partial (each lane a sorted list of its
// own entries, then wave-wide minimum rounds), the sequential FP32 distance
// of each candidate (one lane each), and the two smallest (distance, row) -
// the result of host_brute restricted to the candidates.
__device__ inline void exact_rerank_one(const RerankArgs& a, uint32_t q, int lane) {
    constexpr uint64_t NONE = ~0ull;
    uint64_t l[EXACT_CAND];
#pragma unroll
    for (int j = 0; j < EXACT_CAND; j++) l[j] = NONE;
    auto push = [&](uint32_t key_bits, uint32_t row) {
        if (row == 0xFFFFFFFFu) return;
        const int o = Key<KIND_F32>::ord(__uint_as_float(key_bits));
        uint64_t x = ((uint64_t)((uint32_t)o ^ 0x80000000u) << 32) | row;
#pragma unroll
        for (int j = 0; j < EXACT_CAND; j++) {  // insertion: l stays ascending
            const uint64_t lo = min(l[j], x);
            x = max(l[j], x);
            l[j] = lo;
        }
    };
    for (int s = lane; s < a.n_slabs; s += 64) {
        const uint4 p = a.part[(size_t)s * a.q_stride + q];
        push(p.x, p.y);
        push(p.z, p.w);
    }
    uint32_t cand = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < EXACT_CAND; r++) {
        uint64_t m = l[0];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) m = min(m, __shfl_xor(m, off));
        if (lane == r) cand = m == NONE ? 0xFFFFFFFFu : (uint32_t)m;
        const bool pop = l[0] == m && m != NONE;  // rows are unique: one lane holds m
#pragma unroll
        for (int j = 0; j < EXACT_CAND - 1; j++) l[j] = pop ? l[j + 1] : l[j];
        l[EXACT_CAND - 1] = pop ? NONE : l[EXACT_CAND - 1];
    }
    // distances are >= +0, so their bit patterns order as unsigned integers
    uint64_t m1 = NONE, m2 = NONE;
    if (lane < EXACT_CAND && cand != 0xFFFFFFFFu) {
        const size_t qrow = a.src.q_map ? (size_t)a.src.q_map[q] : q;
        const float d = seq_dist(a.src.q, a.src.q_type, qrow, a.src.r, a.src.r_type, cand);
        m1 = ((uint64_t)__float_as_uint(d) << 32) | cand;
    }
#pragma unroll
    for (int off = 1; off < EXACT_CAND; off <<= 1) {
        const uint64_t b1 = __shfl_xor(m1, off, EXACT_CAND), b2 = __shfl_xor(m2, off, EXACT_CAND);
        const uint64_t n2 = min(max(m1, b1), min(m2, b2));  // 2nd smallest of two sorted pairs
        m1 = min(m1, b1);
        m2 = n2;
    }
    if (lane != 0) return;  // lane 0 writes (the other lanes rejoin after it)
    const bool h1 = m1 != NONE, h2 = m2 != NONE;
    const float d1 = h1 ? __uint_as_float((uint32_t)(m1 >> 32)) : FLT_MAX;
    const float d2 = h2 ? __uint_as_float((uint32_t)(m2 >> 32)) : FLT_MAX;
    const uint32_t i1 = h1 ? (uint32_t)m1 : 0xFFFFFFFFu, i2 = h2 ? (uint32_t)m2 : 0xFFFFFFFFu;
    if (a.idx) {
        a.idx[2 * (size_t)q] = i1;
        a.idx[2 * (size_t)q + 1] = i2;
    }
    if (a.dist) {
        a.dist[2 * (size_t)q] = d1;
        a.dist[2 * (size_t)q + 1] = d2;
    }
    if (a.sorted) {
        sift_float4 o;
        o.x = d1;
        o.y = d2;
        o.z = h1 ? (float)i1 : FLT_MAX;
        o.w = h2 ? (float)i2 : FLT_MAX;
        a.sorted[q] = o;
    }
    if (a.matches) a.matches[q] = (h1 && sqrtf(d1) / sqrtf(d2) < a.ratio) ? i1 : 0xFFFFFFFFu;
}

// Lexicographic merge of the slab partials, then d = key + ||q||^2 and the
// output formats.  A block takes 64 queries; wave w merges slabs w, w + 8, ...
// for them (each load is 1 KB contiguous), and wave 0 merges the 8 results.
// In seeding mode the result is the query's shared bound: its 2nd-best key
// over the rows merged.
constexpr int MERGE_WAVES = 8;
// FP8 coarse pass: relative widening of the seed bounds (coarse distances
// differ from exact ones by a few percent for SIFT data).
constexpr float FP8_SLACK = 1.f / 16.f;

// FP8 coarse pass: the 16 smallest coarse (key, row) over every slab partial
// of a query (two entries each), one wave per query: each lane keeps a sorted
// top-16 of its own entries, then 16 rounds of a wave-wide minimum pop the
// overall order.  Rows come out ascending by coarse key; absent: UINT32_MAX.
__global__ __launch_bounds__(256) void merge_cand16(const uint4* __restrict__ part, uint32_t q_n,
                                                    uint32_t q_stride, int n_slabs, uint32_t* __restrict__ cand) {
    constexpr uint64_t NONE = ~0ull;
    const int lane = threadIdx.x & 63;
    const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= q_n) return;  // whole wave
    uint64_t l[16];
#pragma unroll
    for (int j = 0; j < 16; j++) l[j] = NONE;
    auto push = [&](uint32_t key_bits, uint32_t row) {
        if (row == 0xFFFFFFFFu) return;
        const int o = Key<KIND_FP8>::ord(__uint_as_float(key_bits));
        uint64_t x = ((uint64_t)((uint32_t)o ^ 0x80000000u) << 32) | row;
#pragma unroll
        for (int j = 0; j < 16; j++) {  // insertion: l stays ascending
            const uint64_t lo = min(l[j], x);
            x = max(l[j], x);
            l[j] = lo;
        }
    };
    for (int s = lane; s < n_slabs; s += 64) {
        const uint4 p = part[(size_t)s * q_stride + q];
        push(p.x, p.y);
        push(p.z, p.w);
    }
    uint32_t out = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        uint64_t m = l[0];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) m = min(m, __shfl_xor(m, off));
        if (lane == r) out = m == NONE ? 0xFFFFFFFFu : (uint32_t)m;
        const bool pop = l[0] == m && m != NONE;  // rows are unique: one lane holds m
#pragma unroll
        for (int j = 0; j < 15; j++) l[j] = pop ? l[j + 1] : l[j];
        l[15] = pop ? NONE : l[15];
    }
    if (lane < 16) cand[(size_t)q * 16 + lane] = out;
}
template <int KIND>
__global__ __launch_bounds__(64 * MERGE_WAVES) void merge_top2(MergeArgs a) {
    using key_t = typename Key<KIND>::T;
    __shared__ uint4 red[MERGE_WAVES - 1][64];
    const bool coarse = a.coarse != nullptr && *a.coarse != 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (coarse && !a.bound_out) {  // the exact re-rank of the block's 64 queries, one wave each
        RerankArgs ra{a.coarse, a.part, a.q_n, a.q_stride, a.n_slabs, a.src, a.idx, a.dist, a.sorted, a.matches, a.ratio};
        for (uint32_t i = 0; i < 64 / MERGE_WAVES; i++) {
            const uint32_t q = blockIdx.x * 64u + (uint32_t)w * (64 / MERGE_WAVES) + i;
            if (q < a.q_n) exact_rerank_one(ra, q, lane);
        }
        return;
    }
    const uint32_t q = blockIdx.x * 64u + lane;
    key_t m1 = Key<KIND>::none(), m2 = Key<KIND>::none();
    uint32_t i1 = 0xFFFFFFFFu, i2 = 0xFFFFFFFFu;
    if (q < a.q_n) {
        for (int s = w; s < a.n_slabs; s += MERGE_WAVES) {
            const uint4 p = a.part[(size_t)s * a.q_stride + q];
            merge2(m1, i1, m2, i2, __builtin_bit_cast(key_t, p.x), p.y, __builtin_bit_cast(key_t, p.z), p.w);
        }
    }
    if (w > 0) {
        red[w - 1][lane] = make_uint4(__builtin_bit_cast(uint32_t, m1), i1, __builtin_bit_cast(uint32_t, m2), i2);
    }
    __syncthreads();
    if (w > 0 || q >= a.q_n) return;
#pragma unroll
    for (int v = 0; v < MERGE_WAVES - 1; v++) {
        const uint4 p = red[v][lane];
        merge2(m1, i1, m2, i2, __builtin_bit_cast(key_t, p.x), p.y, __builtin_bit_cast(key_t, p.z), p.w);
    }
    const bool has1 = i1 < a.r_n, has2 = i2 < a.r_n;
    if (a.bound_out) {  // inclusive bound: the 2nd-best key over the rows merged
        key_t b = has2 ? m2 : Key<KIND>::none();
        if constexpr (KIND == KIND_FP8) {
            // coarse keys: widen by FP8_SLACK of the coarse distance so rows
            // whose e4m3 rounding pushed them past the bound still enter
            if (has2) b = b + FP8_SLACK * fabsf(b + a.q_norm[q]) + 1.f;
        } else if constexpr (KIND != KIND_I8) {
            // coarse keys: the 2nd-best exact key is at most b + e and a true
            // top-2 row's coarse key at most that + e, e = rel (|q| + |b|)^2
            // with |b| <= |q| + sqrt(d2)
            if (coarse && has2) {
                const float nq = fmaxf(a.q_norm[q], 0.f);
                const float s = 2.f * sqrtf(nq) + sqrtf(fmaxf(b + nq, 0.f));
                b = b + 2.f * a.slack * s * s;
            }
        }
        a.bound_out[q] = Key<KIND>::ord(b);
        return;
    }
    float d1 = FLT_MAX, d2 = FLT_MAX;
    if constexpr (KIND == KIND_I8) {
        const int nq = __float_as_int(a.q_norm[q]);
        if (has1) d1 = (float)(m1 + nq);
        if (has2) d2 = (float)(m2 + nq);
    } else {
        const float nq = a.q_norm[q];
        if (has1) d1 = fmaxf(m1 + nq, 0.f);
        if (has2) d2 = fmaxf(m2 + nq, 0.f);
    }
    if (!has1) i1 = 0xFFFFFFFFu;
    if (!has2) i2 = 0xFFFFFFFFu;
    if (a.idx) {
        a.idx[2 * (size_t)q] = i1;
        a.idx[2 * (size_t)q + 1] = i2;
    }
    if (a.dist) {
        a.dist[2 * (size_t)q] = d1;
        a.dist[2 * (size_t)q + 1] = d2;
    }
    if (a.sorted) {
        sift_float4 o;
        o.x = d1;
        o.y = d2;
        o.z = has1 ? (float)i1 : FLT_MAX;
        o.w = has2 ? (float)i2 : FLT_MAX;
        a.sorted[q] = o;
    }
    if (a.matches) {
        // IEEE sqrt and division (correctly rounded by default in HIP)
        const bool pass = has1 && (sqrtf(d1) / sqrtf(d2) < a.ratio);
        a.matches[q] = pass ? i1 : 0xFFFFFFFFu;
    }
}

// Seeding.  The base is processed as a chain of launches over growing tile
// ranges [0, e0), [e0, e1), ..., [e_last, n): the first runs unbounded, and
// each later one starts every lane from the bound B = the query's 2nd-best key
// over all rows before its range (a merge of the earlier launches' partials),
// so fewer rows pass the admission test and the slow path runs less.  Every
// launch's partials stay in the final merge: no row is scanned twice.
// Default ends: 128 tiles (4,096 rows), then x8 while the end is at most a
// third of the base.  SIFT_SEED="e0,e1,..." (tile counts) overrides, "0"
// disables seeding.
int seed_ends(int n_tiles, int* ends, int max_ends) {
    int n = 0;
    if (const char* env = std::getenv("SIFT_SEED")) {
        const char* c = env;
        while (*c && n < max_ends) {
            char* e = nullptr;
            const long v = std::strtol(c, &e, 10);
            if (e == c) break;
            if (v <= 0) return 0;
            if (v < n_tiles && (n == 0 || v > ends[n - 1])) ends[n++] = (int)v;
            c = (*e == ',') ? e + 1 : e;
            if (*e != ',') break;
        }
        return n;
    }
    for (long e = 128; 3 * e <= n_tiles && n < max_ends; e *= 8) ends[n++] = (int)e;
    return n;
}

// Slab count for one launch over `n_tiles` tiles: blocks run in rounds of
// `cap` resident blocks; pick the count whose block total fills its last
// round best, preferring 2-8 rounds, counting a block's start-up (query
// fragments, pipeline fill) as a few tiles.  Slabs are a multiple of 8 (one
// XCD each).
void choose_slabs(int n_tiles, int n_groups, int cap, int tps, int* n_slabs, int* tps_slab) {
    constexpr double START_TILES = 6;
    const int max_slabs = std::max(8, std::min(1024, ((n_tiles / tps + 7) / 8) * 8));
    int best_ns = 8;
    double best = 1e30;
    for (int ns = 8; ns <= max_slabs; ns += 8) {
        const int tps_s = ((n_tiles + ns - 1) / ns + tps - 1) / tps * tps;
        const int used = (n_tiles + tps_s - 1) / tps_s;        // non-empty slabs
        const long items = (long)n_groups * used;
        const long rounds = (items + cap - 1) / cap;
        double cost = (double)rounds * (tps_s + START_TILES);  // wall time ~ rounds x slab length
        if (rounds < 2) cost *= 1.02;                          // keep some dynamic balance
        cost *= 1.0 + 0.0005 * ns / 8;                         // the merge grows with slabs
        if (cost < best) {
            best = cost;
            best_ns = ns;
        }
    }
    *n_slabs = best_ns;
    *tps_slab = std::max(tps, ((n_tiles + best_ns - 1) / best_ns + tps - 1) / tps * tps);
}

// The seeded chain over a packed base: launch l covers tiles
// [t0[l], t0[l + 1]) with shape shp[l] (groups[l] query groups) and writes its
// partials from slab p0[l]; every launch uses the common query stride.
struct ChainPlan {
    static constexpr int MAX_LAUNCH = 8;
    int n_launch = 0;
    uint32_t q_stride = 0;
    const FusedShape* shp[MAX_LAUNCH];
    int groups[MAX_LAUNCH];
    int t0[MAX_LAUNCH + 1], slabs[MAX_LAUNCH], tps_slab[MAX_LAUNCH], p0[MAX_LAUNCH + 1];
    // guided slabs (launches after the first): region and chunk schedule
    int grid[MAX_LAUNCH], region[MAX_LAUNCH], n_chunks[MAX_LAUNCH];
    int chunk_start[MAX_LAUNCH][MAX_CHUNKS + 1];
    size_t part_bytes() const { return (size_t)p0[n_launch] * q_stride * sizeof(uint4); }
    size_t bound_bytes() const { return (size_t)q_stride * sizeof(int); }
};

// Guided slabs.  A launch after the first (the seeded ones) splits its tile
// range into 8 regions, one per XCD, and each region into chunks that shrink
// towards its end; the blocks of chunk c run after those of chunk c - 1 (the
// dispatcher starts blocks in index order), so each XCD ends on short blocks
// instead of a last round of full slabs.  Measured on the 1M x 10k BF16 chain
// (same box, alternating runs): 1.733-1.737 ms per step for the four
// launches against 1.759-1.772 ms with uniform slabs; a per-block timeline of
// the main launch put the CU slots' idle share at 3.7% against 7.5-8.8%.
// Chunk sizes for one region of `region` tiles, `groups` blocks per chunk and
// `slots` resident blocks per XCD: each chunk's blocks take about 1/K of the
// work left per slot (G * R / (K * slots) tiles for R tiles left), between
// SMIN and SMAX tiles and a multiple of the stage depth.  More than
// MAX_CHUNKS chunks raise SMIN until the schedule fits.  SIFT_GUIDE="K,SMIN,
// SMAX" overrides the defaults (1.5, 16, 768); SIFT_GUIDE=0 keeps uniform slabs.
bool guided_off() {
    const char* env = std::getenv("SIFT_GUIDE");
    return env && env[0] == '0' && env[1] == 0;
}
int guided_schedule(int region, int groups, int slots, int tps, int* starts) {
    double K = 1.5;
    int smin = 16, smax = 768;
    if (const char* env = std::getenv("SIFT_GUIDE")) {
        double k = 0;
        int a = 0, b = 0;
        if (std::sscanf(env, "%lf,%d,%d", &k, &a, &b) == 3 && k > 0 && a > 0 && b >= a) {
            K = k;
            smin = a;
            smax = b;
        }
    }
    for (;; smin *= 2) {
        int n = 0, pos = 0;
        starts[0] = 0;
        while (pos < region && n < MAX_CHUNKS) {
            const int left = region - pos;
            int sz = (int)std::ceil((double)groups * left / (K * std::max(1, slots)));
            sz = std::max(smin, std::min(smax, sz));
            sz = (sz + tps - 1) / tps * tps;
            pos += std::min(sz, left);
            starts[++n] = pos;
        }
        if (pos >= region) return n;
    }
}

// The first launch of a chain runs unbounded over few tiles (128), so its
// blocks are few and each wave spends most of its time in the slow path.
// There the kind's listed shape with one query block per wave (same waves and
// stage depth), where there is one, runs twice the blocks at half the work
// each, with more waves per SIMD: 25 against 33 us on the 1M x 10k BF16
// chain.  Later launches keep the chain's shape (one query block was no
// faster there), and so does I8, whose chain shape already holds three
// blocks per CU (its one-block shape ran that launch in 28 against 26 us).
// SIFT_SEED_SHAPE="qw,waves,tps" overrides it with another listed shape.
const FusedShape* first_launch_shape(const FusedShape& sh) {
    int qw = 1, waves = sh.waves, tps = sh.tps;
    if (const char* env = std::getenv("SIFT_SEED_SHAPE")) {
        if (std::sscanf(env, "%d,%d,%d", &qw, &waves, &tps) != 3) return &sh;
    }
    const FusedShape* cand[16];
    const int nc = shapes_of_kind(sh.kind, cand, 16);
    for (int i = 0; i < nc; i++)
        if (cand[i]->qw == qw && cand[i]->waves == waves && cand[i]->tps == tps) return cand[i];
    return &sh;
}

ChainPlan plan_chain(const FusedShape& sh, uint32_t n_tiles, uint32_t q_n, uint32_t r_n) {
    ChainPlan c;
    const FusedShape* first = first_launch_shape(sh);
    c.n_launch = 1 + (r_n >= 2 ? seed_ends((int)n_tiles, c.t0 + 1, ChainPlan::MAX_LAUNCH - 1) : 0);
    c.t0[0] = 0;
    c.t0[c.n_launch] = (int)n_tiles;
    c.p0[0] = 0;
    for (int l = 0; l < c.n_launch; l++) {
        const int nt = c.t0[l + 1] - c.t0[l];
        const FusedShape* s = (c.n_launch > 1 && l == 0) ? first : &sh;
        const int qpg = s->qw * s->waves * 32;
        c.shp[l] = s;
        c.groups[l] = (int)((q_n + qpg - 1) / qpg);
        c.q_stride = std::max(c.q_stride, (uint32_t)c.groups[l] * qpg);
        const int cap = device_cu_count() * s->blocks_per_cu();
        if (l > 0 && !guided_off()) {
            c.region[l] = (nt + 7) / 8;
            c.n_chunks[l] = guided_schedule(c.region[l], c.groups[l], std::max(1, cap / 8), s->tps, c.chunk_start[l]);
            c.slabs[l] = 8 * c.n_chunks[l];
            c.grid[l] = c.slabs[l] * c.groups[l];
            c.tps_slab[l] = 0;
        } else {
            c.n_chunks[l] = 0;
            choose_slabs(nt, c.groups[l], cap, s->tps, &c.slabs[l], &c.tps_slab[l]);
            c.grid[l] = c.slabs[l] * c.groups[l];
        }
        c.p0[l + 1] = c.p0[l] + c.slabs[l];
    }
    return c;
}

int launch_fused_shape(sift_context* h, const FusedShape& sh, const PackedSet& qs,
                       const PackedSet& bs, uint32_t q_n, uint32_t r_n, const Knn2Out& out,
                       const Knn2Work* work) {
    const ChainPlan c = plan_chain(sh, bs.ntiles, q_n, r_n);
    if (c.q_stride > qs.rows_pad)
        return set_error(h, SIFT_ERR_INVALID, "packed query buffer smaller than the launch shape");
    const int n_launch = c.n_launch;
    const uint32_t q_stride = c.q_stride;
    const int* t0 = c.t0;
    const int* slabs = c.slabs;
    const int* tps_slab = c.tps_slab;
    const int* p0 = c.p0;
    uint4* part = nullptr;
    int* bound = nullptr;
    if (work) {  // caller scratch: no allocation
        if (work->part_bytes < c.part_bytes() || work->bound_bytes < c.bound_bytes())
            return set_error(h, SIFT_ERR_INVALID, "knn2 workspace smaller than the launch needs");
        part = static_cast<uint4*>(work->part);
        bound = work->bound;
    } else {
        SIFT_TRY(reserve(h, h->part, c.part_bytes()));
        SIFT_TRY(reserve(h, h->thresh, c.bound_bytes()));
        part = static_cast<uint4*>(h->part.p);
        bound = static_cast<int*>(h->thresh.p);
    }
    const bool float_kind = bs.kind == KIND_BF16 || bs.kind == KIND_F16 || bs.kind == KIND_F32;
    const int* coarse = float_kind ? out.coarse : nullptr;
    if (coarse && (!out.src.q || (r_n && !out.src.r)))
        return set_error(h, SIFT_ERR_INVALID, "coarse keys need the rows for the exact re-rank");

    auto mk_fused = [&](int l) {
        FusedArgs fa{};
        fa.q_tiles = static_cast<const uint4*>(qs.tiles.p);
        fa.b_tiles = static_cast<const uint4*>(bs.tiles.p);
        // I8: the MFMAs start at c >> 1; the slow path adds c to twice the product
        fa.b_nl = static_cast<const float*>(bs.kind == KIND_I8 ? bs.norm_half.p : bs.norm_lane.p);
        fa.b_nlc = static_cast<const int*>(bs.norm_lane.p);
        fa.part = part + (size_t)p0[l] * q_stride;
        fa.bound = l > 0 ? bound : nullptr;
        fa.q_n = q_n;
        fa.r_n = r_n;
        fa.q_stride = q_stride;
        fa.tile0 = t0[l];
        fa.end_tile = t0[l + 1];
        fa.tiles_per_slab = tps_slab[l];
        fa.n_slabs = slabs[l];
        fa.n_groups = c.groups[l];
        fa.wait_polls = h->wait_polls;
        fa.status = wait_status(h);
        fa.coarse = coarse;
        if (c.n_chunks[l] > 0) {
            fa.region = c.region[l];
            fa.n_chunks = c.n_chunks[l];
            for (int i = 0; i <= c.n_chunks[l]; i++) fa.chunk_start[i] = c.chunk_start[l][i];
        }
        return fa;
    };
    auto mk_merge = [&](int n_part) {
        MergeArgs ma;
        ma.part = part;
        ma.q_norm = static_cast<const float*>(qs.norm.p);
        ma.q_n = q_n;
        ma.r_n = r_n;
        ma.q_stride = q_stride;
        ma.n_slabs = n_part;
        ma.idx = out.idx;
        ma.dist = out.dist;
        ma.sorted = out.sorted;
        ma.matches = out.matches;
        ma.ratio = out.ratio;
        ma.bound_out = nullptr;
        ma.coarse = coarse;
        ma.slack = coarse_rel(bs.kind);
        ma.src = out.src;
        return ma;
    };
    auto launch_merge = [&](const MergeArgs& ma) {
        const unsigned g = (q_n + 63) / 64, b = 64 * MERGE_WAVES;
        switch (bs.kind) {
            case KIND_BF16: merge_top2<KIND_BF16><<<g, b, 0, h->stream>>>(ma); break;
            case KIND_F16: merge_top2<KIND_F16><<<g, b, 0, h->stream>>>(ma); break;
            case KIND_F32: merge_top2<KIND_F32><<<g, b, 0, h->stream>>>(ma); break;
            case KIND_FP8: merge_top2<KIND_FP8><<<g, b, 0, h->stream>>>(ma); break;
            default: merge_top2<KIND_I8><<<g, b, 0, h->stream>>>(ma); break;
        }
    };

    for (int l = 0; l < n_launch; l++) {
        if (l > 0) {  // bound from everything before this range
            ProfScope ps(h, PROF_SEED);
            MergeArgs ma = mk_merge(p0[l]);
            ma.bound_out = bound;
            launch_merge(ma);
            SIFT_HIP_TRY(h, hipGetLastError());
        }
        ProfScope ps(h, PROF_FUSED);
        c.shp[l]->fn(mk_fused(l), (unsigned)c.grid[l], h->stream);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    if (bs.kind == KIND_FP8) {  // 16 coarse candidates per query, exact re-rank
        if (!qs.u8.p || !bs.u8.p) return set_error(h, SIFT_ERR_INVALID, "FP8 sets need their uint8 copies");
        ProfScope ps(h, PROF_FP8);
        SIFT_TRY(reserve(h, h->cand, (size_t)q_n * 16 * 4));
        auto* cand = static_cast<uint32_t*>(h->cand.p);
        merge_cand16<<<(q_n + 3) / 4, 256, 0, h->stream>>>(part, q_n, q_stride, p0[n_launch], cand);
        SIFT_HIP_TRY(h, hipGetLastError());
        return fp8_rerank_launch(h, cand, qs.u8.p, static_cast<const int*>(qs.n2.p), bs.u8.p,
                                 static_cast<const int*>(bs.n2.p), nullptr, nullptr, q_n, out.idx, out.dist,
                                 out.sorted, out.matches, out.ratio);
    }
    {
        ProfScope ps(h, PROF_MERGE);
        launch_merge(mk_merge(p0[n_launch]));
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    return SIFT_OK;
}

}  // namespace

// Bases of at least this many tiles pick their launch shape by timing: the
// first call of a kind on a handle runs the kind's first listed shape once
// untimed (first-touch costs), then each of its first two listed shapes
// TUNE_ROUNDS times in A B B A order (every shape gives identical results),
// and keeps the one with the lower mean.  A first call finds the GPU's clock
// still rising (by 10% over the first 50 ms of load): with a fastest-run rule
// or A B A B order the later runs' shape won on the clock alone.
// The best shape differs between boxes by a few percent with no static rule,
// except for I8: its first shape, (4, 4, 2), ran the 1M x 10k chain 3-4%
// faster than (2, 4, 2) in sustained runs on every box measured, while the
// first-call timing picked (2, 4, 2), so I8 is not timed.
constexpr uint32_t TUNE_MIN_TILES = 8192;
constexpr int TUNE_ROUNDS = 4;
int tune_candidates(int kind) {
    const FusedShape* cand[16];
    return kind == KIND_I8 ? 1 : std::min(2, shapes_of_kind(kind, cand, 16));
}

const FusedShape* knn2_shape(sift_context* h, int kind, uint32_t n_tiles) {
    const FusedShape* sh = pick_shape(kind);
    if (!sh || std::getenv("SIFT_SHAPE") || n_tiles < TUNE_MIN_TILES || kind < 0 || kind >= 8) return sh;
    const FusedShape* cand[16];
    const int nc = std::min(tune_candidates(kind), shapes_of_kind(kind, cand, 16));
    return (h->tuned_shape[kind] >= 0 && h->tuned_shape[kind] < nc) ? cand[h->tuned_shape[kind]] : sh;
}

size_t knn2_work_bytes(sift_context* h, int kind, uint32_t n_tiles, uint32_t q_n, uint32_t r_n, size_t* bound_bytes) {
    const FusedShape* sh = knn2_shape(h, kind, n_tiles);
    if (!sh) return 0;
    const ChainPlan c = plan_chain(*sh, n_tiles, q_n, r_n);
    if (bound_bytes) *bound_bytes = c.bound_bytes();
    return c.part_bytes();
}

int knn2_packed(sift_context* h, const PackedSet& qs, const PackedSet& bs, uint32_t q_n,
                uint32_t r_n, const Knn2Out& out, const Knn2Work* work) {
    if (qs.kind != bs.kind) return set_error(h, SIFT_ERR_INVALID, "query/base packed with different kinds");
    if (q_n == 0) return SIFT_OK;
    const FusedShape* sh = pick_shape(bs.kind);
    if (!sh) return set_error(h, SIFT_ERR_INVALID, "bad kind");
    const int kind = bs.kind;
    // caller scratch: the shape already tuned on this handle, or the default; never timed here
    if (work) return launch_fused_shape(h, *knn2_shape(h, kind, bs.ntiles), qs, bs, q_n, r_n, out, work);
    if (std::getenv("SIFT_SHAPE") || bs.ntiles < TUNE_MIN_TILES || kind < 0 || kind >= 8)
        return launch_fused_shape(h, *sh, qs, bs, q_n, r_n, out, nullptr);
    const FusedShape* cand[16];
    const int nc = std::min(tune_candidates(kind), shapes_of_kind(kind, cand, 16));
    if (nc <= 1) return launch_fused_shape(h, *sh, qs, bs, q_n, r_n, out, nullptr);
    if (h->tuned_shape[kind] >= 0 && h->tuned_shape[kind] < nc)
        return launch_fused_shape(h, *cand[h->tuned_shape[kind]], qs, bs, q_n, r_n, out, nullptr);
    hipEvent_t e[2];
    SIFT_HIP_TRY(h, hipEventCreate(&e[0]));
    if (hipEventCreate(&e[1]) != hipSuccess) {
        (void)hipEventDestroy(e[0]);
        return set_error(h, SIFT_ERR_HIP, "hipEventCreate");
    }
    int best = 0, rc = launch_fused_shape(h, *cand[0], qs, bs, q_n, r_n, out, nullptr);
    float sum_ms[16];
    for (int i = 0; i < nc; i++) sum_ms[i] = 0.f;
    for (int r = 0; r < TUNE_ROUNDS && rc == SIFT_OK; r++) {
        for (int k = 0; k < nc && rc == SIFT_OK; k++) {
            const int i = (r & 1) ? nc - 1 - k : k;  // A B, B A, ...: unbiased under a linear ramp
            (void)hipEventRecord(e[0], h->stream);
            rc = launch_fused_shape(h, *cand[i], qs, bs, q_n, r_n, out, nullptr);
            (void)hipEventRecord(e[1], h->stream);
            if (rc != SIFT_OK) break;
            float ms = 0.f;
            if (hipEventSynchronize(e[1]) != hipSuccess || hipEventElapsedTime(&ms, e[0], e[1]) != hipSuccess) {
                rc = set_error(h, SIFT_ERR_HIP, "shape timing");
                break;
            }
            sum_ms[i] += ms;
        }
    }
    for (int i = 1; i < nc; i++)
        if (sum_ms[i] < sum_ms[best]) best = i;
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    if (rc == SIFT_OK) h->tuned_shape[kind] = best;  // the last launch already wrote the outputs
    return rc;
}

uint32_t query_rows_pad(int kind, uint32_t q_n) {
    (void)kind;
    return std::max<uint32_t>(1, (q_n + QPAD - 1) / QPAD) * QPAD;
}

}  // namespace sift

using namespace sift;

namespace sift {

bool valid_type(int type) { return type >= SIFT_FLOAT_HOST && type <= SIFT_U8_DEVICE; }

// Resolve SIFT_COMPUTE_* to a packed kind.  AUTO: integer data in [0,255] ->
// I8 (exact, and the fastest exact kind: int8 MFMA at twice the BF16 rate),
// otherwise `fallback` (F32 for the FP32 entry points, F16 for the fp16 one).
// Reads one device flag (a 4-byte synchronous copy).
int resolve_kind(sift_context* h, const void* q, const void* r, int type, uint32_t q_n,
                 uint32_t r_n, int compute, int fallback, int* kind) {
    switch (compute) {
        case SIFT_COMPUTE_BF16: *kind = KIND_BF16; return SIFT_OK;
        case SIFT_COMPUTE_F16: *kind = KIND_F16; return SIFT_OK;
        case SIFT_COMPUTE_F32: *kind = KIND_F32; return SIFT_OK;
        case SIFT_COMPUTE_I8: *kind = KIND_I8; return SIFT_OK;
        case SIFT_COMPUTE_FP8: *kind = KIND_FP8; return SIFT_OK;
        case SIFT_COMPUTE_AUTO: break;
        default: return set_error(h, SIFT_ERR_INVALID, "unknown compute kind " + std::to_string(compute));
    }
    if (type == SIFT_U8_DEVICE) {
        *kind = KIND_I8;
        return SIFT_OK;
    }
    SIFT_TRY(reserve(h, h->flag, 16));
    int* flag = static_cast<int*>(h->flag.p);
    SIFT_HIP_TRY(h, hipMemsetAsync(flag, 0, 4, h->stream));
    SIFT_TRY(check_exact(h, q, type, q_n, flag));
    SIFT_TRY(check_exact(h, r, type, r_n, flag));
    int bad = 0;
    SIFT_HIP_TRY(h, hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    *kind = bad ? fallback : KIND_I8;
    return SIFT_OK;
}

}  // namespace sift

namespace {

int knn2_impl(sift_context* h, const void* q, const void* r, int type, uint32_t q_n, uint32_t r_n,
              int compute, int fallback, const Knn2Out& out_in, bool blocking, const char* what) {
    if (!h) return set_error(nullptr, SIFT_ERR_INVALID, "null handle");
    if (!valid_type(type)) return set_error(h, SIFT_ERR_INVALID, "unknown input type " + std::to_string(type));
    if (q_n == 0) return SIFT_OK;
    if ((!q) || (r_n > 0 && !r)) return set_error(h, SIFT_ERR_INVALID, "null input pointer");
    if (r_n >= (1u << 24)) return set_error(h, SIFT_ERR_UNSUPPORTED, "r_n must be < 2^24 (indices exact in float4 output)");
    if (type == SIFT_FLOAT_HOST) {
        SIFT_TRY(stage_host_input(h, q, q_n, h->stage_q, &q));
        SIFT_TRY(stage_host_input(h, r, r_n, h->stage_r, &r));
        type = SIFT_FLOAT_DEVICE;
    }
    int kind = 0;
    SIFT_TRY(resolve_kind(h, q, r, type, q_n, r_n, compute, fallback, &kind));
    int* flag = static_cast<int*>(h->flag.p);
    SIFT_HIP_TRY(h, hipMemsetAsync(flag, 0, 4, h->stream));
    const uint32_t rows_b = std::max<uint32_t>(32, ((r_n + 31) / 32) * 32);
    SIFT_TRY(pack_side(h, r, type, r_n, rows_b, kind, PACK_BASE, h->b_set, flag));
    SIFT_TRY(pack_side(h, q, type, q_n, query_rows_pad(kind, q_n), kind, PACK_QUERY, h->q_set, flag));
    // float kinds: data that is not all integers in [0,255] (flagged by the
    // packs) is ranked coarsely and re-ranked exactly from the input rows
    Knn2Out out = out_in;
    out.coarse = flag;
    out.src = ExactSrc{q, type, r, type};
    SIFT_TRY(knn2_packed(h, h->q_set, h->b_set, q_n, r_n, out));
    if (blocking) {
        int fl[3] = {0, 0, 0};
        SIFT_HIP_TRY(h, hipMemcpyAsync(fl, flag, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
        SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
        SIFT_TRY(wait_error(h, fl[2], what));
        if (fl[0] && (kind == KIND_I8 || kind == KIND_FP8))
            return set_error(h, SIFT_ERR_NOT_EXACT, kind == KIND_I8 ? "SIFT_COMPUTE_I8 needs integer data in [0,255]"
                                                                    : "SIFT_COMPUTE_FP8 needs integer data in [0,255]");
    }
    return SIFT_OK;
}

std::mutex g_default_mu;
sift_context* g_default[64] = {};

sift_context* default_handle() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default[dev]) sift_create(&g_default[dev]);
    return g_default[dev];
}

}  // namespace

extern "C" {

int sift_knn2(sift_handle_t h, const void* q, const void* r, int type, uint32_t q_n, uint32_t r_n,
              int compute, uint32_t* idx, float* dist, sift_float4* sorted, uint32_t* matches,
              float ratio) {
    Knn2Out o{idx, dist, sorted, matches, ratio};
    return knn2_impl(h, q, r, type, q_n, r_n, compute, KIND_F32, o, true, "sift_knn2");
}

int sift_knn2_shape(sift_handle_t h, int compute, uint32_t r_n, int* qw, int* waves, int* tps) {
    if (!h || !qw || !waves || !tps) return set_error(h, SIFT_ERR_INVALID, "sift_knn2_shape: null argument");
    if (compute == SIFT_COMPUTE_AUTO)
        return set_error(h, SIFT_ERR_INVALID, "sift_knn2_shape: name the compute kind (AUTO depends on the data)");
    int kind = 0;
    SIFT_TRY(resolve_kind(h, nullptr, nullptr, SIFT_FLOAT_DEVICE, 0, 0, compute, KIND_F32, &kind));
    const FusedShape* sh = knn2_shape(h, kind, (r_n + 31) / 32);
    if (!sh) return set_error(h, SIFT_ERR_INVALID, "sift_knn2_shape: no shape for this kind");
    *qw = sh->qw;
    *waves = sh->waves;
    *tps = sh->tps;
    return SIFT_OK;
}

int sift_knn2_async(sift_handle_t h, const void* q, const void* r, int type, uint32_t q_n,
                    uint32_t r_n, int compute, uint32_t* idx, float* dist, sift_float4* sorted,
                    uint32_t* matches, float ratio) {
    Knn2Out o{idx, dist, sorted, matches, ratio};
    return knn2_impl(h, q, r, type, q_n, r_n, compute, KIND_F32, o, false, "sift_knn2_async");
}

// src/knn_brute.cu:24-209.  fp16 semantics for non-integer data (the
// reference converts to half); exact int8 MFMA for integer SIFT data.
int cublas_2nn_sift(void* q_points, void* r_points, int type, uint32_t q_n, uint32_t r_n,
                    uint32_t* matches, float threshold, sift_handle_t handle, int stream_n,
                    size_t use, int warps) {
    (void)stream_n;
    (void)use;
    (void)warps;
    if (!matches && q_n) return set_error(handle, SIFT_ERR_INVALID, "null matches");
    int compute = SIFT_COMPUTE_AUTO;
    if (type == SIFT_HALF_DEVICE) compute = SIFT_COMPUTE_F16;
    if (type == SIFT_BF16_DEVICE) compute = SIFT_COMPUTE_BF16;
    Knn2Out o;
    o.ratio = threshold;
    // host-memory matches (FLOAT_HOST callers) go through a device buffer
    if (type == SIFT_FLOAT_HOST && handle) {
        if (!device_accessible(matches)) {
            DevBuf tmp;
            SIFT_TRY(reserve(handle, tmp, (size_t)q_n * 4));
            o.matches = static_cast<uint32_t*>(tmp.p);
            int rc = knn2_impl(handle, q_points, r_points, type, q_n, r_n, compute, KIND_F16, o, true, "cublas_2nn_sift");
            if (rc == SIFT_OK)
                rc = hipMemcpy(matches, tmp.p, (size_t)q_n * 4, hipMemcpyDeviceToHost) == hipSuccess
                         ? SIFT_OK
                         : set_error(handle, SIFT_ERR_HIP, "copy matches to host");
            release(tmp);
            return rc;
        }
    }
    o.matches = matches;
    return knn2_impl(handle, q_points, r_points, type, q_n, r_n, compute, KIND_F16, o, true, "cublas_2nn_sift");
}

// src/knn_brute.cu:213-278.  R is fp16 on the device; Q and dist are unused.
int cublas_2nn_sift_batch(void* q_points, void* Q, int type, void* R, uint32_t q_n, uint32_t r_n,
                          void* dist, uint32_t* matches, float threshold, sift_handle_t handle,
                          void* stream, int warps) {
    (void)Q;
    (void)dist;
    (void)warps;
    if (!handle) return set_error(nullptr, SIFT_ERR_INVALID, "null handle");
    if (type != SIFT_HALF_DEVICE && type != SIFT_FLOAT_DEVICE)
        return set_error(handle, SIFT_ERR_UNSUPPORTED, "cublas_2nn_sift_batch: q must be float or half on the device");
    if (q_n == 0) return SIFT_OK;
    if (!q_points || (r_n && !R) || !matches) return set_error(handle, SIFT_ERR_INVALID, "null pointer");
    if (r_n >= (1u << 24)) return set_error(handle, SIFT_ERR_UNSUPPORTED, "r_n must be < 2^24");
    // work queued on the caller's stream (e.g. the conversion of R) comes first
    if (stream && stream != (void*)handle->stream)
        SIFT_HIP_TRY(handle, hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    // R arrives as fp16: FP16 MFMAs over R and q rounded to fp16; data that is
    // not all integers in [0,255] is re-ranked exactly from R and q as given
    int* flag = static_cast<int*>(handle->flag.p);
    SIFT_HIP_TRY(handle, hipMemsetAsync(flag, 0, 4, handle->stream));
    const uint32_t rows_b = std::max<uint32_t>(32, ((r_n + 31) / 32) * 32);
    SIFT_TRY(pack_rows(handle, R, SIFT_HALF_DEVICE, r_n, rows_b, KIND_F16, PACK_BASE, handle->b_set, flag));
    SIFT_TRY(pack_rows(handle, q_points, type, q_n, query_rows_pad(KIND_F16, q_n), KIND_F16, PACK_QUERY,
                       handle->q_set, flag));
    Knn2Out o;
    o.matches = matches;
    o.ratio = threshold;
    o.coarse = flag;
    o.src = ExactSrc{q_points, type, R, SIFT_HALF_DEVICE};
    SIFT_TRY(knn2_packed(handle, handle->q_set, handle->b_set, q_n, r_n, o));
    return check_wait(handle, "cublas_2nn_sift_batch");
}

// src/knn_brute.cu:619-667: exact FP32 2-NN into float4.
void device_brute(const des_t_f* q_points, const des_t_f* r_points, int q_n, int r_n, void* sorted) {
    sift_context* h = default_handle();
    if (!h) {
        set_error(nullptr, SIFT_ERR_HIP, "device_brute: no default handle");
        return;
    }
    if (q_n < 0 || r_n < 0) {
        set_error(h, SIFT_ERR_INVALID, "negative size");
        return;
    }
    Knn2Out o;
    o.sorted = static_cast<sift_float4*>(sorted);
    int rc = knn2_impl(h, q_points, r_points, SIFT_FLOAT_DEVICE, (uint32_t)q_n, (uint32_t)r_n,
                       SIFT_COMPUTE_AUTO, KIND_F32, o, true, "device_brute");
    if (rc == SIFT_OK) set_error(h, SIFT_OK, "");
}

// src/knn_brute.cu:523-563.
void cublas_2nn_f(const des_t_f* q_points, const des_t_f* r_points, int q_n, int r_n, void* sorted,
                  sift_handle_t handle) {
    if (!handle) {
        set_error(nullptr, SIFT_ERR_INVALID, "cublas_2nn_f: null handle");
        return;
    }
    if (q_n < 0 || r_n < 0) {
        set_error(handle, SIFT_ERR_INVALID, "negative size");
        return;
    }
    Knn2Out o;
    o.sorted = static_cast<sift_float4*>(sorted);
    int rc = knn2_impl(handle, q_points, r_points, SIFT_FLOAT_DEVICE, (uint32_t)q_n, (uint32_t)r_n,
                       SIFT_COMPUTE_AUTO, KIND_F32, o, true, "cublas_2nn_f");
    if (rc == SIFT_OK) set_error(handle, SIFT_OK, "");
}

// src/knn_brute.cu:567-616 (dist buffer unused).
void cublas_2nn_brute_f(const des_t_f* q_points, const des_t_f* r_points, int q_n, int r_n, void* sorted,
                        float* dist, sift_handle_t handle) {
    (void)dist;
    cublas_2nn_f(q_points, r_points, q_n, r_n, sorted, handle);
}

// ------------------------------------------------- brute-force index ----
}  // extern "C"

extern "C" {

int sift_bf_index_build(sift_handle_t h, const void* r, int type, uint32_t r_n, int compute,
                        sift_bf_index_t* index) {
    if (!h || !index) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (!valid_type(type)) return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    if (r_n >= (1u << 24)) return set_error(h, SIFT_ERR_UNSUPPORTED, "r_n must be < 2^24");
    if (r_n > 0 && !r) return set_error(h, SIFT_ERR_INVALID, "null base");
    *index = nullptr;
    if (type == SIFT_FLOAT_HOST) {
        SIFT_TRY(stage_host_input(h, r, r_n, h->stage_r, &r));
        type = SIFT_FLOAT_DEVICE;
    }
    int kind = 0;
    SIFT_TRY(resolve_kind(h, r, r, type, 0, r_n, compute, KIND_F32, &kind));
    auto* ix = new sift_bf_index();
    ix->h = h;
    ix->r_n = r_n;
    ix->kind = kind;
    int* flag = static_cast<int*>(h->flag.p);
    int rc = hipMemsetAsync(flag, 0, 4, h->stream) == hipSuccess ? SIFT_OK : set_error(h, SIFT_ERR_HIP, "memset");
    const uint32_t rows_b = std::max<uint32_t>(32, ((r_n + 31) / 32) * 32);
    if (rc == SIFT_OK) rc = pack_side(h, r, type, r_n, rows_b, kind, PACK_BASE, ix->base, flag);
    if (rc == SIFT_OK && hipStreamSynchronize(h->stream) != hipSuccess)
        rc = set_error(h, SIFT_ERR_HIP, "index build sync");
    if (rc == SIFT_OK) {
        int bad = 0;
        if (hipMemcpy(&bad, flag, 4, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = set_error(h, SIFT_ERR_HIP, "index build flag");
        } else if (bad && kind != KIND_F32) {
            // the packed rows are all the index keeps: only FP32 holds other data exactly
            rc = set_error(h, SIFT_ERR_NOT_EXACT,
                           "BF16, F16, I8 and FP8 indexes need integer data in [0,255] (use F32 for other data)");
        }
        ix->coarse = bad != 0;  // F32 over other data: coarse keys, exact re-rank from the tiles
    }
    if (rc != SIFT_OK) {
        release_set(ix->base);
        delete ix;
        return rc;
    }
    *index = ix;
    return SIFT_OK;
}

}  // extern "C"

namespace sift {
// Query an index: pack the queries (flag: device int, set when some query
// value is not an integer in [0,255]), then the fused pass; float kinds
// re-rank exactly from the packed base tiles (integer BF16/F16 bases, any
// F32 base) and the query rows as given when the keys are coarse.
int bf_index_run(sift_context* h, sift_bf_index* ix, const void* q, int type, uint32_t q_n, const Knn2Out& out_in,
                 int* flag, PackedSet& qset, const Knn2Work* work) {
    SIFT_TRY(set_flag_word(h, flag, ix->coarse ? 1 : 0));
    SIFT_TRY(work ? pack_launch(h, q, type, q_n, nullptr, ix->kind, PACK_QUERY, qset, flag)
                  : pack_side(h, q, type, q_n, query_rows_pad(ix->kind, q_n), ix->kind, PACK_QUERY, qset, flag));
    Knn2Out o = out_in;
    if (ix->kind == KIND_BF16 || ix->kind == KIND_F16 || ix->kind == KIND_F32) {
        o.coarse = flag;
        o.src = ExactSrc{q, type, ix->base.tiles.p,
                         ix->kind == KIND_F32 ? SRC_F32_TILES : (ix->kind == KIND_BF16 ? SRC_BF16_TILES : SRC_F16_TILES)};
    }
    return knn2_packed(h, qset, ix->base, q_n, ix->r_n, o, work);
}
}  // namespace sift

extern "C" {

int sift_bf_index_query(sift_handle_t h, sift_bf_index_t ix, const void* q, int type, uint32_t q_n,
                        uint32_t* idx, float* dist, sift_float4* sorted, uint32_t* matches,
                        float ratio, int blocking) {
    if (!h || !ix) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (!valid_type(type)) return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    if (q_n == 0) return SIFT_OK;
    if (!q) return set_error(h, SIFT_ERR_INVALID, "null queries");
    if (type == SIFT_FLOAT_HOST) {
        SIFT_TRY(stage_host_input(h, q, q_n, h->stage_q, &q));
        type = SIFT_FLOAT_DEVICE;
    }
    int* flag = static_cast<int*>(h->flag.p);
    Knn2Out o{idx, dist, sorted, matches, ratio};
    SIFT_TRY(bf_index_run(h, ix, q, type, q_n, o, flag, h->q_set, nullptr));
    if (blocking) {
        int fl[3] = {0, 0, 0};
        SIFT_HIP_TRY(h, hipMemcpyAsync(fl, flag, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
        SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
        SIFT_TRY(wait_error(h, fl[2], "sift_bf_index_query"));
        if (fl[0] && (ix->kind == KIND_I8 || ix->kind == KIND_FP8))
            return set_error(h, SIFT_ERR_NOT_EXACT, "I8 and FP8 indexes need integer queries in [0,255]");
    }
    return SIFT_OK;
}

int sift_bf_index_destroy(sift_bf_index_t ix) {
    if (!ix) return SIFT_OK;
    (void)hipStreamSynchronize(ix->h->stream);
    release_set(ix->base);
    delete ix;
    return SIFT_OK;
}

}  // extern "C"
