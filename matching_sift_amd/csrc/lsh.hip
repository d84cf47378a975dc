// lsh.hip — random-hyperplane LSH: build (projection + sign packing + radix
// sort into buckets) and query (bucket lookup + coalesced gather + exact int32
// re-rank), and the reference-named lsh_gpu.
//
// Reference: lsh_gpu (src/lsh_h.cu:722-1015) hashes Q and R with a cuBLAS
// HGEMM per table (:909, :914), packs sign bits (set_bit_h :265-317), sorts
// with thrust and run-length encodes buckets on the host (:565-629),
// merge-joins q/r buckets on the host (:657-681) and brute-forces each bucket
// pair in FP16 with a 64-bit CAS top-2 (:441-562, :103-134).
//
// Here:
//   hyperplanes : L tables x b planes, integers in [-127,127] (Gaussian sd 32
//                 or the reference's +-1), exact in bf16 and int8
//   project     : one MFMA 32x32x16 tile per (32 points, table): hyperplane
//                 h lives in accumulator lane h, so __ballot(acc[j] < 0) is
//                 directly the packed code of two points (bit h = sign of
//                 h.x < 0, the reference's bit order :302-309)
//   buckets     : keys (table << b | code) of all tables sorted together with
//                 a stable LSD radix sort; a query's bucket in table t is the
//                 equal range of (t << b | code_q), found by binary search
//   re-rank     : one wave per query, 8 candidates per step (8 lanes x 16 B
//                 per candidate row, fully coalesced), exact int32 squared L2
//                 via udot4, top-2 in registers with lexicographic (d, index)
//                 order and duplicate suppression across tables — no atomics.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fused.h"
#include "rows.h"

namespace sift {
int exclusive_scan(sift_context* h, uint32_t* data, uint32_t n, DevBuf* levels, int depth);
int scan_reserve(sift_context* h, uint32_t n, DevBuf* levels);
int sort_reserve(sift_context* h, uint32_t seg_n, uint32_t nseg, DevBuf& hist, DevBuf* levels);
int radix_sort_pairs(sift_context* h, uint32_t** keys, uint32_t** vals, uint32_t** tk, uint32_t** tv,
                     uint32_t seg_n, uint32_t nseg, int bits, DevBuf& hist, DevBuf* levels, bool implicit_vals);
}  // namespace sift

using namespace sift;

struct sift_lsh_index {
    sift_context* h = nullptr;
    uint32_t n = 0;
    int nbits = 0, L = 0, hash_kind = 0;
    uint64_t seed = 0;
    std::vector<float> planes;  // [L][nbits][128]
    PackedSet hp;               // hyperplanes packed as B operands, 32 per table
    DevBuf hp_i8;               // the same as int8 B operands ([L][4][64] x 16 B), then thr = -128 sum(h) [L][32]
    std::vector<sift_mlp*> mlps;  // learned hash: one network per table (owned by the caller,
                                  // or by the index when it was loaded from a file)
    bool owns_mlps = false;
    int probes = 0;  // multi-probe: extra buckets per table (the least confident bits flipped)
    double cand_per_lookup = 0.0;  // mean bucket size seen by a base-like point (re-rank order rule)
    DevBuf keys, vals;          // sorted (table << b | code), base row index; L*n each
    DevBuf offs;                // b <= OFFS_MAX_BITS: bucket starts, [L][2^b + 1] absolute positions
    DevBuf base_u8, base_nrm;   // gather copy (integer data)
    // Data that is not all integers in [0,255]: hashed as clamp(rint(x * 2^qexp),
    // -255, 255) (integers, so codes stay exact) and re-ranked from a float32
    // copy with the reference's sequential FP32 distance.
    bool float_mode = false;
    int qexp = 0;
    DevBuf base_f32;
    DevBuf scratch_k, scratch_v, hist, levels[4];
};

namespace {

// Per-handle query scratch.
struct LshScratch {
    DevBuf q_u8, q_nrm, qcodes, flag;
    DevBuf qbase, qmarg;  // multi-probe: base codes [L][q_n] and margins [L][q_n][32]
    PackedSet q_tiles;
    // bucket-ordered re-rank: unit keys / ids (+ sort scratch), partial top-2 and counts
    DevBuf ukeys, uvals, utk, utv, uhist, ulevels[4], upart, ucnt;
    // exact fallback
    DevBuf fb_nc, fb_dist, fb_list, fb_count, fb_map, fb_out;
    PackedSet fb_q;
};

LshScratch* scratch_of(sift_context* h) {
    if (!h->lsh_state) h->lsh_state = new LshScratch();
    return static_cast<LshScratch*>(h->lsh_state);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ projection --
// One wave per 32-point tile; every table's 32 hyperplanes in turn.
// codes[t * stride + row] = packed sign bits of table t.
__global__ __launch_bounds__(256) void lsh_project(const uint4* __restrict__ x_tiles,
                                                   uint32_t n_rows, uint32_t n_tiles,
                                                   const uint4* __restrict__ h_tiles, int L,
                                                   uint32_t* __restrict__ codes, uint32_t stride,
                                                   int b, int combine, float* __restrict__ margins) {
    const int lane = threadIdx.x & 63;
    const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= n_tiles) return;  // whole wave
    uint4 a[8];
#pragma unroll
    for (int g = 0; g < 8; g++) a[g] = x_tiles[((size_t)tile * 9 + g) * 64 + lane];  // 9 groups per tile
    const uint32_t row = tile * 32 + (uint32_t)lane;
    for (int t = 0; t < L; t++) {
        f32x16 acc = f32x16{};
#pragma unroll
        for (int g = 0; g < 8; g++) {
            const uint4 hb = h_tiles[((size_t)t * 9 + g) * 64 + lane];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a[g]),
                                                          __builtin_bit_cast(bf16x8_t, hb), acc, 0, 0, 0);
        }
        uint32_t code = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint64_t m = __ballot(acc[j] < 0.f);
            const int r0 = 8 * (j >> 2) + (j & 3);
            if (lane == r0) code = (uint32_t)m;
            if (lane == r0 + 4) code = (uint32_t)(m >> 32);
        }
        if (lane < 32 && row < n_rows)
            codes[(size_t)t * stride + row] = (combine && b < 32) ? (((uint32_t)t << b) | code) : code;
        if (margins) {  // |h.x| of hyperplane lane % 32 for the lane's 16 rows: [t][row][32]
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const uint32_t r = tile * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                if (r < n_rows) margins[((size_t)t * stride + r) * 32 + (lane & 31)] = fabsf(acc[j]);
            }
        }
    }
}

// Build pass over integer data, reading each input row once.  A block of
// PH_WAVES waves, one 32-row tile per wave.  Phase 1 reads the tile coalesced
// (16 lanes per row, 8 dims per lane, four rows per step), checks that every
// value is an integer in [0,255] (else *bad is set and the build redoes the
// pass in float mode), writes the uint8 gather copy and the int32 squared
// norms, and stages the rows as int8 x - 128 in LDS.  Phase 2 reads them back
// in the int8 MFMA A layout (lane l: row l % 32, dims 32g + 16(l / 32) + 0..15
// for g = 0..3).  Phase 3 reuses the LDS for the int8 hyperplanes of up to 8
// tables at a time, shared by the block's waves: plane p of table t gives
// bit p of the code where h.x < 0, i.e. h.(x - 128) < thr = -128 sum(h) (all
// integers, exact).  It writes for every table t the key (t << b | code) at
// t * n + row (the unsorted keys of the radix sort, whose values are the
// positions within each table).
// Code words of a 32-row x 32-plane accumulator tile (lane l: plane l % 32;
// element j: rows 8(j/4) + (j%4) in lanes 0-31, that + 4 in lanes 32-63):
// bit p of row r's word is acc(r, p) < th(p); lane r (< 32) receives row r's
// word.  One ballot per element; the two halves of each go into their rows'
// lanes by v_writelane.  The s_nop covers the wait states between the
// compares that wrote the SGPRs and the v_writelane that reads them.
__device__ inline uint32_t code_words(const i32x16& acc, int th) {
    uint32_t w[32];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const uint64_t mb = __ballot(acc[j] < th);
        w[2 * j] = (uint32_t)mb;
        w[2 * j + 1] = (uint32_t)(mb >> 32);
    }
    uint32_t code = 0;
    asm volatile(
        "s_nop 4\n"
        "v_writelane_b32 %0, %1, 0\n"
        "v_writelane_b32 %0, %2, 4\n"
        "v_writelane_b32 %0, %3, 1\n"
        "v_writelane_b32 %0, %4, 5\n"
        "v_writelane_b32 %0, %5, 2\n"
        "v_writelane_b32 %0, %6, 6\n"
        "v_writelane_b32 %0, %7, 3\n"
        "v_writelane_b32 %0, %8, 7\n"
        "v_writelane_b32 %0, %9, 8\n"
        "v_writelane_b32 %0, %10, 12\n"
        "v_writelane_b32 %0, %11, 9\n"
        "v_writelane_b32 %0, %12, 13\n"
        "v_writelane_b32 %0, %13, 10\n"
        "v_writelane_b32 %0, %14, 14\n"
        "v_writelane_b32 %0, %15, 11\n"
        "v_writelane_b32 %0, %16, 15\n"
        "v_writelane_b32 %0, %17, 16\n"
        "v_writelane_b32 %0, %18, 20\n"
        "v_writelane_b32 %0, %19, 17\n"
        "v_writelane_b32 %0, %20, 21\n"
        "v_writelane_b32 %0, %21, 18\n"
        "v_writelane_b32 %0, %22, 22\n"
        "v_writelane_b32 %0, %23, 19\n"
        "v_writelane_b32 %0, %24, 23\n"
        "v_writelane_b32 %0, %25, 24\n"
        "v_writelane_b32 %0, %26, 28\n"
        "v_writelane_b32 %0, %27, 25\n"
        "v_writelane_b32 %0, %28, 29\n"
        "v_writelane_b32 %0, %29, 26\n"
        "v_writelane_b32 %0, %30, 30\n"
        "v_writelane_b32 %0, %31, 27\n"
        "v_writelane_b32 %0, %32, 31\n"
        : "+v"(code)
        : "s"(w[0]), "s"(w[1]), "s"(w[2]), "s"(w[3]), "s"(w[4]), "s"(w[5]), "s"(w[6]), "s"(w[7]), "s"(w[8]), "s"(w[9]), "s"(w[10]), "s"(w[11]), "s"(w[12]), "s"(w[13]), "s"(w[14]), "s"(w[15]), "s"(w[16]), "s"(w[17]), "s"(w[18]), "s"(w[19]), "s"(w[20]), "s"(w[21]), "s"(w[22]), "s"(w[23]), "s"(w[24]), "s"(w[25]), "s"(w[26]), "s"(w[27]), "s"(w[28]), "s"(w[29]), "s"(w[30]), "s"(w[31]));
    return code;
}

constexpr int PH_WAVES = 8;
constexpr int PH_ROW = 9;         // uint4 per staged row: 8 of data + 1 of padding (conflict-free reads)
constexpr int PH_TABLE = 4 * 64;  // uint4 of one table's int8 plane groups

template <int TYPE>
__global__ __launch_bounds__(64 * PH_WAVES) __attribute__((amdgpu_waves_per_eu(6))) void lsh_pack_hash(const void* __restrict__ in, uint32_t n,
                                                               uint32_t n_tiles, const uint4* __restrict__ planes,
                                                               const int* __restrict__ thr, int L, int b,
                                                               uint8_t* __restrict__ u8_out,
                                                               int* __restrict__ nrm_out, int* __restrict__ bad,
                                                               uint32_t* __restrict__ keys) {
    static_assert(PH_WAVES * 32 * PH_ROW >= 8 * PH_TABLE, "the planes of 8 tables fit the staging area");
    __shared__ uint4 lds[PH_WAVES * 32 * PH_ROW];
    __shared__ int lthr[8 * 32];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t tile = blockIdx.x * PH_WAVES + (uint32_t)wave;
    const bool active = tile < n_tiles;  // idle waves still take part in the block's barriers
    const uint32_t base = tile * 32;
    uint4 a[4];
    if (active) {
        uint4* st = lds + wave * 32 * PH_ROW;
        // phase 1: step k covers rows base + 4k .. base + 4k + 3; lane = 16 (row % 4) + chunk
        const int c = lane & 15, rq = lane >> 4;
        bool ok = true;
#pragma unroll
        for (int k0 = 0; k0 < 8; k0 += 4) {  // two batches of four steps, each batch's loads issued together
            float v[4][8];
#pragma unroll
            for (int k = 0; k < 4; k++) {  // padding rows read row n - 1
                const uint32_t row = base + 4 * (k0 + k) + rq;
                load8_src(in, TYPE, row < n ? row : n - 1, c, v[k]);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                // rows past n repeat row n - 1 and write its (identical) bytes and norm again
                const int r = 4 * (k0 + k) + rq;
                const uint32_t grow = base + r < n ? base + r : n - 1;
                uint32_t q[2] = {0u, 0u};
#pragma unroll
                for (int i = 0; i < 8; i++) {  // round and clamp to a byte, then check the byte is the value
                    const float x = v[k][i];
                    q[i >> 2] = __builtin_amdgcn_cvt_pk_u8_f32(x, i & 3, q[i >> 2]);
                    ok &= x == (float)((q[i >> 2] >> (8 * (i & 3))) & 255u);
                }
                int sq = (int)__builtin_amdgcn_udot4(q[0], q[0], __builtin_amdgcn_udot4(q[1], q[1], 0u, false), false);
                reinterpret_cast<uint2*>(st)[r * 2 * PH_ROW + c] = make_uint2(q[0] ^ 0x80808080u, q[1] ^ 0x80808080u);
                // sum over the row's 16 lanes (one DPP row): rotations by 8, 4, 2, 1
                sq += __builtin_amdgcn_update_dpp(0, sq, 0x128, 0xF, 0xF, false);
                sq += __builtin_amdgcn_update_dpp(0, sq, 0x124, 0xF, 0xF, false);
                sq += __builtin_amdgcn_update_dpp(0, sq, 0x122, 0xF, 0xF, false);
                sq += __builtin_amdgcn_update_dpp(0, sq, 0x121, 0xF, 0xF, false);
                reinterpret_cast<uint2*>(u8_out)[(size_t)grow * 16 + c] = make_uint2(q[0], q[1]);
                if (c == 0) nrm_out[grow] = sq;
            }
            // the batch's checks done here (its values dead), its loads before the next batch's
            int okv = ok;
            asm volatile("" : "+v"(okv)::"memory");
            ok = okv != 0;
        }
        const unsigned long long m = __ballot(!ok);
        if (m && lane == __ffsll((long long)m) - 1) atomicOr(bad, 1);
        // phase 2: the wave reads back what all its lanes wrote
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int r = lane & 31, h = lane >> 5;
#pragma unroll
        for (int g = 0; g < 4; g++) a[g] = st[r * PH_ROW + 2 * g + h];
    }
    const uint32_t row = base + (uint32_t)(lane & 31);
    for (int t0 = 0; t0 < L; t0 += 8) {
        const int nt = min(8, L - t0);
        __syncthreads();  // the staging area (or the previous planes) is no longer read
        for (int e = threadIdx.x; e < nt * PH_TABLE; e += 64 * PH_WAVES) lds[e] = planes[(size_t)t0 * PH_TABLE + e];
        if (threadIdx.x < nt * 32) lthr[threadIdx.x] = thr[t0 * 32 + threadIdx.x];
        __syncthreads();
        if (!active) continue;
        for (int k = 0; k < nt; k++) {
            const int t = t0 + k;
            i32x16 acc = i32x16{};
#pragma unroll
            for (int g = 0; g < 4; g++) acc = KT<KIND_I8>::mma(a[g], lds[k * PH_TABLE + g * 64 + lane], acc);
            const int th = lthr[k * 32 + (lane & 31)];  // accumulator column = plane lane % 32
            const uint32_t code = code_words(acc, th);
            if (lane < 32 && row < n) keys[(size_t)t * n + row] = b < 32 ? (((uint32_t)t << b) | code) : code;
        }
    }
}

// Queries the LSH answer does not settle: fewer than min_cand candidates, or
// (accept_ratio > 0) a best candidate that does not pass the ratio test
// sqrt(d1)/sqrt(d2) < accept_ratio against the next one (IEEE sqrt and
// division, as the match output).
__device__ inline bool needs_exact(const uint32_t* nc, const float* d, uint32_t q, uint32_t min_cand,
                                   float accept_ratio) {
    bool low = nc[q] < min_cand;
    if (accept_ratio > 0.f) low |= !(sqrtf(d[2 * (size_t)q]) / sqrtf(d[2 * (size_t)q + 1]) < accept_ratio);
    return low;
}

__global__ void fallback_flags(const uint32_t* __restrict__ nc, const float* __restrict__ d, uint32_t q_n,
                               uint32_t min_cand, float accept_ratio, uint32_t* __restrict__ flag) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < q_n) flag[q] = needs_exact(nc, d, q, min_cand, accept_ratio) ? 1u : 0u;
}

// pos = exclusive scan of the flags: the fallback queries in increasing order
// become map[0 .. count); the last thread writes count.
__global__ void fallback_compact(const uint32_t* __restrict__ nc, const float* __restrict__ d, uint32_t q_n,
                                 uint32_t min_cand, float accept_ratio, const uint32_t* __restrict__ pos,
                                 int32_t* __restrict__ map, uint32_t* __restrict__ count) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= q_n) return;
    const bool low = needs_exact(nc, d, q, min_cand, accept_ratio);
    if (low) map[pos[q]] = (int32_t)q;
    if (q == q_n - 1) *count = pos[q] + (low ? 1u : 0u);
}

// Pad rows of the fallback query pack: map[cnt .. rows_pad) = -1.
__global__ void map_tail(int32_t* __restrict__ map, uint32_t cnt, uint32_t rows_pad) {
    const uint32_t i = cnt + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows_pad) map[i] = -1;
}

// Exact results of the fallback queries into the caller's outputs.
__global__ void scatter_exact(const int32_t* __restrict__ map, uint32_t cnt, const uint32_t* __restrict__ oi,
                              const float* __restrict__ od, const uint32_t* __restrict__ om,
                              uint32_t* __restrict__ idx, float* __restrict__ dist, uint32_t* __restrict__ matches) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const size_t q = (size_t)map[i];
    if (idx) {
        idx[2 * q] = oi[2 * (size_t)i];
        idx[2 * q + 1] = oi[2 * (size_t)i + 1];
    }
    if (dist) {
        dist[2 * q] = od[2 * (size_t)i];
        dist[2 * q + 1] = od[2 * (size_t)i + 1];
    }
    if (matches) matches[q] = om[i];
}

// ----------------------------------------------------------------- query --
struct RerankArgs {
    const uint8_t* base_u8;
    const int* base_nrm;
    const uint8_t* q_u8;
    const uint32_t* keys;
    const uint32_t* vals;
    const uint32_t* qcodes;  // [L][P + 1][q_n]: per table the base code, then the probes
    const uint32_t* offs;    // bucket starts [L][2^b + 1], or null (binary search)
    uint32_t n, q_n;
    int L, b, P;
    uint32_t* idx;
    float* dist;
    uint32_t* ncand;
    uint32_t* matches;
    float ratio;
};

// First position in [lo, hi) whose key is >= k (upper: > k).
__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t k) {
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < k) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ inline uint32_t upper_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t k) {
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] <= k) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// top-2 of two lexicographically sorted lists that may share elements.
__device__ inline void merge2_dedup(uint32_t& a1, uint32_t& ia1, uint32_t& a2, uint32_t& ia2,
                                    uint32_t b1, uint32_t ib1, uint32_t b2, uint32_t ib2) {
    if (ia1 == ib1 && ia1 != 0xFFFFFFFFu) {  // same element heads both lists
        if (lex_less(b2, ib2, a2, ia2)) { a2 = b2; ia2 = ib2; }
        return;
    }
    if (lex_less(a1, ia1, b1, ib1)) {
        if (lex_less(b1, ib1, a2, ia2) && ib1 != ia2) { a2 = b1; ia2 = ib1; }
    } else {
        const uint32_t oa1 = a1, oia1 = ia1;
        a1 = b1; ia1 = ib1;
        if (lex_less(b2, ib2, oa1, oia1)) { a2 = b2; ia2 = ib2; } else { a2 = oa1; ia2 = oia1; }
    }
}

// Bucket range [lo, hi) of (table, probe) tp for query q: the offset table,
// or a binary search of the sorted keys.
__device__ inline void bucket_range(const RerankArgs& a, int tp, uint32_t q, uint32_t& lo, uint32_t& hi) {
    const int t = tp / (a.P + 1);
    const uint32_t code = a.qcodes[(size_t)tp * a.q_n + q];
    if (a.offs) {
        const uint32_t* o = a.offs + (size_t)t * ((1u << a.b) + 1u) + code;
        lo = o[0];
        hi = o[1];
    } else {
        const uint32_t tkey = a.b < 32 ? ((uint32_t)t << a.b) : 0u;  // b == 32 implies L == 1
        const uint32_t key = tkey | code;
        const uint32_t t0 = (uint32_t)t * a.n, t1 = t0 + a.n;
        lo = lower_bound_u32(a.keys, t0, t1, key);
        hi = upper_bound_u32(a.keys, lo, t1, key);
    }
}

// One wave scans the candidates at sorted positions [lo, hi): lane (g, s)
// takes rows lo + g, lo + g + 8, ... and their 16-byte segment s, UNR rows per
// slot at a time (8 UNR rows in flight per wave), the next chunk's row ids
// loaded before the current rows are used.  Each slot keeps its own top-2
// (lexicographic on (distance, row), rows already held are skipped).
__device__ inline void scan_bucket(const RerankArgs& a, const uint4& qv, uint32_t nq, uint32_t lo, uint32_t hi,
                                   int g, int s, uint32_t& m1, uint32_t& i1, uint32_t& m2, uint32_t& i2) {
    constexpr int UNR = 8;
    uint32_t p0 = lo + g;
    uint32_t jv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) jv[u] = p0 + 8u * u < hi ? a.vals[p0 + 8u * u] : 0xFFFFFFFFu;
    while (p0 < hi) {  // jv[0] is a real row here
        uint4 bv[UNR];
        int bn[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint32_t j = jv[u] != 0xFFFFFFFFu ? jv[u] : jv[0];
            bv[u] = reinterpret_cast<const uint4*>(a.base_u8 + (size_t)j * 128)[s];
            bn[u] = a.base_nrm[j];
        }
        const uint32_t pn = p0 + 8u * UNR;
        uint32_t jn[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) jn[u] = pn + 8u * u < hi ? a.vals[pn + 8u * u] : 0xFFFFFFFFu;
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            uint32_t dot = __builtin_amdgcn_udot4(qv.x, bv[u].x, 0u, false);
            dot = __builtin_amdgcn_udot4(qv.y, bv[u].y, dot, false);
            dot = __builtin_amdgcn_udot4(qv.z, bv[u].z, dot, false);
            dot = __builtin_amdgcn_udot4(qv.w, bv[u].w, dot, false);
            dot += __shfl_xor(dot, 1);
            dot += __shfl_xor(dot, 2);
            dot += __shfl_xor(dot, 4);
            const uint32_t d = nq + (uint32_t)bn[u] - 2u * dot;
            const uint32_t j = jv[u];
            if (j != 0xFFFFFFFFu && j != i1 && j != i2) {
                const bool c1 = lex_less(d, j, m1, i1), c2 = lex_less(d, j, m2, i2);
                m2 = c1 ? m1 : (c2 ? d : m2);
                i2 = c1 ? i1 : (c2 ? j : i2);
                m1 = c1 ? d : m1;
                i1 = c1 ? j : i1;
            }
        }
        p0 = pn;
#pragma unroll
        for (int u = 0; u < UNR; u++) jv[u] = jn[u];
    }
}

// Query row segment s and the query's squared norm (all 64 lanes).
__device__ inline uint32_t load_query(const RerankArgs& a, uint32_t q, int s, uint4& qv) {
    qv = reinterpret_cast<const uint4*>(a.q_u8 + (size_t)q * 128)[s];
    uint32_t nq = __builtin_amdgcn_udot4(qv.x, qv.x, 0u, false);
    nq = __builtin_amdgcn_udot4(qv.y, qv.y, nq, false);
    nq = __builtin_amdgcn_udot4(qv.z, qv.z, nq, false);
    nq = __builtin_amdgcn_udot4(qv.w, qv.w, nq, false);
    nq += __shfl_xor(nq, 1);
    nq += __shfl_xor(nq, 2);
    nq += __shfl_xor(nq, 4);
    return nq;
}

// The 8 slots of a wave hold disjoint or overlapping lists: merge them.
__device__ inline void merge_slots(uint32_t& m1, uint32_t& i1, uint32_t& m2, uint32_t& i2) {
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) {
        const uint32_t b1 = __shfl_xor(m1, off), ib1 = __shfl_xor(i1, off);
        const uint32_t b2 = __shfl_xor(m2, off), ib2 = __shfl_xor(i2, off);
        merge2_dedup(m1, i1, m2, i2, b1, ib1, b2, ib2);
    }
}

__device__ inline void write_result(const RerankArgs& a, uint32_t q, uint32_t m1, uint32_t i1, uint32_t m2,
                                    uint32_t i2, uint32_t nc) {
    const bool h1 = i1 != 0xFFFFFFFFu, h2 = i2 != 0xFFFFFFFFu;
    const float d1 = h1 ? (float)m1 : 3.402823466e+38f, d2 = h2 ? (float)m2 : 3.402823466e+38f;
    if (a.idx) {
        a.idx[2 * (size_t)q] = i1;
        a.idx[2 * (size_t)q + 1] = i2;
    }
    if (a.dist) {
        a.dist[2 * (size_t)q] = d1;
        a.dist[2 * (size_t)q + 1] = d2;
    }
    if (a.ncand) a.ncand[q] = nc;
    if (a.matches) a.matches[q] = (h1 && sqrtf(d1) / sqrtf(d2) < a.ratio) ? i1 : 0xFFFFFFFFu;
}

// One wave per query: every (table, probe) bucket in turn.  Bucket ranges are
// looked up for up to 64 (table, probe) pairs at once, one per lane.
__global__ __launch_bounds__(256) void lsh_rerank(RerankArgs a) {
    const int lane = threadIdx.x & 63;
    const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= a.q_n) return;  // whole wave
    const int g = lane >> 3, s = lane & 7;  // candidate slot, 16-byte segment
    uint4 qv;
    const uint32_t nq = load_query(a, q, s, qv);
    uint32_t m1 = 0xFFFFFFFFu, m2 = 0xFFFFFFFFu, i1 = 0xFFFFFFFFu, i2 = 0xFFFFFFFFu;
    uint32_t nc = 0;
    const int TP = a.L * (a.P + 1);
    for (int tb = 0; tb < TP; tb += 64) {
        uint32_t rlo = 0, rhi = 0;
        if (tb + lane < TP) bucket_range(a, tb + lane, q, rlo, rhi);
        const int nt = min(64, TP - tb);
        for (int i = 0; i < nt; i++) {
            const uint32_t lo = __shfl(rlo, i), hi = __shfl(rhi, i);
            nc += hi - lo;
            scan_bucket(a, qv, nq, lo, hi, g, s, m1, i1, m2, i2);
        }
    }
    merge_slots(m1, i1, m2, i2);
    if (lane == 0) write_result(a, q, m1, i1, m2, i2, nc);
}

// Float data: one wave per query, each lane one candidate row at a time (the
// sequential FP32 distance of host_lenght over the float32 copy of the base
// and the query row as given), top-2 per lane with rows already held skipped,
// then a wave-wide merge that drops rows held twice.  Distances are >= +0, so
// their bit patterns order as unsigned integers and the (distance, row) order
// is the integer path's.
struct RerankF32Args {
    RerankArgs a;
    const void* q;  // query rows of type q_type
    int q_type;
    const float* base;
};
__global__ __launch_bounds__(256) void lsh_rerank_f32(RerankF32Args f) {
    const RerankArgs& a = f.a;
    const int lane = threadIdx.x & 63;
    const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= a.q_n) return;  // whole wave
    uint32_t m1 = 0xFFFFFFFFu, m2 = 0xFFFFFFFFu, i1 = 0xFFFFFFFFu, i2 = 0xFFFFFFFFu;
    uint32_t nc = 0;
    const int TP = a.L * (a.P + 1);
    for (int tb = 0; tb < TP; tb += 64) {
        uint32_t rlo = 0, rhi = 0;
        if (tb + lane < TP) bucket_range(a, tb + lane, q, rlo, rhi);
        const int nt = min(64, TP - tb);
        for (int i = 0; i < nt; i++) {
            const uint32_t lo = __shfl(rlo, i), hi = __shfl(rhi, i);
            nc += hi - lo;
            for (uint32_t p = lo + lane; p < hi; p += 64) {
                const uint32_t j = a.vals[p];
                if (j == i1 || j == i2) continue;
                const uint32_t d = __float_as_uint(seq_dist(f.q, f.q_type, q, f.base, SIFT_FLOAT_DEVICE, j));
                const bool c1 = lex_less(d, j, m1, i1), c2 = lex_less(d, j, m2, i2);
                m2 = c1 ? m1 : (c2 ? d : m2);
                i2 = c1 ? i1 : (c2 ? j : i2);
                m1 = c1 ? d : m1;
                i1 = c1 ? j : i1;
            }
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t b1 = __shfl_xor(m1, off), ib1 = __shfl_xor(i1, off);
        const uint32_t b2 = __shfl_xor(m2, off), ib2 = __shfl_xor(i2, off);
        merge2_dedup(m1, i1, m2, i2, b1, ib1, b2, ib2);
    }
    if (lane != 0) return;
    const bool h1 = i1 != 0xFFFFFFFFu, h2 = i2 != 0xFFFFFFFFu;
    const float d1 = h1 ? __uint_as_float(m1) : 3.402823466e+38f, d2 = h2 ? __uint_as_float(m2) : 3.402823466e+38f;
    if (a.idx) {
        a.idx[2 * (size_t)q] = i1;
        a.idx[2 * (size_t)q + 1] = i2;
    }
    if (a.dist) {
        a.dist[2 * (size_t)q] = d1;
        a.dist[2 * (size_t)q + 1] = d2;
    }
    if (a.ncand) a.ncand[q] = nc;
    if (a.matches) a.matches[q] = (h1 && sqrtf(d1) / sqrtf(d2) < a.ratio) ? i1 : 0xFFFFFFFFu;
}

// Bucket-ordered variant: a unit is one (table, probe, query) lookup, u =
// tp * q_n + q; units are visited in order of their bucket (sorted keys
// (table << b) | code), so the waves that read one bucket's rows run side by
// side and share them in cache.  Each unit leaves its top-2 and candidate
// count; lsh_merge_units combines a query's units.
__global__ void unit_keys(const uint32_t* __restrict__ qcodes, uint32_t nu, uint32_t q_n, int P1, int b,
                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint32_t t = (u / q_n) / (uint32_t)P1;
    keys[u] = (b < 32 ? (t << b) : 0u) | qcodes[u];
    vals[u] = u;
}

__global__ __launch_bounds__(256) void lsh_rerank_units(RerankArgs a, const uint32_t* __restrict__ order,
                                                        uint32_t nu, uint4* __restrict__ part,
                                                        uint32_t* __restrict__ pcnt) {
    const int lane = threadIdx.x & 63;
    const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nu) return;  // whole wave
    const uint32_t u = order[i];
    const uint32_t tp = u / a.q_n, q = u - tp * a.q_n;
    const int g = lane >> 3, s = lane & 7;
    uint4 qv;
    const uint32_t nq = load_query(a, q, s, qv);
    uint32_t lo, hi;
    bucket_range(a, (int)tp, q, lo, hi);
    uint32_t m1 = 0xFFFFFFFFu, m2 = 0xFFFFFFFFu, i1 = 0xFFFFFFFFu, i2 = 0xFFFFFFFFu;
    scan_bucket(a, qv, nq, lo, hi, g, s, m1, i1, m2, i2);
    merge_slots(m1, i1, m2, i2);
    if (lane == 0) {
        part[u] = make_uint4(m1, i1, m2, i2);
        pcnt[u] = hi - lo;
    }
}

__global__ void lsh_merge_units(RerankArgs a, const uint4* __restrict__ part, const uint32_t* __restrict__ pcnt) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.q_n) return;
    uint32_t m1 = 0xFFFFFFFFu, m2 = 0xFFFFFFFFu, i1 = 0xFFFFFFFFu, i2 = 0xFFFFFFFFu;
    uint32_t nc = 0;
    const int TP = a.L * (a.P + 1);
    for (int tp = 0; tp < TP; tp++) {
        const uint4 p = part[(size_t)tp * a.q_n + q];
        merge2_dedup(m1, i1, m2, i2, p.x, p.y, p.z, p.w);
        nc += pcnt[(size_t)tp * a.q_n + q];
    }
    write_result(a, q, m1, i1, m2, i2, nc);
}

// ---------------------------------------------------------- hyperplanes --
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) {
            uint64_t z = (x += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            v = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(6.283185307179586 * u2);
    }
};

// Gaussian (sd 32, rounded, clipped to +-127) or the reference's +-1
// hyperplanes (make_vector_h_sum_1, src/lsh_h.cu:51-84).
void make_planes(std::vector<float>& p, int L, int nbits, int kind, uint64_t seed) {
    p.assign((size_t)L * nbits * 128, 0.f);
    Rng r(seed);
    for (auto& v : p) {
        if (kind == SIFT_LSH_HYPERPLANE_PM1) {
            v = r.uniform() >= 0.5 ? 1.f : -1.f;
        } else {
            double x = std::nearbyint(32.0 * r.normal());
            v = (float)std::max(-127.0, std::min(127.0, x));
        }
    }
}

int project(sift_context* h, const PackedSet& xs, uint32_t n, const PackedSet& hp, int L, int b,
            uint32_t* codes, uint32_t stride, bool combine, float* margins = nullptr) {
    if (n == 0) return SIFT_OK;
    if (xs.ntiles == 0) return SIFT_OK;
    const unsigned grid = (xs.ntiles + 3) / 4;
    lsh_project<<<grid, 256, 0, h->stream>>>(static_cast<const uint4*>(xs.tiles.p), n, xs.ntiles,
                                             static_cast<const uint4*>(hp.tiles.p), L, codes, stride,
                                             b, combine ? 1 : 0, margins);
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

// Codes of every table: hyperplanes (one launch) or one network per table.
// margins (optional, [L][stride][32]): |h.x| or |z| per bit, signed z for
// the learned kind (the probe kernel takes the magnitude).
int hash_codes(sift_context* h, const sift_lsh_index* ix, const PackedSet& xs, uint32_t n, uint32_t* codes,
               uint32_t stride, bool combine, float* margins = nullptr) {
    if (ix->hash_kind != SIFT_LSH_LEARNED)
        return project(h, xs, n, ix->hp, ix->L, ix->nbits, codes, stride, combine, margins);
    return mlp_forward_tables(ix->mlps.data(), ix->L, xs, n, margins, codes, stride, combine);
}

// Multi-probe codes: per (table, query) the base code, then the code with
// each of the P least confident bits flipped (smallest |margin| first, lower
// bit first on ties): pcodes[(t * (P + 1) + p) * q_n + q].
__global__ void make_probes(const uint32_t* __restrict__ codes, const float* __restrict__ margins, uint32_t q_n,
                            int L, int b, int P, uint32_t* __restrict__ pcodes) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)L * q_n) return;
    const uint32_t t = (uint32_t)(i / q_n), q = (uint32_t)(i % q_n);
    const uint32_t code = codes[i];
    const float* m = margins + i * 32;
    float mv[32];
#pragma unroll
    for (int k = 0; k < 32; k++) mv[k] = k < b ? fabsf(m[k]) : __builtin_huge_valf();
    uint32_t* out = pcodes + (size_t)t * (P + 1) * q_n + q;
    out[0] = code;
    uint32_t used = 0;
    for (int p = 1; p <= P; p++) {
        int best = -1;
        float bv = __builtin_huge_valf();
#pragma unroll
        for (int k = 0; k < 32; k++) {
            const bool free_bit = !((used >> k) & 1u) && k < b;
            if (free_bit && (best < 0 || mv[k] < bv)) {
                best = k;
                bv = mv[k];
            }
        }
        used |= 1u << best;
        out[(size_t)p * q_n] = code ^ (1u << best);
    }
}

constexpr int MAX_PROBES = 16;

// Query codes for the re-rank, [L][P + 1][q_n], from the packed queries in
// sc->q_tiles: the base codes alone without probes, else base codes and
// margins, then make_probes.
int query_codes(sift_context* h, const sift_lsh_index* ix, LshScratch* sc, uint32_t q_n, const uint32_t** out) {
    const int L = ix->L, P = ix->probes;
    SIFT_TRY(reserve(h, sc->qcodes, (size_t)q_n * L * (P + 1) * 4));
    auto* qcodes = static_cast<uint32_t*>(sc->qcodes.p);
    if (P == 0) {
        SIFT_TRY(hash_codes(h, ix, sc->q_tiles, q_n, qcodes, q_n, false));
    } else {
        SIFT_TRY(reserve(h, sc->qbase, (size_t)q_n * L * 4));
        SIFT_TRY(reserve(h, sc->qmarg, (size_t)q_n * L * 32 * 4));
        auto* base = static_cast<uint32_t*>(sc->qbase.p);
        SIFT_TRY(hash_codes(h, ix, sc->q_tiles, q_n, base, q_n, false, static_cast<float*>(sc->qmarg.p)));
        const size_t nt = (size_t)L * q_n;
        make_probes<<<(unsigned)((nt + 255) / 256), 256, 0, h->stream>>>(base, static_cast<const float*>(sc->qmarg.p),
                                                                         q_n, L, ix->nbits, P, qcodes);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    *out = qcodes;
    return SIFT_OK;
}

// Bucket-ordered re-rank (lsh_rerank_units) when the lookups are expected to
// bring many candidates each (at least 32 per (table, probe) lookup and 4M in
// all): the unit sort and merge then cost little against reading each
// bucket's rows once per group of waves instead of once per query.  Small
// buckets go one wave per query (lsh_rerank).  SIFT_LSH_ORDER=0/1 forces
// either.
bool bucket_ordered(const sift_lsh_index* ix, uint32_t q_n) {
    const double lookups = (double)q_n * ix->L * (ix->probes + 1);
    if (lookups >= 2147483648.0) return false;  // unit ids are 32-bit
    if (const char* e = std::getenv("SIFT_LSH_ORDER")) return e[0] == '1';
    return ix->cand_per_lookup >= 32.0 && ix->cand_per_lookup * lookups >= 4e6;
}

int rerank_units(sift_context* h, const sift_lsh_index* ix, LshScratch* sc, const RerankArgs& a) {
    const uint32_t TP = (uint32_t)(ix->L * (ix->probes + 1));
    const uint32_t nu = TP * a.q_n;
    for (DevBuf* b : {&sc->ukeys, &sc->uvals, &sc->utk, &sc->utv, &sc->ucnt}) SIFT_TRY(reserve(h, *b, (size_t)nu * 4));
    SIFT_TRY(reserve(h, sc->upart, (size_t)nu * 16));
    auto* k = static_cast<uint32_t*>(sc->ukeys.p);
    auto* v = static_cast<uint32_t*>(sc->uvals.p);
    auto* tk = static_cast<uint32_t*>(sc->utk.p);
    auto* tv = static_cast<uint32_t*>(sc->utv.p);
    unit_keys<<<(nu + 255) / 256, 256, 0, h->stream>>>(a.qcodes, nu, a.q_n, ix->probes + 1, ix->nbits, k, v);
    SIFT_HIP_TRY(h, hipGetLastError());
    int tbits = 0;
    while ((1 << tbits) < ix->L) tbits++;
    const int bits = std::min(32, ix->nbits + tbits);
    SIFT_TRY(radix_sort_pairs(h, &k, &v, &tk, &tv, nu, 1, bits, sc->uhist, sc->ulevels, false));
    auto* part = static_cast<uint4*>(sc->upart.p);
    auto* cnt = static_cast<uint32_t*>(sc->ucnt.p);
    lsh_rerank_units<<<(nu + 3) / 4, 256, 0, h->stream>>>(a, v, nu, part, cnt);
    SIFT_HIP_TRY(h, hipGetLastError());
    lsh_merge_units<<<(a.q_n + 255) / 256, 256, 0, h->stream>>>(a, part, cnt);
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

void free_index(sift_lsh_index* ix) {
    if (ix->owns_mlps)
        for (sift_mlp* m : ix->mlps) sift_mlp_destroy(m);
    for (DevBuf* b : {&ix->keys, &ix->vals, &ix->offs, &ix->base_u8, &ix->base_nrm, &ix->base_f32, &ix->scratch_k,
                      &ix->scratch_v, &ix->hist, &ix->hp.tiles, &ix->hp.norm, &ix->hp.norm_lane, &ix->hp_i8})
        release(*b);
    for (auto& l : ix->levels) release(l);
    delete ix;
}

int check_flag(sift_context* h, int* flag, const char* what) {
    int bad = 0;
    SIFT_HIP_TRY(h, hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (bad)
        return set_error(h, SIFT_ERR_NOT_EXACT, std::string(what) +
                                                    ": an index built from integer descriptors in [0,255] needs "
                                                    "queries of that kind (build from float data for float queries)");
    return SIFT_OK;
}

// Bucket starts for direct lookup: offs[t][c] = first position of table t's
// sorted range [t*n, (t+1)*n) whose key is >= (t << b | c); offs[t][2^b] is
// the table's end.  Small table sets: one thread per (table, code), each a
// binary search; larger ones: one pass over the sorted keys (offs_fill).
constexpr int OFFS_MAX_BITS = 20;
constexpr uint32_t OFFS_SEARCH_MAX = 1u << 20;  // entries up to which bucket starts are searched
// Bucket starts [L][2^b + 1] of a small table set: one thread per entry, a
// binary search of its code in its table's sorted keys.
__global__ void bucket_offsets(const uint32_t* __restrict__ keys, uint32_t n, int b, uint32_t L,
                               uint32_t* __restrict__ offs) {
    const uint32_t per = (1u << b) + 1u;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)L * per) return;
    const uint32_t t = (uint32_t)(i / per), c = (uint32_t)(i % per);
    uint32_t lo = t * n, hi = lo + n;
    if (c == per - 1) {
        lo = hi;  // the table's end
    } else {
        const uint32_t k = (t << b) | c;
        while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (keys[mid] < k) lo = mid + 1; else hi = mid;
        }
    }
    offs[i] = lo;
}

// The same for large table sets (b = 20: 2^20 entries and more per index), in
// one pass over the sorted keys: the thread of position i of table t (key
// code k, previous code p) writes t*n + i to the entries of every code in
// (p, k] (in [0, k] at i = 0), and the thread of position n writes the
// table's end to those after the last code, entry 2^b included, so every
// entry is written exactly once.  Gaps of up to 4 codes are written by their
// thread (a wave loops as long as its longest such gap); longer ones are
// queued in LDS and each written by one wave of the block, 64 entries per
// store.  (One thread per entry with a binary search, as below 2^20
// entries, costs ~20 dependent loads per entry here; a load-balanced fill
// of the block's entries stalls on blocks whose positions own long runs of
// empty codes, which clustered codes produce.)
constexpr uint32_t FILL_OWN = 4;
__global__ __launch_bounds__(256) void offs_fill(const uint32_t* __restrict__ keys, uint32_t n, uint32_t mask,
                                                 uint32_t per, uint32_t* __restrict__ offs) {
    __shared__ uint32_t q_lo[256], q_hi[256], q_v[256];
    __shared__ uint32_t q_n;
    if (threadIdx.x == 0) q_n = 0;
    __syncthreads();
    const uint32_t t = blockIdx.y, i = blockIdx.x * 256u + threadIdx.x;  // positions 0..n
    const uint32_t* tk = keys + (size_t)t * n;
    uint32_t* to = offs + (size_t)t * per;
    const uint32_t v = t * n + i;
    uint32_t lo = 0, hi = 0;  // codes [lo, hi) take v
    if (i <= n) {
        const uint32_t prev = i > 0 ? (tk[i - 1] & mask) : 0u;
        const uint32_t cur = i < n ? (tk[i] & mask) : 0u;
        if (i == n) {
            lo = prev + 1u;  // n > 0 here
            hi = per;
        } else if (i == 0 || cur != prev) {
            lo = i == 0 ? 0u : prev + 1u;
            hi = cur + 1u;
        }
    }
    if (hi - lo <= FILL_OWN) {
        for (uint32_t c = lo; c < hi; c++) to[c] = v;
    } else {
        const uint32_t slot = atomicAdd(&q_n, 1u);
        q_lo[slot] = lo;
        q_hi[slot] = hi;
        q_v[slot] = v;
    }
    __syncthreads();
    const uint32_t nq = q_n, lane = threadIdx.x & 63u;
    for (uint32_t g = threadIdx.x >> 6; g < nq; g += 4u)
        for (uint32_t c = q_lo[g] + lane; c < q_hi[g]; c += 64u) to[c] = q_v[g];
}

// Sum over buckets of size^2 (= sum over sorted positions of the size of the
// bucket holding them).  Each block writes its partial sum to part[block]
// (the host adds them: no atomics, the same total on every run).
__device__ inline void block_sum_u64(unsigned long long s, unsigned long long* part) {
    __shared__ unsigned long long wsum[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void bucket_sq(const uint32_t* __restrict__ keys, uint32_t n,
                                                 unsigned long long* __restrict__ part) {
    // sum over sorted positions p of 2 (p - s(p)) + 1, s(p) the first position
    // of p's bucket: over a bucket of size m that is m^2.  Block = 4096
    // positions of one table (grid row), 16 consecutive ones per thread.
    __shared__ long long wmax[4];
    __shared__ uint32_t first_start;
    const uint32_t* tk = keys + (size_t)blockIdx.y * n;  // this table's sorted keys
    const uint32_t cbase = blockIdx.x * 4096u, p0 = cbase + threadIdx.x * 16u;
    uint32_t k[16];
#pragma unroll
    for (int i = 0; i < 16; i++) k[i] = p0 + i < n ? tk[p0 + i] : 0u;
    const uint32_t prev = p0 > 0 && p0 - 1 < n ? tk[p0 - 1] : 0u;
    // last bucket start at or before each position inside the thread (-1: before p0)
    long long last = -1;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t p = p0 + i;
        if (p < n && (p == 0 || k[i] != (i ? k[i - 1] : prev))) last = (long long)p;
    }
    // exclusive max over the block's threads of their last start
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long incl = last;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(incl, o);
        if (lane >= o) incl = max(incl, y);
    }
    if (lane == 63) wmax[w] = incl;
    if (threadIdx.x == 0) {  // start of the bucket holding the chunk's first position
        const uint32_t k0 = tk[cbase];
        first_start = lower_bound_u32(tk, 0, cbase, k0);
    }
    __syncthreads();
    long long carry = __shfl_up(incl, 1);
    if (lane == 0) carry = -1;
    for (int v = 0; v < w; v++) carry = max(carry, wmax[v]);
    uint32_t s_run = carry >= 0 ? (uint32_t)carry : first_start;
    unsigned long long sum = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t p = p0 + i;
        if (p < n) {
            if (p == 0 || k[i] != (i ? k[i - 1] : prev)) s_run = p;
            sum += 2ull * (p - s_run) + 1ull;
        }
    }
    block_sum_u64(sum, part);
}

// The same sum from the bucket starts ([L][2^b + 1] positions).
__global__ __launch_bounds__(256) void bucket_sq_offs(const uint32_t* __restrict__ offs, int b, uint32_t L,
                                                      unsigned long long* __restrict__ part) {
    const uint64_t nb = (uint64_t)L << b;
    unsigned long long s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t o = (i >> b) * ((1ull << b) + 1) + (i & ((1ull << b) - 1));
        const unsigned long long len = offs[o + 1] - offs[o];
        s += len * len;
    }
    block_sum_u64(s, part);
}

// The hyperplanes of every table as MFMA B operands: 32 rows per table,
// planes [nbits, 32) zero.
int pack_planes(sift_context* h, sift_lsh_index* ix, int* flag) {
    const int L = ix->L, nbits = ix->nbits;
    std::vector<float> hp((size_t)L * 32 * 128, 0.f);
    for (int t = 0; t < L; t++)
        std::copy(ix->planes.begin() + (size_t)t * nbits * 128, ix->planes.begin() + (size_t)(t + 1) * nbits * 128,
                  hp.begin() + (size_t)t * 32 * 128);
    DevBuf hp_dev;
    SIFT_TRY(reserve(h, hp_dev, hp.size() * 4));
    int rc = SIFT_OK;
    if (hipMemcpyAsync(hp_dev.p, hp.data(), hp.size() * 4, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        rc = set_error(h, SIFT_ERR_HIP, "copy hyperplanes");
    if (rc == SIFT_OK)
        rc = pack_rows(h, hp_dev.p, SIFT_FLOAT_DEVICE, (uint32_t)L * 32, (uint32_t)L * 32, KIND_BF16, PACK_PLAIN,
                       ix->hp, flag);
    // int8 B operands for the build pass (the planes are integers in [-127,127]):
    // lane l holds plane l % 32, dims 32g + 16(l / 32) + 0..15 of group g
    const size_t plane_bytes = (size_t)L * PH_TABLE * 16;
    std::vector<int8_t> p8(plane_bytes + (size_t)L * 32 * 4, 0);
    std::vector<int> thr((size_t)L * 32);
    for (int t = 0; t < L; t++)
        for (int p = 0; p < 32; p++) {
            int sum = 0;
            for (int d = 0; d < 128; d++) {
                const int v = (int)hp[((size_t)t * 32 + p) * 128 + d];
                sum += v;
                const int g = d / 32, half = (d / 16) & 1, lane = p + 32 * half;
                p8[(((size_t)t * 4 + g) * 64 + lane) * 16 + (d & 15)] = (int8_t)v;
            }
            thr[t * 32 + p] = -128 * sum;
        }
    std::memcpy(p8.data() + plane_bytes, thr.data(), thr.size() * 4);
    if (rc == SIFT_OK) rc = reserve(h, ix->hp_i8, p8.size());
    if (rc == SIFT_OK && hipMemcpyAsync(ix->hp_i8.p, p8.data(), p8.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess)
        rc = set_error(h, SIFT_ERR_HIP, "copy hyperplanes");
    if (rc == SIFT_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = set_error(h, SIFT_ERR_HIP, "sync");
    release(hp_dev);
    return rc;
}

// Bucket starts (b <= OFFS_MAX_BITS) and the mean candidates per lookup of a
// point drawn like the base rows: sum over buckets of size^2 / (L n), i.e.
// the mean over sorted positions of the size of the bucket holding them.
int build_offsets(sift_context* h, sift_lsh_index* ix) {
    ix->cand_per_lookup = 0.0;
    if (ix->n == 0) return SIFT_OK;
    const size_t total = (size_t)ix->L * ix->n;
    const bool direct = ix->nbits <= OFFS_MAX_BITS;
    const uint32_t per = direct ? (1u << ix->nbits) + 1u : 0u;
    const uint32_t no = (uint32_t)ix->L * per;
    dim3 grid;
    if (direct)
        grid = dim3((unsigned)std::min<uint64_t>((((uint64_t)ix->L << ix->nbits) + 255) / 256, 4096));
    else
        grid = dim3((ix->n + 4095) / 4096, (unsigned)ix->L);
    const size_t nparts = (size_t)grid.x * grid.y;
    // every buffer first: an allocation between launches would stall the queue
    SIFT_TRY(reserve(h, ix->hist, nparts * 8));
    if (direct) SIFT_TRY(reserve(h, ix->offs, (size_t)no * 4));
    auto* part = static_cast<unsigned long long*>(ix->hist.p);
    {
        ProfScope ps(h, PROF_LSH_SORT);  // bucket structures count with the sort
        if (direct) {
            auto* offs = static_cast<uint32_t*>(ix->offs.p);
            const auto* keys = static_cast<const uint32_t*>(ix->keys.p);
            if (no <= OFFS_SEARCH_MAX) {  // few entries: a search each beats a pass over the keys
                bucket_offsets<<<(no + 255) / 256, 256, 0, h->stream>>>(keys, ix->n, ix->nbits, (uint32_t)ix->L,
                                                                          offs);
            } else {
                offs_fill<<<dim3(ix->n / 256 + 1, (unsigned)ix->L), 256, 0, h->stream>>>(
                    keys, ix->n, (1u << ix->nbits) - 1u, per, offs);
            }
            SIFT_HIP_TRY(h, hipGetLastError());
            bucket_sq_offs<<<grid, 256, 0, h->stream>>>(offs, ix->nbits, (uint32_t)ix->L, part);
        } else {
            bucket_sq<<<grid, 256, 0, h->stream>>>(static_cast<const uint32_t*>(ix->keys.p), ix->n, part);
        }
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    std::vector<unsigned long long> sums(nparts);
    SIFT_HIP_TRY(h, hipMemcpyAsync(sums.data(), part, nparts * 8, hipMemcpyDeviceToHost, h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    unsigned long long sum = 0;
    for (unsigned long long v : sums) sum += v;
    ix->cand_per_lookup = (double)sum / (double)total;
    return SIFT_OK;
}

// Codes of every base row for all tables, stable radix sort of
// (table << b | code, row), bucket starts, and the uint8 gather copy of the
// base.
// Scale exponent of the hash input for data that is not all integers in
// [0,255]: the largest k with max|x| * 2^k <= 255, within [-100, 100] (0 for
// an all-zero base), so the hashed integers use the range of SIFT bytes.
int hash_scale_exp(float mx) {
    if (!(mx > 0.f) || !std::isfinite(mx)) return 0;
    int k = (int)std::floor(std::log2(255.0 / (double)mx));
    while (k < 100 && std::ldexp((double)mx, k + 1) <= 255.0) k++;
    while (k > -100 && std::ldexp((double)mx, k) > 255.0) k--;
    return std::max(-100, std::min(100, k));
}

// The fused integer build pass (lsh_pack_hash) over the index's base rows.
int pack_hash_launch(sift_context* h, sift_lsh_index* ix, const void* r, int type, int* flag, uint32_t* keys) {
    const uint32_t n = ix->n, n_tiles = (n + 31) / 32;
    const unsigned grid = (n_tiles + PH_WAVES - 1) / PH_WAVES;
    const auto* pl = static_cast<const uint4*>(ix->hp_i8.p);
    const auto* th = reinterpret_cast<const int*>(pl + (size_t)ix->L * PH_TABLE);
    auto* u8 = static_cast<uint8_t*>(ix->base_u8.p);
    auto* nrm = static_cast<int*>(ix->base_nrm.p);
    const dim3 blk(64 * PH_WAVES);
    switch (type) {
        case SIFT_FLOAT_HOST:
        case SIFT_FLOAT_DEVICE:
            lsh_pack_hash<SIFT_FLOAT_DEVICE><<<grid, blk, 0, h->stream>>>(r, n, n_tiles, pl, th, ix->L, ix->nbits,
                                                                         u8, nrm, flag, keys);
            break;
        case SIFT_HALF_DEVICE:
            lsh_pack_hash<SIFT_HALF_DEVICE><<<grid, blk, 0, h->stream>>>(r, n, n_tiles, pl, th, ix->L, ix->nbits,
                                                                        u8, nrm, flag, keys);
            break;
        case SIFT_BF16_DEVICE:
            lsh_pack_hash<SIFT_BF16_DEVICE><<<grid, blk, 0, h->stream>>>(r, n, n_tiles, pl, th, ix->L, ix->nbits,
                                                                        u8, nrm, flag, keys);
            break;
        case SIFT_U8_DEVICE:
            lsh_pack_hash<SIFT_U8_DEVICE><<<grid, blk, 0, h->stream>>>(r, n, n_tiles, pl, th, ix->L, ix->nbits,
                                                                      u8, nrm, flag, keys);
            break;
        default: return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    }
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

// flag: the handle's LSH flag words ([0] exactness, [1] scratch, [3] max |x|).
// Sorts every table's (key, row) pairs (keys written, rows implicit) and
// builds the bucket structures; the sorted arrays end in ix->keys / ix->vals.
int sort_tables(sift_context* h, sift_lsh_index* ix) {
    SIFT_TRY(sort_reserve(h, ix->n, (uint32_t)ix->L, ix->hist, ix->levels));
    auto* keys = static_cast<uint32_t*>(ix->keys.p);
    auto* vals = static_cast<uint32_t*>(ix->vals.p);
    auto* tk = static_cast<uint32_t*>(ix->scratch_k.p);
    auto* tv = static_cast<uint32_t*>(ix->scratch_v.p);
    {  // each table sorted in its own segment; the row ids come from positions
        ProfScope ps(h, PROF_LSH_SORT);
        SIFT_TRY(radix_sort_pairs(h, &keys, &vals, &tk, &tv, ix->n, (uint32_t)ix->L, ix->nbits, ix->hist,
                                  ix->levels, true));
    }
    if (keys != ix->keys.p) {
        std::swap(ix->keys, ix->scratch_k);
        std::swap(ix->vals, ix->scratch_v);
    }
    return build_offsets(h, ix);  // ends with the stream synchronized
}

int build_tables(sift_context* h, sift_lsh_index* ix, const void* r, int type, int* flag, bool force_float) {
    const uint32_t r_n = ix->n;
    int rc = SIFT_OK;
    do {
        const uint32_t rows_pad = std::max<uint32_t>(32, ((r_n + 31) / 32) * 32);
        const size_t total = (size_t)r_n * ix->L;
        for (DevBuf* b : {&ix->keys, &ix->vals, &ix->scratch_k, &ix->scratch_v})
            if ((rc = reserve(h, *b, std::max<size_t>(total, 1) * 4)) != SIFT_OK) break;
        if (rc != SIFT_OK) break;
        // the sort's and the bucket starts' scratch too, before the first launch
        if ((rc = sort_reserve(h, r_n, (uint32_t)ix->L, ix->hist, ix->levels)) != SIFT_OK) break;
        if (r_n && ix->nbits <= OFFS_MAX_BITS) {
            const uint32_t no = (uint32_t)ix->L * ((1u << ix->nbits) + 1u);
            if ((rc = reserve(h, ix->offs, (size_t)no * 4)) != SIFT_OK) break;
        }
        auto* keys = static_cast<uint32_t*>(ix->keys.p);
        if (!force_float && ix->hash_kind != SIFT_LSH_LEARNED) {
            // hyperplanes: one pass writes the gather copy and the unsorted keys of
            // every table; the sort follows without waiting for the exactness flag,
            // and a base that is not all integers in [0,255] is built again in float
            // mode
            if ((rc = reserve(h, ix->base_u8, std::max<size_t>((size_t)r_n * 128, 1))) != SIFT_OK) break;
            if ((rc = reserve(h, ix->base_nrm, std::max<size_t>((size_t)r_n * 4, 1))) != SIFT_OK) break;
            if (r_n) {
                ProfScope ps(h, PROF_LSH_PROJ);
                if ((rc = pack_hash_launch(h, ix, r, type, flag, keys)) != SIFT_OK) break;
            }
            if ((rc = sort_tables(h, ix)) != SIFT_OK) break;
            int bad = 0;
            if (hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess) {
                rc = set_error(h, SIFT_ERR_HIP, "build flag");
                break;
            }
            if (bad) return build_tables(h, ix, r, type, flag, true);
            release(ix->scratch_k);
            release(ix->scratch_v);
            return SIFT_OK;
        }
        bool float_mode = force_float;
        if (!float_mode) {  // learned hash: uint8 gather copy, exactness check, plain tiles
            {
                ProfScope ps(h, PROF_PACK);
                if ((rc = to_u8(h, r, type, r_n, ix->base_u8, ix->base_nrm, flag)) != SIFT_OK) break;
            }
            if ((rc = pack_rows(h, r, type, r_n, rows_pad, KIND_BF16, PACK_PLAIN, h->b_set, flag)) != SIFT_OK) break;
            int bad = 0;
            if (hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess) {
                rc = set_error(h, SIFT_ERR_HIP, "build flag");
                break;
            }
            float_mode = bad != 0;
        }
        if (float_mode) {  // quantised hash input, float32 gather copy
            if (ix->hash_kind == SIFT_LSH_LEARNED) {
                rc = set_error(h, SIFT_ERR_NOT_EXACT, "sift_lsh_build_learned: the learned hash needs integer descriptors in [0,255]");
                break;
            }
            release(ix->base_u8);
            release(ix->base_nrm);
            uint32_t mx = 0;
            auto* mx_dev = reinterpret_cast<uint32_t*>(flag + 3);
            if (hipMemsetAsync(mx_dev, 0, 4, h->stream) != hipSuccess) { rc = set_error(h, SIFT_ERR_HIP, "memset"); break; }
            {
                ProfScope ps(h, PROF_PACK);
                if ((rc = abs_max(h, r, type, r_n, mx_dev)) != SIFT_OK) break;
            }
            if (hipMemcpyAsync(&mx, mx_dev, 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess) {
                rc = set_error(h, SIFT_ERR_HIP, "build max");
                break;
            }
            ix->float_mode = true;
            ix->qexp = hash_scale_exp(__builtin_bit_cast(float, mx));
            ProfScope ps(h, PROF_PACK);
            if ((rc = to_f32(h, r, type, r_n, ix->base_f32)) != SIFT_OK) break;
            if ((rc = pack_quant(h, r, type, r_n, rows_pad, ix->qexp, h->b_set, flag + 1)) != SIFT_OK) break;
        }
        {
            ProfScope ps(h, PROF_LSH_PROJ);
            if ((rc = hash_codes(h, ix, h->b_set, r_n, keys, r_n, true)) != SIFT_OK) break;
        }
        if ((rc = sort_tables(h, ix)) != SIFT_OK) break;
        release(ix->scratch_k);
        release(ix->scratch_v);
    } while (0);
    return rc;
}

// Query rows into sc->q_tiles the way the index hashes its base: plain BF16
// (integer index) or the quantised input of a float index.
int pack_for_hash(sift_context* h, const sift_lsh_index* ix, const void* q, int type, uint32_t q_n, LshScratch* sc,
                  int* flag) {
    const uint32_t rows_pad = ((q_n + 31) / 32) * 32;
    if (ix->float_mode) return pack_quant(h, q, type, q_n, rows_pad, ix->qexp, sc->q_tiles, flag + 1);
    return pack_rows(h, q, type, q_n, rows_pad, KIND_BF16, PACK_PLAIN, sc->q_tiles, flag);
}

}  // namespace

extern "C" {

void sift_lsh_state_free(sift_context* h) {
    if (!h || !h->lsh_state) return;
    auto* s = static_cast<LshScratch*>(h->lsh_state);
    for (DevBuf* b : {&s->q_u8, &s->q_nrm, &s->qcodes, &s->qbase, &s->qmarg, &s->flag, &s->q_tiles.tiles, &s->q_tiles.norm,
                      &s->q_tiles.norm_lane, &s->fb_nc, &s->fb_dist, &s->fb_list, &s->fb_count, &s->fb_map, &s->fb_out,
                      &s->fb_q.tiles, &s->fb_q.norm, &s->fb_q.norm_lane, &s->ukeys, &s->uvals, &s->utk, &s->utv,
                      &s->uhist, &s->upart, &s->ucnt})
        release(*b);
    for (auto& l : s->ulevels) release(l);
    delete s;
    h->lsh_state = nullptr;
}

}  // extern "C"

namespace {
// force_float: build the float mode whatever the base holds (lsh_gpu, when the
// queries are not all integers in [0,255]).
int lsh_build_impl(sift_handle_t h, const void* r, int type, uint32_t r_n, int nbits, int L, int hash_kind,
                   uint64_t seed, bool force_float, sift_lsh_index_t* out) {
    if (!h || !out) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    *out = nullptr;
    if (type < SIFT_FLOAT_HOST || type > SIFT_U8_DEVICE) return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    if (nbits < 1 || nbits > 32) return set_error(h, SIFT_ERR_INVALID, "nbits must be in [1,32]");
    if (L < 1 || L > 256) return set_error(h, SIFT_ERR_INVALID, "L must be in [1,256]");
    int tbits = 0;
    while ((1 << tbits) < L) tbits++;
    if (nbits + tbits > 32) return set_error(h, SIFT_ERR_UNSUPPORTED, "nbits + log2(L) must be <= 32");
    if (hash_kind != SIFT_LSH_HYPERPLANE && hash_kind != SIFT_LSH_HYPERPLANE_PM1)
        return set_error(h, SIFT_ERR_UNSUPPORTED, "hash kind not supported by sift_lsh_build");
    if ((uint64_t)r_n * (uint64_t)L >= (1ull << 32)) return set_error(h, SIFT_ERR_UNSUPPORTED, "r_n * L too large");
    if (r_n > 0 && !r) return set_error(h, SIFT_ERR_INVALID, "null base");
    if (type == SIFT_FLOAT_HOST) {
        SIFT_TRY(stage_host_input(h, r, r_n, h->stage_r, &r));
        type = SIFT_FLOAT_DEVICE;
    }

    auto* ix = new sift_lsh_index();
    ix->h = h;
    ix->n = r_n;
    ix->nbits = nbits;
    ix->L = L;
    ix->hash_kind = hash_kind;
    ix->seed = seed;
    make_planes(ix->planes, L, nbits, hash_kind, seed);
    int rc = SIFT_OK;
    do {
        LshScratch* sc = scratch_of(h);
        if ((rc = reserve(h, sc->flag, 16)) != SIFT_OK) break;
        int* flag = static_cast<int*>(sc->flag.p);
        if (hipMemsetAsync(flag, 0, 8, h->stream) != hipSuccess) { rc = set_error(h, SIFT_ERR_HIP, "memset"); break; }
        if ((rc = pack_planes(h, ix, flag + 1)) != SIFT_OK) break;
        rc = build_tables(h, ix, r, type, flag, force_float);
    } while (0);
    if (rc != SIFT_OK) {
        free_index(ix);
        return rc;
    }
    *out = ix;
    return SIFT_OK;
}
}  // namespace

extern "C" {

int sift_lsh_build(sift_handle_t h, const void* r, int type, uint32_t r_n, int nbits, int L,
                   int hash_kind, uint64_t seed, sift_lsh_index_t* out) {
    return lsh_build_impl(h, r, type, r_n, nbits, L, hash_kind, seed, false, out);
}

int sift_lsh_codes(sift_handle_t h, sift_lsh_index_t ix, const void* q, int type, uint32_t q_n,
                   uint32_t* codes) {
    if (!h || !ix) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (q_n == 0) return SIFT_OK;
    if (type == SIFT_FLOAT_HOST) SIFT_TRY(stage_host_input(h, q, q_n, h->stage_q, &q));
    LshScratch* sc = scratch_of(h);
    SIFT_TRY(reserve(h, sc->flag, 16));
    int* flag = static_cast<int*>(sc->flag.p);
    SIFT_TRY(pack_for_hash(h, ix, q, type, q_n, sc, flag));
    SIFT_TRY(hash_codes(h, ix, sc->q_tiles, q_n, codes, q_n, false));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SIFT_OK;
}

int sift_lsh_query(sift_handle_t h, sift_lsh_index_t ix, const void* q, int type, uint32_t q_n,
                   uint32_t* idx, float* dist, uint32_t* ncand, uint32_t* matches, float ratio,
                   int blocking) {
    if (!h || !ix) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (type < SIFT_FLOAT_HOST || type > SIFT_U8_DEVICE) return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    if (q_n == 0) return SIFT_OK;
    if (!q) return set_error(h, SIFT_ERR_INVALID, "null queries");
    if (type == SIFT_FLOAT_HOST) {
        SIFT_TRY(stage_host_input(h, q, q_n, h->stage_q, &q));
        type = SIFT_FLOAT_DEVICE;
    }
    LshScratch* sc = scratch_of(h);
    SIFT_TRY(reserve(h, sc->flag, 16));
    int* flag = static_cast<int*>(sc->flag.p);
    SIFT_HIP_TRY(h, hipMemsetAsync(flag, 0, 8, h->stream));
    if (!ix->float_mode) {
        ProfScope ps(h, PROF_LSH_AUX);  // the queries' uint8 copy
        SIFT_TRY(to_u8(h, q, type, q_n, sc->q_u8, sc->q_nrm, flag));
    }
    SIFT_TRY(pack_for_hash(h, ix, q, type, q_n, sc, flag));
    const uint32_t* qcodes = nullptr;
    {
        ProfScope ps(h, PROF_LSH_PROJ);
        SIFT_TRY(query_codes(h, ix, sc, q_n, &qcodes));
    }
    RerankArgs a;
    a.base_u8 = static_cast<const uint8_t*>(ix->base_u8.p);
    a.base_nrm = static_cast<const int*>(ix->base_nrm.p);
    a.q_u8 = static_cast<const uint8_t*>(sc->q_u8.p);
    a.keys = static_cast<const uint32_t*>(ix->keys.p);
    a.vals = static_cast<const uint32_t*>(ix->vals.p);
    a.qcodes = qcodes;
    a.offs = ix->offs.p ? static_cast<const uint32_t*>(ix->offs.p) : nullptr;
    a.n = ix->n;
    a.q_n = q_n;
    a.L = ix->L;
    a.b = ix->nbits;
    a.P = ix->probes;
    a.idx = idx;
    a.dist = dist;
    a.ncand = ncand;
    a.matches = matches;
    a.ratio = ratio;
    if (ix->float_mode) {  // any query data; sequential FP32 distances from the float copy
        ProfScope ps(h, PROF_LSH_QUERY);
        RerankF32Args f{a, q, type, static_cast<const float*>(ix->base_f32.p)};
        lsh_rerank_f32<<<(q_n + 3) / 4, 256, 0, h->stream>>>(f);
        SIFT_HIP_TRY(h, hipGetLastError());
        if (blocking) SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
        return SIFT_OK;
    }
    {
        ProfScope ps(h, PROF_LSH_QUERY);
        if (bucket_ordered(ix, q_n)) {
            SIFT_TRY(rerank_units(h, ix, sc, a));
        } else {
            lsh_rerank<<<(q_n + 3) / 4, 256, 0, h->stream>>>(a);
            SIFT_HIP_TRY(h, hipGetLastError());
        }
    }
    if (blocking) SIFT_TRY(check_flag(h, flag, "sift_lsh_query"));
    return SIFT_OK;
}

int sift_lsh_set_probes(sift_lsh_index_t ix, int probes) {
    if (!ix) return set_error(nullptr, SIFT_ERR_INVALID, "null index");
    if (probes < 0 || probes > std::min(ix->nbits, MAX_PROBES))
        return set_error(ix->h, SIFT_ERR_INVALID, "probes must be in [0, min(nbits, 16)]");
    ix->probes = probes;
    return SIFT_OK;
}

int sift_lsh_bucket_stat(sift_lsh_index_t ix, double* cand_per_lookup) {
    if (!ix || !cand_per_lookup) return set_error(ix ? ix->h : nullptr, SIFT_ERR_INVALID, "null index/output");
    *cand_per_lookup = ix->cand_per_lookup;
    return SIFT_OK;
}

int sift_lsh_probe_codes(sift_handle_t h, sift_lsh_index_t ix, const void* q, int type, uint32_t q_n,
                         uint32_t* codes) {
    if (!h || !ix) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (type < SIFT_FLOAT_HOST || type > SIFT_U8_DEVICE) return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    if (q_n == 0) return SIFT_OK;
    if (!q || !codes) return set_error(h, SIFT_ERR_INVALID, "null pointer");
    if (type == SIFT_FLOAT_HOST) SIFT_TRY(stage_host_input(h, q, q_n, h->stage_q, &q));
    LshScratch* sc = scratch_of(h);
    SIFT_TRY(reserve(h, sc->flag, 16));
    SIFT_TRY(pack_for_hash(h, ix, q, type, q_n, sc, static_cast<int*>(sc->flag.p)));
    const uint32_t* pc = nullptr;
    SIFT_TRY(query_codes(h, ix, sc, q_n, &pc));
    SIFT_HIP_TRY(h, hipMemcpyAsync(codes, pc, (size_t)ix->L * (ix->probes + 1) * q_n * 4, hipMemcpyDeviceToDevice,
                                   h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SIFT_OK;
}

int sift_lsh_build_learned(sift_handle_t h, const void* r, int type, uint32_t r_n, const sift_mlp_t* mlps,
                           int L, sift_lsh_index_t* out) {
    if (!h || !out || !mlps) return set_error(h, SIFT_ERR_INVALID, "null handle/index/networks");
    *out = nullptr;
    if (type < SIFT_FLOAT_HOST || type > SIFT_U8_DEVICE) return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    if (L < 1 || L > 256) return set_error(h, SIFT_ERR_INVALID, "L must be in [1,256]");
    for (int t = 0; t < L; t++)
        if (!mlps[t]) return set_error(h, SIFT_ERR_INVALID, "null network");
    const int nbits = mlp_nbits(mlps[0]);
    for (int t = 1; t < L; t++)
        if (mlp_nbits(mlps[t]) != nbits) return set_error(h, SIFT_ERR_INVALID, "networks differ in bits");
    int tbits = 0;
    while ((1 << tbits) < L) tbits++;
    if (nbits + tbits > 32) return set_error(h, SIFT_ERR_UNSUPPORTED, "nbits + log2(L) must be <= 32");
    if ((uint64_t)r_n * (uint64_t)L >= (1ull << 32)) return set_error(h, SIFT_ERR_UNSUPPORTED, "r_n * L too large");
    if (r_n > 0 && !r) return set_error(h, SIFT_ERR_INVALID, "null base");
    if (type == SIFT_FLOAT_HOST) SIFT_TRY(stage_host_input(h, r, r_n, h->stage_r, &r));
    auto* ix = new sift_lsh_index();
    ix->h = h;
    ix->n = r_n;
    ix->nbits = nbits;
    ix->L = L;
    ix->hash_kind = SIFT_LSH_LEARNED;
    ix->mlps.assign(mlps, mlps + L);
    int rc = SIFT_OK;
    do {
        LshScratch* sc = scratch_of(h);
        if ((rc = reserve(h, sc->flag, 16)) != SIFT_OK) break;
        int* flag = static_cast<int*>(sc->flag.p);
        if (hipMemsetAsync(flag, 0, 8, h->stream) != hipSuccess) { rc = set_error(h, SIFT_ERR_HIP, "memset"); break; }
        rc = build_tables(h, ix, r, type, flag, false);
    } while (0);
    if (rc != SIFT_OK) {
        free_index(ix);
        return rc;
    }
    *out = ix;
    return SIFT_OK;
}

int sift_lsh_query_fallback(sift_handle_t h, sift_lsh_index_t ix, sift_bf_index_t bf, const void* q, int type,
                            uint32_t q_n, uint32_t min_cand, float accept_ratio, uint32_t* idx, float* dist,
                            uint32_t* ncand, uint32_t* matches, float ratio, uint32_t* n_fallback) {
    if (!h || !ix || !bf) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (bf->r_n != ix->n) return set_error(h, SIFT_ERR_INVALID, "the brute-force index holds another base");
    if (!(accept_ratio >= 0.f && accept_ratio <= 1.f))
        return set_error(h, SIFT_ERR_INVALID, "accept_ratio must be in [0, 1] (0: count rule only)");
    if (n_fallback) *n_fallback = 0;
    if (q_n == 0) return SIFT_OK;
    if (!q) return set_error(h, SIFT_ERR_INVALID, "null queries");
    if (type == SIFT_FLOAT_HOST) {
        SIFT_TRY(stage_host_input(h, q, q_n, h->stage_q, &q));
        type = SIFT_FLOAT_DEVICE;
    }
    LshScratch* sc = scratch_of(h);
    SIFT_TRY(reserve(h, sc->fb_nc, (size_t)q_n * 4));
    uint32_t* nc = ncand ? ncand : static_cast<uint32_t*>(sc->fb_nc.p);
    float* dq = dist;  // the ratio rule reads the LSH distances
    if (!dq && accept_ratio > 0.f) {
        SIFT_TRY(reserve(h, sc->fb_dist, (size_t)q_n * 8));
        dq = static_cast<float*>(sc->fb_dist.p);
    }
    SIFT_TRY(sift_lsh_query(h, ix, q, type, q_n, idx, dq, nc, matches, ratio, 0));
    // queries the LSH answer does not settle, in increasing order (flags,
    // scan, compaction on the device; only the count comes back)
    SIFT_TRY(reserve(h, sc->fb_list, (size_t)q_n * 4 + 4));
    auto* pos = static_cast<uint32_t*>(sc->fb_list.p);
    SIFT_TRY(reserve(h, sc->fb_count, 4));
    SIFT_TRY(reserve(h, sc->fb_map, (size_t)query_rows_pad(bf->kind, q_n) * 4));
    auto* map = static_cast<int32_t*>(sc->fb_map.p);
    const unsigned gq = (q_n + 255) / 256;
    {
        ProfScope ps(h, PROF_LSH_AUX);  // the list of queries for the exact path
        fallback_flags<<<gq, 256, 0, h->stream>>>(nc, dq, q_n, min_cand, accept_ratio, pos);
        SIFT_HIP_TRY(h, hipGetLastError());
        SIFT_TRY(exclusive_scan(h, pos, q_n, sc->ulevels, 0));
        fallback_compact<<<gq, 256, 0, h->stream>>>(nc, dq, q_n, min_cand, accept_ratio, pos, map,
                                                    static_cast<uint32_t*>(sc->fb_count.p));
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    uint32_t cnt = 0;
    SIFT_HIP_TRY(h, hipMemcpyAsync(&cnt, sc->fb_count.p, 4, hipMemcpyDeviceToHost, h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    // the LSH pass's own input check (the non-blocking query skipped it)
    if (!ix->float_mode) SIFT_TRY(check_flag(h, static_cast<int*>(sc->flag.p), "sift_lsh_query_fallback"));
    if (n_fallback) *n_fallback = cnt;
    if (cnt == 0) return SIFT_OK;
    const uint32_t rows_pad = query_rows_pad(bf->kind, cnt);
    if (rows_pad > cnt) {
        ProfScope ps(h, PROF_LSH_AUX);
        map_tail<<<(rows_pad - cnt + 255) / 256, 256, 0, h->stream>>>(map, cnt, rows_pad);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    // exact 2-NN of those queries against the packed base, then scatter back
    PackedSet& qs = sc->fb_q;
    SIFT_TRY(reserve(h, qs.tiles, (size_t)(rows_pad / 32) * tile_bytes(bf->kind)));
    SIFT_TRY(reserve(h, qs.norm, (size_t)rows_pad * 4));
    qs.rows = cnt;
    qs.rows_pad = rows_pad;
    qs.ntiles = rows_pad / 32;
    qs.kind = bf->kind;
    int* flag = static_cast<int*>(sc->flag.p);
    SIFT_HIP_TRY(h, hipMemsetAsync(flag, bf->coarse ? 1 : 0, 4, h->stream));
    const auto* fb_map = static_cast<const int32_t*>(sc->fb_map.p);
    {
        ProfScope ps(h, PROF_PACK);
        SIFT_TRY(pack_launch(h, q, type, 0, fb_map, bf->kind, PACK_QUERY, qs, flag));
    }
    SIFT_TRY(reserve(h, sc->fb_out, (size_t)cnt * 20));
    auto* oi = static_cast<uint32_t*>(sc->fb_out.p);
    auto* od = reinterpret_cast<float*>(oi + 2 * (size_t)cnt);
    auto* om = oi + 4 * (size_t)cnt;
    Knn2Out o{oi, od, nullptr, om, ratio};
    if (bf->kind == KIND_BF16 || bf->kind == KIND_F16 || bf->kind == KIND_F32) {
        // coarse keys for data that is not all integers in [0,255]: exact re-rank from the index's tiles
        o.coarse = flag;
        o.src = ExactSrc{q, type, bf->base.tiles.p,
                         bf->kind == KIND_F32 ? SRC_F32_TILES : (bf->kind == KIND_BF16 ? SRC_BF16_TILES : SRC_F16_TILES),
                         fb_map};
    }
    SIFT_TRY(knn2_packed(h, qs, bf->base, cnt, bf->r_n, o));
    {
        ProfScope ps(h, PROF_LSH_AUX);
        scatter_exact<<<(cnt + 255) / 256, 256, 0, h->stream>>>(fb_map, cnt, oi, od, om, idx, dist, matches);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    SIFT_TRY(check_wait(h, "sift_lsh_query_fallback"));
    if (bf->kind == KIND_I8 || bf->kind == KIND_FP8) {  // those kinds cannot hold other data
        int bad = 0;
        SIFT_HIP_TRY(h, hipMemcpy(&bad, flag, 4, hipMemcpyDeviceToHost));
        if (bad) return set_error(h, SIFT_ERR_NOT_EXACT, "sift_lsh_query_fallback: an I8 or FP8 brute-force index needs integer queries in [0,255]");
    }
    return SIFT_OK;
}

// ------------------------------------------------------- persistence ----
// (the probe count is a query setting, not saved)
// File: header, then the hash (hyperplanes [L][nbits][128] float32, or the
// L networks' parameters), the sorted keys and rows [L*n] uint32, the uint8
// gather copy [n][128] and its int32 norms [n].  Bucket offsets are rebuilt
// on load.  Native byte order.
namespace {
struct LshFileHeader {
    char magic[8];
    uint32_t version, n;
    int32_t nbits, L, hash_kind;
    uint32_t nparam;  // learned: parameters per network; else 0
    uint64_t seed;
};
constexpr char LSH_MAGIC[8] = {'S', 'I', 'F', 'T', 'L', 'S', 'H', '1'};

// Version 1: integer index (uint8 rows + int32 norms).  Version 2: float
// index (int32 scale exponent + 4 bytes reserved after the header, float32
// rows).
size_t lsh_file_bytes(const LshFileHeader& hd) {
    const size_t hash = hd.hash_kind == SIFT_LSH_LEARNED ? (size_t)hd.L * hd.nparam * 4
                                                          : (size_t)hd.L * hd.nbits * 128 * 4;
    const size_t rows = hd.version == 2 ? 8 + (size_t)hd.n * 512 : (size_t)hd.n * 132;
    return sizeof(LshFileHeader) + hash + (size_t)hd.L * hd.n * 8 + rows;
}
}  // namespace

int sift_lsh_save(sift_lsh_index_t ix, const char* path) {
    if (!ix || !path) return set_error(ix ? ix->h : nullptr, SIFT_ERR_INVALID, "null index/path");
    sift_context* h = ix->h;
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    LshFileHeader hd{};
    std::memcpy(hd.magic, LSH_MAGIC, 8);
    hd.version = ix->float_mode ? 2 : 1;
    hd.n = ix->n;
    hd.nbits = ix->nbits;
    hd.L = ix->L;
    hd.hash_kind = ix->hash_kind;
    hd.nparam = ix->hash_kind == SIFT_LSH_LEARNED ? (uint32_t)SIFT_MLP_NPARAM : 0u;
    hd.seed = ix->seed;
    std::vector<float> hash;
    if (ix->hash_kind == SIFT_LSH_LEARNED) {
        hash.resize((size_t)ix->L * SIFT_MLP_NPARAM);
        for (int t = 0; t < ix->L; t++) SIFT_TRY(sift_mlp_get_params(ix->mlps[t], hash.data() + (size_t)t * SIFT_MLP_NPARAM));
    } else {
        hash = ix->planes;
    }
    const size_t nk = (size_t)ix->L * ix->n;
    std::vector<uint32_t> keys(nk), vals(nk);
    const bool fm = ix->float_mode;
    std::vector<uint8_t> u8(fm ? 0 : (size_t)ix->n * 128);
    std::vector<int32_t> nrm(fm ? 0 : ix->n);
    std::vector<float> f32(fm ? (size_t)ix->n * 128 : 0);
    if (nk) {
        SIFT_HIP_TRY(h, hipMemcpy(keys.data(), ix->keys.p, nk * 4, hipMemcpyDeviceToHost));
        SIFT_HIP_TRY(h, hipMemcpy(vals.data(), ix->vals.p, nk * 4, hipMemcpyDeviceToHost));
    }
    if (ix->n && fm) {
        SIFT_HIP_TRY(h, hipMemcpy(f32.data(), ix->base_f32.p, f32.size() * 4, hipMemcpyDeviceToHost));
    } else if (ix->n) {
        SIFT_HIP_TRY(h, hipMemcpy(u8.data(), ix->base_u8.p, u8.size(), hipMemcpyDeviceToHost));
        SIFT_HIP_TRY(h, hipMemcpy(nrm.data(), ix->base_nrm.p, nrm.size() * 4, hipMemcpyDeviceToHost));
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(h, SIFT_ERR_INVALID, std::string("cannot write ") + path);
    bool ok = std::fwrite(&hd, sizeof hd, 1, f) == 1;
    const int32_t ext[2] = {ix->qexp, 0};
    ok = ok && (!fm || std::fwrite(ext, 4, 2, f) == 2);
    ok = ok && (hash.empty() || std::fwrite(hash.data(), 4, hash.size(), f) == hash.size());
    ok = ok && (nk == 0 || (std::fwrite(keys.data(), 4, nk, f) == nk && std::fwrite(vals.data(), 4, nk, f) == nk));
    ok = ok && (ix->n == 0 || fm || (std::fwrite(u8.data(), 1, u8.size(), f) == u8.size() &&
                                     std::fwrite(nrm.data(), 4, nrm.size(), f) == nrm.size()));
    ok = ok && (ix->n == 0 || !fm || std::fwrite(f32.data(), 4, f32.size(), f) == f32.size());
    ok = (std::fclose(f) == 0) && ok;
    return ok ? SIFT_OK : set_error(h, SIFT_ERR_INVALID, std::string("write failed: ") + path);
}

}  // extern "C"

namespace {
// An index file read into host memory and checked.
struct LshFile {
    LshFileHeader hd{};
    std::vector<float> hash;
    std::vector<uint32_t> keys, vals;
    std::vector<uint8_t> u8;
    std::vector<int32_t> nrm;
    int32_t qexp = 0;        // version 2
    std::vector<float> f32;  // version 2
};

// Read and check an index file without touching the device.  The contents
// must be what a build writes before anything reaches the device: the re-rank
// gathers rows by the stored ids and the bucket lookups assume each table's
// keys sorted under its own prefix.  Returns an empty string or the reason.
std::string read_lsh_file(const char* path, LshFile& lf) {
    const std::string p(path);
    FILE* f = std::fopen(path, "rb");
    if (!f) return "cannot open " + p;
    struct Closer { FILE* f; ~Closer() { std::fclose(f); } } closer{f};
    LshFileHeader& hd = lf.hd;
    if (std::fread(&hd, sizeof hd, 1, f) != 1 || std::memcmp(hd.magic, LSH_MAGIC, 8) != 0 ||
        (hd.version != 1 && hd.version != 2))
        return p + " is not an LSH index file";
    const bool fm = hd.version == 2;
    int tbits = 0;
    while (hd.L > 0 && (1 << tbits) < hd.L) tbits++;
    const bool learned = hd.hash_kind == SIFT_LSH_LEARNED;
    if (hd.nbits < 1 || hd.nbits > 32 || hd.L < 1 || hd.L > 256 || hd.nbits + tbits > 32 ||
        (uint64_t)hd.n * (uint64_t)hd.L >= (1ull << 32) || hd.n >= (1u << 24) ||
        (hd.hash_kind != SIFT_LSH_HYPERPLANE && hd.hash_kind != SIFT_LSH_HYPERPLANE_PM1 && !learned) ||
        (learned && hd.nparam != (uint32_t)SIFT_MLP_NPARAM) || (!learned && hd.nparam != 0) || (fm && learned))
        return p + ": bad index header";
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    if (size < 0 || (size_t)size != lsh_file_bytes(hd)) return p + ": size does not match its header";
    std::fseek(f, (long)sizeof hd, SEEK_SET);
    int32_t ext[2] = {0, 0};
    if (fm && std::fread(ext, 4, 2, f) != 2) return "short read: " + p;
    lf.qexp = ext[0];
    if (fm && (lf.qexp < -100 || lf.qexp > 100)) return p + ": bad scale exponent";
    lf.hash.resize(learned ? (size_t)hd.L * hd.nparam : (size_t)hd.L * hd.nbits * 128);
    const size_t nk = (size_t)hd.L * hd.n;
    lf.keys.resize(nk);
    lf.vals.resize(nk);
    lf.u8.resize(fm ? 0 : (size_t)hd.n * 128);
    lf.nrm.resize(fm ? 0 : hd.n);
    lf.f32.resize(fm ? (size_t)hd.n * 128 : 0);
    bool ok = std::fread(lf.hash.data(), 4, lf.hash.size(), f) == lf.hash.size();
    ok = ok && (nk == 0 || (std::fread(lf.keys.data(), 4, nk, f) == nk && std::fread(lf.vals.data(), 4, nk, f) == nk));
    ok = ok && (hd.n == 0 || fm || (std::fread(lf.u8.data(), 1, lf.u8.size(), f) == lf.u8.size() &&
                                    std::fread(lf.nrm.data(), 4, lf.nrm.size(), f) == lf.nrm.size()));
    ok = ok && (hd.n == 0 || !fm || std::fread(lf.f32.data(), 4, lf.f32.size(), f) == lf.f32.size());
    if (!ok) return "short read: " + p;
    for (float v : lf.f32)
        if (!std::isfinite(v)) return p + ": a stored row holds a value that is not finite";
    for (size_t t = 0; t < (size_t)hd.L; t++) {
        const uint32_t prefix = hd.nbits < 32 ? (uint32_t)t : 0u;
        for (size_t i = t * hd.n; i < (t + 1) * hd.n; i++) {
            if (lf.vals[i] >= hd.n) return p + ": a bucket names a row past the base";
            if ((hd.nbits < 32 ? lf.keys[i] >> hd.nbits : 0u) != prefix) return p + ": a key lies outside its table";
            if (i > t * hd.n && lf.keys[i] < lf.keys[i - 1]) return p + ": the keys of a table are not sorted";
        }
    }
    for (size_t j = 0; j < (fm ? 0 : hd.n); j++) {
        int32_t s = 0;
        for (int d = 0; d < 128; d++) s += (int32_t)lf.u8[j * 128 + d] * lf.u8[j * 128 + d];
        if (s != lf.nrm[j]) return p + ": a stored norm does not match its row";
    }
    if (!learned)
        for (float v : lf.hash)
            if (!(v >= -127.f && v <= 127.f && v == std::nearbyint(v)))
                return p + ": a hyperplane value is not an integer in [-127, 127]";
    return std::string();
}
}  // namespace

extern "C" {

int sift_lsh_check_file(const char* path) {
    if (!path) return set_error(nullptr, SIFT_ERR_INVALID, "null path");
    LshFile lf;
    const std::string why = read_lsh_file(path, lf);
    return why.empty() ? SIFT_OK : set_error(nullptr, SIFT_ERR_INVALID, why);
}

int sift_lsh_load(sift_handle_t h, const char* path, sift_lsh_index_t* out) {
    if (!h || !path || !out) return set_error(h, SIFT_ERR_INVALID, "null handle/path/index");
    *out = nullptr;
    LshFile lf;
    const std::string why = read_lsh_file(path, lf);
    if (!why.empty()) return set_error(h, SIFT_ERR_INVALID, why);
    const LshFileHeader& hd = lf.hd;
    const bool learned = hd.hash_kind == SIFT_LSH_LEARNED;
    const size_t nk = (size_t)hd.L * hd.n;
    std::vector<float>& hash = lf.hash;
    std::vector<uint32_t>& keys = lf.keys;
    std::vector<uint32_t>& vals = lf.vals;
    std::vector<uint8_t>& u8 = lf.u8;
    std::vector<int32_t>& nrm = lf.nrm;

    auto* ix = new sift_lsh_index();
    ix->h = h;
    ix->n = hd.n;
    ix->nbits = hd.nbits;
    ix->L = hd.L;
    ix->hash_kind = hd.hash_kind;
    ix->seed = hd.seed;
    int rc = SIFT_OK;
    do {
        if (learned) {
            ix->owns_mlps = true;
            for (int t = 0; t < hd.L && rc == SIFT_OK; t++) {
                sift_mlp* m = nullptr;
                if ((rc = sift_mlp_create(h, hd.nbits, 0, &m)) != SIFT_OK) break;
                ix->mlps.push_back(m);
                rc = sift_mlp_set_params(m, hash.data() + (size_t)t * hd.nparam);
            }
            if (rc != SIFT_OK) break;
        } else {
            ix->planes = hash;
            LshScratch* sc = scratch_of(h);
            if ((rc = reserve(h, sc->flag, 16)) != SIFT_OK) break;
            if (hipMemsetAsync(sc->flag.p, 0, 8, h->stream) != hipSuccess) { rc = set_error(h, SIFT_ERR_HIP, "memset"); break; }
            if ((rc = pack_planes(h, ix, static_cast<int*>(sc->flag.p) + 1)) != SIFT_OK) break;
        }
        for (DevBuf* b : {&ix->keys, &ix->vals})
            if ((rc = reserve(h, *b, std::max<size_t>(nk, 1) * 4)) != SIFT_OK) break;
        if (rc != SIFT_OK) break;
        ix->float_mode = hd.version == 2;
        ix->qexp = lf.qexp;
        if (ix->float_mode) {
            if ((rc = reserve(h, ix->base_f32, std::max<size_t>(lf.f32.size(), 1) * 4)) != SIFT_OK) break;
        } else {
            if ((rc = reserve(h, ix->base_u8, std::max<size_t>(u8.size(), 1))) != SIFT_OK) break;
            if ((rc = reserve(h, ix->base_nrm, std::max<size_t>(nrm.size(), 1) * 4)) != SIFT_OK) break;
        }
        if ((nk && (hipMemcpy(ix->keys.p, keys.data(), nk * 4, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(ix->vals.p, vals.data(), nk * 4, hipMemcpyHostToDevice) != hipSuccess)) ||
            (hd.n && !ix->float_mode &&
             (hipMemcpy(ix->base_u8.p, u8.data(), u8.size(), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(ix->base_nrm.p, nrm.data(), nrm.size() * 4, hipMemcpyHostToDevice) != hipSuccess)) ||
            (hd.n && ix->float_mode &&
             hipMemcpy(ix->base_f32.p, lf.f32.data(), lf.f32.size() * 4, hipMemcpyHostToDevice) != hipSuccess)) {
            rc = set_error(h, SIFT_ERR_HIP, "copy index to the device");
            break;
        }
        if ((rc = build_offsets(h, ix)) != SIFT_OK) break;
        if (hipStreamSynchronize(h->stream) != hipSuccess) rc = set_error(h, SIFT_ERR_HIP, "sync");
    } while (0);
    if (rc != SIFT_OK) {
        free_index(ix);
        return rc;
    }
    *out = ix;
    return SIFT_OK;
}

int sift_lsh_destroy(sift_lsh_index_t ix) {
    if (!ix) return SIFT_OK;
    (void)hipStreamSynchronize(ix->h->stream);
    free_index(ix);
    return SIFT_OK;
}

int sift_lsh_hash_input(sift_lsh_index_t ix, int* float_mode, int* scale_exp) {
    if (!ix) return set_error(nullptr, SIFT_ERR_INVALID, "null index");
    if (float_mode) *float_mode = ix->float_mode ? 1 : 0;
    if (scale_exp) *scale_exp = ix->qexp;
    return SIFT_OK;
}

int sift_lsh_info(sift_lsh_index_t ix, uint32_t* r_n, int* nbits, int* L) {
    if (!ix) return SIFT_ERR_INVALID;
    if (r_n) *r_n = ix->n;
    if (nbits) *nbits = ix->nbits;
    if (L) *L = ix->L;
    return SIFT_OK;
}

int sift_lsh_get_hyperplanes(sift_lsh_index_t ix, float* out) {
    if (!ix || !out) return SIFT_ERR_INVALID;
    std::copy(ix->planes.begin(), ix->planes.end(), out);
    return SIFT_OK;
}

int sift_lsh_get_buckets(sift_lsh_index_t ix, uint32_t* keys, uint32_t* vals) {
    if (!ix) return SIFT_ERR_INVALID;
    const size_t bytes = (size_t)ix->n * ix->L * 4;
    if (bytes == 0) return SIFT_OK;
    if (keys && hipMemcpy(keys, ix->keys.p, bytes, hipMemcpyDeviceToHost) != hipSuccess) return SIFT_ERR_HIP;
    if (vals && hipMemcpy(vals, ix->vals.p, bytes, hipMemcpyDeviceToHost) != hipSuccess) return SIFT_ERR_HIP;
    return SIFT_OK;
}

// src/lsh_h.cu:722-1015: build + query + destroy.  stream_n is accepted for
// source compatibility; the threshold is applied as a ratio test.
int lsh_gpu(void* q_points, void* r_points, int type, uint32_t q_n, uint32_t r_n,
            uint32_t* matches, float threshold, sift_handle_t handle, int stream_n, int l,
            int lsh_type, int nbits) {
    (void)stream_n;
    if (!handle) return set_error(nullptr, SIFT_ERR_INVALID, "lsh_gpu: null handle");
    if (lsh_type != SIFT_LSH_HYPERPLANE && lsh_type != SIFT_LSH_HYPERPLANE_PM1)
        return set_error(handle, SIFT_ERR_UNSUPPORTED, "lsh_gpu: lsh_type must be a hyperplane kind");
    if (!matches && q_n) return set_error(handle, SIFT_ERR_INVALID, "lsh_gpu: null matches");
    if (!valid_type(type)) return set_error(handle, SIFT_ERR_INVALID, "lsh_gpu: unknown input type");
    if ((q_n && !q_points) || (r_n && !r_points)) return set_error(handle, SIFT_ERR_INVALID, "lsh_gpu: null input");
    const void* q = q_points;
    const void* r = r_points;
    if (type == SIFT_FLOAT_HOST) {
        SIFT_TRY(stage_host_input(handle, q, q_n, handle->stage_q, &q));
        SIFT_TRY(stage_host_input(handle, r, r_n, handle->stage_r, &r));
        type = SIFT_FLOAT_DEVICE;
    }
    // queries that are not all integers in [0,255] need the float mode of the
    // index whatever the base holds (the reference converts both to fp16,
    // src/lsh_h.cu:748-780)
    LshScratch* sc = scratch_of(handle);
    SIFT_TRY(reserve(handle, sc->flag, 16));
    int* flag = static_cast<int*>(sc->flag.p);
    SIFT_HIP_TRY(handle, hipMemsetAsync(flag, 0, 4, handle->stream));
    SIFT_TRY(check_exact(handle, q, type, q_n, flag));
    int q_bad = 0;
    SIFT_HIP_TRY(handle, hipMemcpyAsync(&q_bad, flag, 4, hipMemcpyDeviceToHost, handle->stream));
    SIFT_HIP_TRY(handle, hipStreamSynchronize(handle->stream));
    sift_lsh_index_t ix = nullptr;
    SIFT_TRY(lsh_build_impl(handle, r, type, r_n, nbits, l, lsh_type, handle->seed, q_bad != 0, &ix));
    DevBuf tmp;
    uint32_t* dst = matches;
    int rc = SIFT_OK;
    if (!device_accessible(matches)) {  // plain host memory (FLOAT_HOST callers)
        rc = reserve(handle, tmp, (size_t)q_n * 4 + 4);
        dst = static_cast<uint32_t*>(tmp.p);
    }
    if (rc == SIFT_OK)
        rc = sift_lsh_query(handle, ix, q, type, q_n, nullptr, nullptr, nullptr, dst, threshold, 1);
    if (rc == SIFT_OK && dst != matches && q_n &&
        hipMemcpy(matches, dst, (size_t)q_n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_error(handle, SIFT_ERR_HIP, "lsh_gpu: copy matches to host");
    release(tmp);
    sift_lsh_destroy(ix);
    return rc;
}

}  // extern "C"
