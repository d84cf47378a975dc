// knn2.hip — exact 2-NN: fused MFMA distance + top-2 kernel, slab merge,
// and the brute-force entry points.
//
// The reference materialises a q x r distance matrix with cuBLAS and re-reads
// it in a top-2 kernel (src/knn_brute.cu:244 then :265; :592 then :606;
// :644 then :648).  Here one kernel computes, per 32x32 MFMA tile,
//     t = ||b||^2 - 2 q.b          (rank key; ||q||^2 is constant per query)
// and folds it into a per-lane top-2, so no distance ever reaches HBM.
//
// Work decomposition (one launch):
//   block  = (query group, base slab); WAVES waves x QW blocks of 32 queries
//   wave   : query operands resident in VGPRs for the whole slab
//   block  : base tiles streamed HBM/L2 -> VGPR -> LDS (one copy shared by
//            all waves), TPS tiles per stage, next stage prefetched into
//            registers while the current one is consumed
//   lane   : accumulator layout of v_mfma_*_32x32: lane l holds query l%32 and
//            16 base rows (8*(j/4) + 4*(l/32) + j%4), so each lane keeps the
//            top-2 of its own rows for its own query; lanes l and l^32 merge
//            at the end.
//   block -> (group, slab) is XCD-aware: consecutive blocks land on different
//            XCDs, so all groups of one slab are placed on the same XCD and
//            share that XCD's L2 copy of the slab.
// Tie-break: rows reach a lane in increasing index order and the insert uses
// strict '<', merges compare (value, index) lexicographically — the result is
// the reference host_brute's (src/knn_brute.cu:816-831) lowest-index rule.
#include <algorithm>
#include <cfloat>
#include <mutex>

#include "internal.h"

namespace sift {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

template <int KIND>
struct KT;
template <>
struct KT<KIND_BF16> {
    using acc_t = f32x16;
    static constexpr int G = 8;
    __device__ static acc_t mma(uint4 a, uint4 b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
};
template <>
struct KT<KIND_F16> {
    using acc_t = f32x16;
    static constexpr int G = 8;
    __device__ static acc_t mma(uint4 a, uint4 b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    }
};
template <>
struct KT<KIND_F32> {
    using acc_t = f32x16;
    static constexpr int G = 16;
    __device__ static acc_t mma(uint4 a, uint4 b, acc_t c) {
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.x), __uint_as_float(b.x), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.y), __uint_as_float(b.y), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.z), __uint_as_float(b.z), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.w), __uint_as_float(b.w), c, 0, 0, 0);
        return c;
    }
};
template <>
struct KT<KIND_I8> {
    using acc_t = i32x16;
    static constexpr int G = 4;
    __device__ static acc_t mma(uint4 a, uint4 b, acc_t c) {
        return __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i32x4, a),
                                                     __builtin_bit_cast(i32x4, b), c, 0, 0, 0);
    }
};

// Rank keys are float for the float kinds and int32 for KIND_I8; both are
// carried as 32-bit patterns.  "none" = +inf / INT32_MAX with index UINT32_MAX.
template <int KIND>
struct Key {
    using T = float;
    __device__ static float none() { return __builtin_huge_valf(); }
};
template <>
struct Key<KIND_I8> {
    using T = int;
    __device__ static int none() { return 0x7FFFFFFF; }
};

template <typename T>
__device__ inline bool lex_less(T x, uint32_t ix, T y, uint32_t iy) {
    return x < y || (x == y && ix < iy);
}

// top-2 of {a1,a2} U {b1,b2}, both lists sorted lexicographically.
template <typename T>
__device__ inline void merge2(T& a1, uint32_t& ia1, T& a2, uint32_t& ia2, T b1, uint32_t ib1, T b2,
                              uint32_t ib2) {
    const bool bf = lex_less(b1, ib1, a1, ia1);
    const T f = bf ? b1 : a1;
    const uint32_t fi = bf ? ib1 : ia1;
    const T c1 = bf ? a1 : b1;
    const uint32_t ci1 = bf ? ia1 : ib1;
    const T c2 = bf ? b2 : a2;
    const uint32_t ci2 = bf ? ib2 : ia2;
    const bool sf = lex_less(c2, ci2, c1, ci1);
    a2 = sf ? c2 : c1;
    ia2 = sf ? ci2 : ci1;
    a1 = f;
    ia1 = fi;
}

struct FusedArgs {
    const uint4* q_tiles;   // packed queries, [qblocks][G][64]
    const uint4* b_tiles;   // packed base, [n_tiles][G][64]
    const float* b_nl;      // base norms in accumulator-lane order, [n_tiles][32]
    uint4* part;            // [n_slabs][q_stride] {key1, idx1, key2, idx2}
    uint32_t q_n, q_stride;
    int n_tiles, tiles_per_slab, n_slabs, n_groups;
};

template <int KIND, int QW, int WAVES, int TPS>
__global__ __launch_bounds__(WAVES * 64) void fused_top2(FusedArgs a) {
    using KTk = KT<KIND>;
    using acc_t = typename KTk::acc_t;
    using key_t = typename Key<KIND>::T;
    constexpr int G = KTk::G;
    constexpr int UPT = G * 64;               // 16-byte units per tile
    constexpr int NT = WAVES * 64;
    constexpr int STAGE = TPS * UPT;
    constexpr int PER_THREAD = STAGE / NT;
    static_assert(STAGE % NT == 0, "stage must split evenly over the block");
    static_assert(TPS * 32 <= NT, "one norm per thread per stage");

    __shared__ __align__(16) uint4 s_tiles[STAGE];
    __shared__ __align__(16) float s_nl[TPS * 32];

    const int b = blockIdx.x;
    const int xcd = b & 7, r = b >> 3;
    const int slab = (r / a.n_groups) * 8 + xcd;
    const int group = r % a.n_groups;
    if (slab >= a.n_slabs) return;  // uniform per block
    const int t_begin = slab * a.tiles_per_slab;
    const int t_end = min(t_begin + a.tiles_per_slab, a.n_tiles);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6, h = lane >> 5;
    const int qblk0 = (group * WAVES + wave) * QW;

    uint4 qop[QW][G];
#pragma unroll
    for (int qb = 0; qb < QW; qb++)
#pragma unroll
        for (int g = 0; g < G; g++) qop[qb][g] = a.q_tiles[((size_t)(qblk0 + qb) * G + g) * 64 + lane];

    key_t m1[QW], m2[QW];
    uint32_t i1[QW], i2[QW];
#pragma unroll
    for (int qb = 0; qb < QW; qb++) {
        m1[qb] = m2[qb] = Key<KIND>::none();
        i1[qb] = i2[qb] = 0xFFFFFFFFu;
    }

    uint4 pre[PER_THREAD];
    float pre_nl = 0.f;
    auto load_stage = [&](int ts) {
        const int avail = min(TPS, t_end - ts);
        const uint4* src = a.b_tiles + (size_t)ts * UPT;
#pragma unroll
        for (int u = 0; u < PER_THREAD; u++) {
            const int idx = u * NT + tid;
            pre[u] = (idx < avail * UPT) ? src[idx] : make_uint4(0u, 0u, 0u, 0u);
        }
        if (tid < TPS * 32) pre_nl = (tid < avail * 32) ? a.b_nl[(size_t)ts * 32 + tid] : 0.f;
    };

    if (t_begin < t_end) load_stage(t_begin);
    for (int ts = t_begin; ts < t_end; ts += TPS) {
        __syncthreads();  // previous stage fully consumed
#pragma unroll
        for (int u = 0; u < PER_THREAD; u++) s_tiles[u * NT + tid] = pre[u];
        if (tid < TPS * 32) s_nl[tid] = pre_nl;
        __syncthreads();
        if (ts + TPS < t_end) load_stage(ts + TPS);  // in flight during this stage
        const int ntile = min(TPS, t_end - ts);
#pragma unroll
        for (int tt = 0; tt < TPS; tt++) {
            if (tt >= ntile) break;
            acc_t acc[QW];
#pragma unroll
            for (int qb = 0; qb < QW; qb++) acc[qb] = acc_t{};
#pragma unroll
            for (int g = 0; g < G; g++) {
                const uint4 av = s_tiles[tt * UPT + g * 64 + lane];
#pragma unroll
                for (int qb = 0; qb < QW; qb++) acc[qb] = KTk::mma(av, qop[qb][g], acc[qb]);
            }
            // norms of this lane's 16 rows
            const float4* nlp = reinterpret_cast<const float4*>(s_nl + tt * 32 + h * 16);
            float nl[16];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float4 v = nlp[c];
                nl[4 * c] = v.x; nl[4 * c + 1] = v.y; nl[4 * c + 2] = v.z; nl[4 * c + 3] = v.w;
            }
            const uint32_t rowbase = (uint32_t)(ts + tt) * 32u + 4u * h;
#pragma unroll
            for (int qb = 0; qb < QW; qb++) {
                key_t tv[16];
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    if constexpr (KIND == KIND_I8)
                        tv[j] = __float_as_int(nl[j]) - 2 * acc[qb][j];
                    else
                        tv[j] = fmaf(acc[qb][j], -2.f, nl[j]);
                }
                key_t mn = tv[0];
#pragma unroll
                for (int j = 1; j < 16; j++) mn = min(mn, tv[j]);
                if (mn < m2[qb]) {  // rare after warm-up: some row beats the lane's 2nd best
#pragma unroll
                    for (int j = 0; j < 16; j++) {
                        const key_t v = tv[j];
                        const uint32_t row = rowbase + 8u * (j >> 2) + (j & 3);
                        const bool c1 = v < m1[qb], c2 = v < m2[qb];
                        m2[qb] = c1 ? m1[qb] : (c2 ? v : m2[qb]);
                        i2[qb] = c1 ? i1[qb] : (c2 ? row : i2[qb]);
                        m1[qb] = c1 ? v : m1[qb];
                        i1[qb] = c1 ? row : i1[qb];
                    }
                }
            }
        }
    }

    // lanes l and l^32 hold disjoint rows of the same query
#pragma unroll
    for (int qb = 0; qb < QW; qb++) {
        const key_t o1 = __shfl_xor(m1[qb], 32);
        const key_t o2 = __shfl_xor(m2[qb], 32);
        const uint32_t oi1 = __shfl_xor(i1[qb], 32);
        const uint32_t oi2 = __shfl_xor(i2[qb], 32);
        merge2(m1[qb], i1[qb], m2[qb], i2[qb], o1, oi1, o2, oi2);
        const uint32_t q = (uint32_t)(qblk0 + qb) * 32u + (uint32_t)lane;
        if (h == 0 && q < a.q_n) {
            uint4 o;
            o.x = __builtin_bit_cast(uint32_t, m1[qb]);
            o.y = i1[qb];
            o.z = __builtin_bit_cast(uint32_t, m2[qb]);
            o.w = i2[qb];
            a.part[(size_t)slab * a.q_stride + q] = o;
        }
    }
}

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
};

// One thread per query: lexicographic merge of the slab partials, then
// d = key + ||q||^2 and the output formats.
template <int KIND>
__global__ __launch_bounds__(256) void merge_top2(MergeArgs a) {
    using key_t = typename Key<KIND>::T;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.q_n) return;
    key_t m1 = Key<KIND>::none(), m2 = Key<KIND>::none();
    uint32_t i1 = 0xFFFFFFFFu, i2 = 0xFFFFFFFFu;
    for (int s = 0; s < a.n_slabs; s++) {
        const uint4 p = a.part[(size_t)s * a.q_stride + q];
        merge2(m1, i1, m2, i2, __builtin_bit_cast(key_t, p.x), p.y, __builtin_bit_cast(key_t, p.z), p.w);
    }
    float d1 = FLT_MAX, d2 = FLT_MAX;
    const bool has1 = i1 < a.r_n, has2 = i2 < a.r_n;
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

// Launch shapes per kind.  BF16/F16: 2 query blocks per wave (64 queries),
// 4 waves, 2 base tiles (64 rows) per LDS stage.
template <int KIND>
struct Shape;
template <> struct Shape<KIND_BF16> { static constexpr int QW = 2, WAVES = 4, TPS = 2; };
template <> struct Shape<KIND_F16> { static constexpr int QW = 2, WAVES = 4, TPS = 2; };
template <> struct Shape<KIND_F32> { static constexpr int QW = 2, WAVES = 4, TPS = 1; };
template <> struct Shape<KIND_I8> { static constexpr int QW = 4, WAVES = 4, TPS = 4; };

inline int queries_per_group(int kind) {
    switch (kind) {
        case KIND_BF16: return Shape<KIND_BF16>::QW * Shape<KIND_BF16>::WAVES * 32;
        case KIND_F16: return Shape<KIND_F16>::QW * Shape<KIND_F16>::WAVES * 32;
        case KIND_F32: return Shape<KIND_F32>::QW * Shape<KIND_F32>::WAVES * 32;
        default: return Shape<KIND_I8>::QW * Shape<KIND_I8>::WAVES * 32;
    }
}

template <int KIND>
int launch_fused(sift_context* h, const PackedSet& qs, const PackedSet& bs, uint32_t q_n,
                 uint32_t r_n, const Knn2Out& out) {
    using S = Shape<KIND>;
    const int qpg = S::QW * S::WAVES * 32;
    const int n_groups = (int)((q_n + qpg - 1) / qpg);
    const int n_tiles = (int)bs.ntiles;
    // ~8 resident blocks' worth per CU across 256 CUs, slabs a multiple of 8
    const int target_blocks = 256 * 8;
    int want_slabs = std::max(1, std::min(n_tiles, (target_blocks + n_groups - 1) / n_groups));
    int tiles_per_slab = std::max(1, (n_tiles + want_slabs - 1) / want_slabs);
    tiles_per_slab = ((tiles_per_slab + S::TPS - 1) / S::TPS) * S::TPS;
    int n_slabs = std::max(1, (n_tiles + tiles_per_slab - 1) / tiles_per_slab);
    n_slabs = ((n_slabs + 7) / 8) * 8;
    const uint32_t q_stride = (uint32_t)n_groups * qpg;
    SIFT_TRY(reserve(h, h->part, (size_t)n_slabs * q_stride * sizeof(uint4)));

    FusedArgs fa;
    fa.q_tiles = static_cast<const uint4*>(qs.tiles.p);
    fa.b_tiles = static_cast<const uint4*>(bs.tiles.p);
    fa.b_nl = static_cast<const float*>(bs.norm_lane.p);
    fa.part = static_cast<uint4*>(h->part.p);
    fa.q_n = q_n;
    fa.q_stride = q_stride;
    fa.n_tiles = n_tiles;
    fa.tiles_per_slab = tiles_per_slab;
    fa.n_slabs = n_slabs;
    fa.n_groups = n_groups;
    const unsigned grid = (unsigned)(n_slabs * n_groups);
    {
        ProfScope ps(h, PROF_FUSED);
        fused_top2<KIND, S::QW, S::WAVES, S::TPS><<<grid, S::WAVES * 64, 0, h->stream>>>(fa);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    MergeArgs ma;
    ma.part = static_cast<const uint4*>(h->part.p);
    ma.q_norm = static_cast<const float*>(qs.norm.p);
    ma.q_n = q_n;
    ma.r_n = r_n;
    ma.q_stride = q_stride;
    ma.n_slabs = n_slabs;
    ma.idx = out.idx;
    ma.dist = out.dist;
    ma.sorted = out.sorted;
    ma.matches = out.matches;
    ma.ratio = out.ratio;
    {
        ProfScope ps(h, PROF_MERGE);
        merge_top2<KIND><<<(q_n + 255) / 256, 256, 0, h->stream>>>(ma);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    return SIFT_OK;
}

}  // namespace

int knn2_packed(sift_context* h, const PackedSet& qs, const PackedSet& bs, uint32_t q_n,
                uint32_t r_n, const Knn2Out& out) {
    if (qs.kind != bs.kind) return set_error(h, SIFT_ERR_INVALID, "query/base packed with different kinds");
    if (q_n == 0) return SIFT_OK;
    switch (bs.kind) {
        case KIND_BF16: return launch_fused<KIND_BF16>(h, qs, bs, q_n, r_n, out);
        case KIND_F16: return launch_fused<KIND_F16>(h, qs, bs, q_n, r_n, out);
        case KIND_F32: return launch_fused<KIND_F32>(h, qs, bs, q_n, r_n, out);
        case KIND_I8: return launch_fused<KIND_I8>(h, qs, bs, q_n, r_n, out);
        default: return set_error(h, SIFT_ERR_INVALID, "bad kind");
    }
}

uint32_t query_rows_pad(int kind, uint32_t q_n) {
    const uint32_t qpg = (uint32_t)queries_per_group(kind);
    return std::max<uint32_t>(1, (q_n + qpg - 1) / qpg) * qpg;
}

}  // namespace sift

using namespace sift;

namespace {

bool valid_type(int type) { return type >= SIFT_FLOAT_HOST && type <= SIFT_U8_DEVICE; }

// Resolve SIFT_COMPUTE_* to a packed kind.  AUTO: integer data in [0,255] ->
// BF16 (exact), otherwise `fallback` (F32 for the FP32 entry points, F16 for
// the fp16 one).  Reads one device flag (a 4-byte synchronous copy).
int resolve_kind(sift_context* h, const void* q, const void* r, int type, uint32_t q_n,
                 uint32_t r_n, int compute, int fallback, int* kind) {
    switch (compute) {
        case SIFT_COMPUTE_BF16: *kind = KIND_BF16; return SIFT_OK;
        case SIFT_COMPUTE_F16: *kind = KIND_F16; return SIFT_OK;
        case SIFT_COMPUTE_F32: *kind = KIND_F32; return SIFT_OK;
        case SIFT_COMPUTE_I8: *kind = KIND_I8; return SIFT_OK;
        case SIFT_COMPUTE_AUTO: break;
        default: return set_error(h, SIFT_ERR_INVALID, "unknown compute kind " + std::to_string(compute));
    }
    if (type == SIFT_U8_DEVICE) {
        *kind = KIND_BF16;
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
    *kind = bad ? fallback : KIND_BF16;
    return SIFT_OK;
}

int knn2_impl(sift_context* h, const void* q, const void* r, int type, uint32_t q_n, uint32_t r_n,
              int compute, int fallback, const Knn2Out& out, bool blocking) {
    if (!h) return set_error(nullptr, SIFT_ERR_INVALID, "null handle");
    if (!valid_type(type)) return set_error(h, SIFT_ERR_INVALID, "unknown input type " + std::to_string(type));
    if (q_n == 0) return SIFT_OK;
    if ((!q) || (r_n > 0 && !r)) return set_error(h, SIFT_ERR_INVALID, "null input pointer");
    if (r_n >= (1u << 24)) return set_error(h, SIFT_ERR_UNSUPPORTED, "r_n must be < 2^24 (indices exact in float4 output)");
    if (type == SIFT_FLOAT_HOST) {
        SIFT_TRY(stage_host_input(h, q, q_n, h->stage_q, &q));
        SIFT_TRY(stage_host_input(h, r, r_n, h->stage_r, &r));
    }
    int kind = 0;
    SIFT_TRY(resolve_kind(h, q, r, type, q_n, r_n, compute, fallback, &kind));
    SIFT_TRY(reserve(h, h->flag, 16));
    int* flag = static_cast<int*>(h->flag.p);
    SIFT_HIP_TRY(h, hipMemsetAsync(flag, 0, 4, h->stream));
    const uint32_t rows_b = std::max<uint32_t>(32, ((r_n + 31) / 32) * 32);
    SIFT_TRY(pack_rows(h, r, type, r_n, rows_b, kind, true, h->b_set, flag));
    SIFT_TRY(pack_rows(h, q, type, q_n, query_rows_pad(kind, q_n), kind, false, h->q_set, flag));
    SIFT_TRY(knn2_packed(h, h->q_set, h->b_set, q_n, r_n, out));
    if (blocking) {
        int bad = 0;
        SIFT_HIP_TRY(h, hipMemcpyAsync(&bad, flag, 4, hipMemcpyDeviceToHost, h->stream));
        SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
        if (bad && kind == KIND_I8)
            return set_error(h, SIFT_ERR_NOT_EXACT, "SIFT_COMPUTE_I8 needs integer data in [0,255]");
    }
    return SIFT_OK;
}

std::mutex g_default_mu;
sift_context* g_default[64] = {};

sift_context* default_handle() {
    int dev = 0;
    hipGetDevice(&dev);
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
    return knn2_impl(h, q, r, type, q_n, r_n, compute, KIND_F32, o, true);
}

int sift_knn2_async(sift_handle_t h, const void* q, const void* r, int type, uint32_t q_n,
                    uint32_t r_n, int compute, uint32_t* idx, float* dist, sift_float4* sorted,
                    uint32_t* matches, float ratio) {
    Knn2Out o{idx, dist, sorted, matches, ratio};
    return knn2_impl(h, q, r, type, q_n, r_n, compute, KIND_F32, o, false);
}

// src/knn_brute.cu:24-209.  fp16 semantics for non-integer data (the
// reference converts to half); exact BF16 for integer SIFT data.
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
        hipPointerAttribute_t attr;
        bool device_ok = hipPointerGetAttributes(&attr, matches) == hipSuccess &&
                         (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
                          attr.type == hipMemoryTypeHost);
        (void)hipGetLastError();
        if (!device_ok) {
            DevBuf tmp;
            SIFT_TRY(reserve(handle, tmp, (size_t)q_n * 4));
            o.matches = static_cast<uint32_t*>(tmp.p);
            int rc = knn2_impl(handle, q_points, r_points, type, q_n, r_n, compute, KIND_F16, o, true);
            if (rc == SIFT_OK)
                rc = hipMemcpy(matches, tmp.p, (size_t)q_n * 4, hipMemcpyDeviceToHost) == hipSuccess
                         ? SIFT_OK
                         : set_error(handle, SIFT_ERR_HIP, "copy matches to host");
            release(tmp);
            return rc;
        }
    }
    o.matches = matches;
    return knn2_impl(handle, q_points, r_points, type, q_n, r_n, compute, KIND_F16, o, true);
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
    if (stream && stream != (void*)handle->stream)
        SIFT_HIP_TRY(handle, hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    const int compute = SIFT_COMPUTE_F16;
    // R arrives as fp16: pack it as such, pack q from its own type
    if (!handle || q_n == 0) return SIFT_OK;
    SIFT_TRY(reserve(handle, handle->flag, 16));
    int* flag = static_cast<int*>(handle->flag.p);
    const uint32_t rows_b = std::max<uint32_t>(32, ((r_n + 31) / 32) * 32);
    SIFT_TRY(pack_rows(handle, R, SIFT_HALF_DEVICE, r_n, rows_b, KIND_F16, true, handle->b_set, flag));
    SIFT_TRY(pack_rows(handle, q_points, type, q_n, query_rows_pad(KIND_F16, q_n), KIND_F16, false,
                       handle->q_set, flag));
    (void)compute;
    Knn2Out o;
    o.matches = matches;
    o.ratio = threshold;
    SIFT_TRY(knn2_packed(handle, handle->q_set, handle->b_set, q_n, r_n, o));
    SIFT_HIP_TRY(handle, hipStreamSynchronize(handle->stream));
    return SIFT_OK;
}

// src/knn_brute.cu:619-667: exact FP32 2-NN into float4.
void device_brute(const float* q_points, const float* r_points, int q_n, int r_n,
                  sift_float4* sorted) {
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
    o.sorted = sorted;
    int rc = knn2_impl(h, q_points, r_points, SIFT_FLOAT_DEVICE, (uint32_t)q_n, (uint32_t)r_n,
                       SIFT_COMPUTE_AUTO, KIND_F32, o, true);
    if (rc == SIFT_OK) set_error(h, SIFT_OK, "");
}

// src/knn_brute.cu:523-563.
void cublas_2nn_f(const float* q_points, const float* r_points, int q_n, int r_n,
                  sift_float4* sorted, sift_handle_t handle) {
    if (!handle) {
        set_error(nullptr, SIFT_ERR_INVALID, "cublas_2nn_f: null handle");
        return;
    }
    if (q_n < 0 || r_n < 0) {
        set_error(handle, SIFT_ERR_INVALID, "negative size");
        return;
    }
    Knn2Out o;
    o.sorted = sorted;
    int rc = knn2_impl(handle, q_points, r_points, SIFT_FLOAT_DEVICE, (uint32_t)q_n, (uint32_t)r_n,
                       SIFT_COMPUTE_AUTO, KIND_F32, o, true);
    if (rc == SIFT_OK) set_error(handle, SIFT_OK, "");
}

// src/knn_brute.cu:567-616 (dist buffer unused).
void cublas_2nn_brute_f(const float* q_points, const float* r_points, int q_n, int r_n,
                        sift_float4* sorted, float* dist, sift_handle_t handle) {
    (void)dist;
    cublas_2nn_f(q_points, r_points, q_n, r_n, sorted, handle);
}

// ------------------------------------------------- brute-force index ----
}  // extern "C"

struct sift_bf_index {
    sift_context* h;
    PackedSet base;
    uint32_t r_n;
    int kind;
};

extern "C" {

int sift_bf_index_build(sift_handle_t h, const void* r, int type, uint32_t r_n, int compute,
                        sift_bf_index_t* index) {
    if (!h || !index) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (!valid_type(type)) return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    if (r_n >= (1u << 24)) return set_error(h, SIFT_ERR_UNSUPPORTED, "r_n must be < 2^24");
    *index = nullptr;
    if (type == SIFT_FLOAT_HOST) SIFT_TRY(stage_host_input(h, r, r_n, h->stage_r, &r));
    int kind = 0;
    SIFT_TRY(resolve_kind(h, r, r, type, 0, r_n, compute, KIND_F32, &kind));
    auto* ix = new sift_bf_index();
    ix->h = h;
    ix->r_n = r_n;
    ix->kind = kind;
    SIFT_TRY(reserve(h, h->flag, 16));
    int* flag = static_cast<int*>(h->flag.p);
    hipMemsetAsync(flag, 0, 4, h->stream);
    const uint32_t rows_b = std::max<uint32_t>(32, ((r_n + 31) / 32) * 32);
    int rc = pack_rows(h, r, type, r_n, rows_b, kind, true, ix->base, flag);
    if (rc == SIFT_OK && hipStreamSynchronize(h->stream) != hipSuccess)
        rc = set_error(h, SIFT_ERR_HIP, "index build sync");
    if (rc != SIFT_OK) {
        release(ix->base.tiles);
        release(ix->base.norm);
        release(ix->base.norm_lane);
        delete ix;
        return rc;
    }
    *index = ix;
    return SIFT_OK;
}

int sift_bf_index_query(sift_handle_t h, sift_bf_index_t ix, const void* q, int type, uint32_t q_n,
                        uint32_t* idx, float* dist, sift_float4* sorted, uint32_t* matches,
                        float ratio, int blocking) {
    if (!h || !ix) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (!valid_type(type)) return set_error(h, SIFT_ERR_INVALID, "unknown input type");
    if (q_n == 0) return SIFT_OK;
    if (type == SIFT_FLOAT_HOST) SIFT_TRY(stage_host_input(h, q, q_n, h->stage_q, &q));
    SIFT_TRY(reserve(h, h->flag, 16));
    int* flag = static_cast<int*>(h->flag.p);
    SIFT_TRY(pack_rows(h, q, type, q_n, query_rows_pad(ix->kind, q_n), ix->kind, false, h->q_set, flag));
    Knn2Out o{idx, dist, sorted, matches, ratio};
    SIFT_TRY(knn2_packed(h, h->q_set, ix->base, q_n, ix->r_n, o));
    if (blocking) SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SIFT_OK;
}

int sift_bf_index_destroy(sift_bf_index_t ix) {
    if (!ix) return SIFT_OK;
    hipStreamSynchronize(ix->h->stream);
    release(ix->base.tiles);
    release(ix->base.norm);
    release(ix->base.norm_lane);
    delete ix;
    return SIFT_OK;
}

}  // extern "C"
