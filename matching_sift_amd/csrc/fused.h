// fused.h — device types shared by the fused distance + top-2 kernel
// (fused.hip) and the slab merge (knn2.hip).
#pragma once

#include "internal.h"

namespace sift {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

template <int KIND>
struct KT;
template <>
struct KT<KIND_BF16> {
    using acc_t = f32x16;
    static constexpr int TG = 9;  // groups per packed tile: 8 data groups + the norm group
    static constexpr int G = 8;   // groups the fused kernel multiplies (K = 128)
    __device__ static acc_t mma(uint4 a, uint4 b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
    static constexpr int NM = G;  // MFMA steps per tile
    __device__ static acc_t mma_at(const uint4 (&a)[G], const uint4 (&b)[G], int m, acc_t c) {
        return mma(a[m], b[m], c);
    }
};
template <>
struct KT<KIND_F16> {
    using acc_t = f32x16;
    static constexpr int TG = 9;
    static constexpr int G = 8;
    __device__ static acc_t mma(uint4 a, uint4 b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    }
    static constexpr int NM = G;  // MFMA steps per tile
    __device__ static acc_t mma_at(const uint4 (&a)[G], const uint4 (&b)[G], int m, acc_t c) {
        return mma(a[m], b[m], c);
    }
};
template <>
struct KT<KIND_F32> {
    using acc_t = f32x16;
    static constexpr int TG = 17;
    static constexpr int G = 16;
    __device__ static acc_t mma(uint4 a, uint4 b, acc_t c) {
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.x), __uint_as_float(b.x), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.y), __uint_as_float(b.y), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.z), __uint_as_float(b.z), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a.w), __uint_as_float(b.w), c, 0, 0, 0);
        return c;
    }
    static constexpr int NM = G;  // MFMA steps per tile
    __device__ static acc_t mma_at(const uint4 (&a)[G], const uint4 (&b)[G], int m, acc_t c) {
        return mma(a[m], b[m], c);
    }
};
template <>
struct KT<KIND_I8> {
    using acc_t = i32x16;
    static constexpr int TG = 4;
    static constexpr int G = 4;
    __device__ static acc_t mma(uint4 a, uint4 b, acc_t c) {
        return __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i32x4, a),
                                                     __builtin_bit_cast(i32x4, b), c, 0, 0, 0);
    }
    static constexpr int NM = G;  // MFMA steps per tile
    __device__ static acc_t mma_at(const uint4 (&a)[G], const uint4 (&b)[G], int m, acc_t c) {
        return mma(a[m], b[m], c);
    }
};

// FP8 (OCP e4m3) with a power-of-two scale per row: groups 0-3 hold the
// data (two K = 64 steps of 2 groups each), group 4 the row's scale byte in
// its first byte (e8m0, 127 = 1.0), which the scaled MFMA applies.
template <>
struct KT<KIND_FP8> {
    using acc_t = f32x16;
    static constexpr int TG = 5;
    static constexpr int G = 5;
    static constexpr int NM = 2;
    __device__ static acc_t mma_at(const uint4 (&a)[G], const uint4 (&b)[G], int m, acc_t c) {
        const uint4 a0 = a[2 * m], a1 = a[2 * m + 1], b0 = b[2 * m], b1 = b[2 * m + 1];
        const i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
        const i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w, (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 0, 0, 0, (int)a[4].x, 0, (int)b[4].x);
    }
};

// Rank keys are float for the float kinds and int32 for KIND_I8; both are
// carried as 32-bit patterns.  "none" = +inf / INT32_MAX with index UINT32_MAX.
// ord(): an int with the same order as the key (atomicMin-able); float keys
// flip the magnitude bits of negatives.
template <int KIND>
struct Key {
    using T = float;
    __device__ static float none() { return __builtin_huge_valf(); }
    __device__ static float lowest() { return -__builtin_huge_valf(); }  // admits nothing
    __device__ static int ord(float f) {
        const int i = __float_as_int(f);
        return i >= 0 ? i : i ^ 0x7FFFFFFF;
    }
    __device__ static float from_ord(int o) { return __int_as_float(o >= 0 ? o : o ^ 0x7FFFFFFF); }
};
template <>
struct Key<KIND_I8> {
    using T = int;
    __device__ static int none() { return 0x7FFFFFFF; }
    __device__ static int lowest() { return (int)0x80000000; }
    __device__ static int ord(int k) { return k; }
    __device__ static int from_ord(int o) { return o; }
};

// Smallest key strictly greater than x (+inf / INT32_MAX stay).
__device__ inline float next_up(float x) {
    if (x == __builtin_huge_valf()) return x;
    if (x == 0.f) return __uint_as_float(1u);
    const uint32_t u = __float_as_uint(x);
    return __uint_as_float(x > 0.f ? u + 1u : u - 1u);
}
__device__ inline int next_up(int x) { return x == 0x7FFFFFFF ? x : x + 1; }
// Threshold of the fast test: float kinds test the key itself; I8 tests
// floor(key / 2) (its raw accumulator) against ceil(thr / 2).
__device__ inline float half_up(float x) { return x; }
__device__ inline int half_up(int x) { return (x >> 1) + (x & 1); }

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

// One unit of the persistent pairs kernel (see fused.hip SlabTask).
struct PairTask {
    int32_t t_begin, t_end, qblk0;
    uint32_t r_lim, row_off, q_row0, q_count, out_off;
    int32_t slot;  // FP8 coarse pass: which quarter of the base image (candidate slot)
};

// Chunks per region of a launch with guided slabs (at most).
constexpr int MAX_CHUNKS = 32;

struct FusedArgs {
    const uint4* q_tiles;   // packed queries, [qblocks][G][64]
    const uint4* b_tiles;   // packed base, [n_tiles][G][64]
    const float* b_nl;      // base norms in accumulator-lane order, [n_tiles][32] (I8: int32;
                            // the brute force passes c >> 1, the pairs kernel c)
    const int* b_nlc;       // I8 brute force: c in accumulator-lane order (exact keys in the slow path)
    uint4* part;            // [n_slabs][q_stride] {key1, idx1, key2, idx2}
    int* bound;             // per-query inclusive upper bound on the 2nd-best key
                            // (Key::ord form), shared by all blocks; or null
    uint32_t q_n, q_stride;
    uint32_t r_n;           // real base rows; rows >= r_n in the last tile are masked
    // brute-force kernel: slab s covers tiles [tile0 + s*tiles_per_slab, ...)
    // clipped to end_tile (absolute tile numbers, so row indices are global)
    int tile0, end_tile, tiles_per_slab, n_slabs, n_groups;
    // Guided slabs (n_chunks > 0, instead of tiles_per_slab): the range
    // [tile0, end_tile) is cut into 8 regions of `region` tiles, one per XCD,
    // and region x into n_chunks chunks, tiles [chunk_start[c],
    // chunk_start[c + 1]) of it, that shrink towards its end.  Block b runs
    // on XCD x = b % 8 and covers chunk c = (b / 8) / n_groups of region x for
    // group (b / 8) % n_groups, writing slab x * n_chunks + c.
    int region, n_chunks;
    int chunk_start[MAX_CHUNKS + 1];
    // pairs kernel only
    const PairTask* tasks;
    const int32_t* task_off;  // block b runs tasks [task_off[b], task_off[b + 1])
    int n_tasks;
    const float* q_norm;  // packed query norms (int32 bits for KIND_I8)
    uint32_t* out_idx;
    float* out_dist;
    uint32_t* out_matches;
    float ratio;
    int comp_mul;  // pairs kernel, I8: 256, passed at run time so a composite takes one v_mad_i32_i24
    uint32_t* out_cand;  // FP8 coarse pass: 16 candidate rows per query
    // 8-wave blocks: polls allowed per stage hand-off wait, and the status word
    // a wave ORs 1 into when a wait runs out of them (its results are then void)
    int wait_polls;
    int* status;
    // brute force, float kinds: device flag, nonzero when some input value is
    // not an integer in [0,255]; the keys are then coarse and the shared bound
    // is not lowered inside the launch (null: exact keys)
    const int* coarse;
};


// Launch shapes: (QW query blocks of 32 per wave, WAVES per block, TPS base
// tiles per LDS stage).  Queries are padded to a multiple of QPAD so every
// shape's group count fits the packed query buffer.
constexpr uint32_t QPAD = 1024;
using FusedFn = void (*)(const FusedArgs&, unsigned, hipStream_t);
using OccFn = int (*)();  // resident blocks per CU of the instantiation
struct FusedShape {
    int kind, qw, waves, tps;
    FusedFn fn;
    OccFn blocks_per_cu;
    FusedFn pairs_fn;
    OccFn pairs_blocks_per_cu;
};
// Every listed shape of a kind (for the pairs autotuner).
int shapes_of_kind(int kind, const FusedShape** out, int max);
int device_cu_count();
// SIFT_SHAPE="qw,waves,tps" picks a listed shape of the kind (for tuning);
// otherwise the kind's default (first listed).
const FusedShape* pick_shape(int kind);

}  // namespace sift
