// fused.hip — the fused MFMA distance + top-2 kernel.
//
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
//
// Compiled with -fno-honor-nans: keys are never NaN here, which lets min()
// map to v_min/v_min3 without canonicalisation.
#include <cstdio>
#include <cstdlib>

#include "fused.h"

namespace sift {
namespace {

template <int KIND, int QW, int WAVES, int TPS>
__global__ __launch_bounds__(WAVES * 64) void fused_top2(FusedArgs a) {
    using KTk = KT<KIND>;
    using acc_t = typename KTk::acc_t;
    using key_t = typename Key<KIND>::T;
    constexpr int G = KTk::G;
    constexpr bool NORM_IN_K = KIND != KIND_I8;  // float kinds: |b|^2 comes out of the MFMA
    constexpr int UPT = G * 64;                   // 16-byte units per tile
    constexpr int NT = WAVES * 64;
    constexpr int STAGE = TPS * UPT;
    constexpr int PER_THREAD = (STAGE + NT - 1) / NT;
    static_assert(TPS * 32 <= NT, "one norm per thread per stage");

    // two LDS stages: stage s+1 is written while stage s is consumed
    __shared__ __align__(16) uint4 s_tiles[2][STAGE];
    __shared__ __align__(16) float s_nl[2][NORM_IN_K ? 1 : TPS * 32];

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

    // Per-lane top-2.  With a seeded bound T (strictly above the query's
    // 2nd-best key over a base sample, hence above its true 2nd-best), both
    // slots start at (T, none): only rows with key < T are ever inserted, and
    // every true top-2 row qualifies.
    key_t m1[QW], m2[QW];
    uint32_t i1[QW], i2[QW];
#pragma unroll
    for (int qb = 0; qb < QW; qb++) {
        const uint32_t q = (uint32_t)(qblk0 + qb) * 32u + (uint32_t)(lane & 31);
        key_t init = Key<KIND>::none();
        if (a.thresh && q < a.q_n) init = __builtin_bit_cast(key_t, a.thresh[q]);
        m1[qb] = m2[qb] = init;
        i1[qb] = i2[qb] = 0xFFFFFFFFu;
    }

    uint4 pre[PER_THREAD];
    float pre_nl = 0.f;
    auto load_stage = [&](int ts) {
        const int avail = min(TPS, t_end - ts) * UPT;
        const uint4* src = a.b_tiles + (size_t)ts * UPT;
#pragma unroll
        for (int u = 0; u < PER_THREAD; u++) {
            const int idx = u * NT + tid;
            pre[u] = (idx < avail) ? src[idx] : make_uint4(0u, 0u, 0u, 0u);
        }
        if constexpr (!NORM_IN_K)
            if (tid < TPS * 32) pre_nl = (tid < min(TPS, t_end - ts) * 32) ? a.b_nl[(size_t)ts * 32 + tid] : 0.f;
    };
    auto store_stage = [&](int buf) {
#pragma unroll
        for (int u = 0; u < PER_THREAD; u++) {
            const int idx = u * NT + tid;
            if (PER_THREAD * NT == STAGE || idx < STAGE) s_tiles[buf][idx] = pre[u];
        }
        if constexpr (!NORM_IN_K)
            if (tid < TPS * 32) s_nl[buf][tid] = pre_nl;
    };

    if (t_begin < t_end) {
        load_stage(t_begin);
        store_stage(0);
    }
    __syncthreads();
    int buf = 0;
    for (int ts = t_begin; ts < t_end; ts += TPS) {
        const bool more = ts + TPS < t_end;
        if (more) load_stage(ts + TPS);  // in flight while this stage is consumed
        const int ntile = min(TPS, t_end - ts);
#pragma unroll
        for (int tt = 0; tt < TPS; tt++) {
            if (tt >= ntile) break;
            const uint4* tp = &s_tiles[buf][tt * UPT];
            acc_t acc[QW];
#pragma unroll
            for (int qb = 0; qb < QW; qb++) acc[qb] = acc_t{};
#pragma unroll
            for (int g = 0; g < G; g++) {
                const uint4 av = tp[g * 64 + lane];
#pragma unroll
                for (int qb = 0; qb < QW; qb++) acc[qb] = KTk::mma(av, qop[qb][g], acc[qb]);
            }
            const uint32_t tile_row0 = (uint32_t)(ts + tt) * 32u;
            const uint32_t rowbase = tile_row0 + 4u * h;
            // keys: float kinds t = |b|^2 - 2 q.b straight from the MFMA (the
            // queries are packed as -2q, the norm group adds |b|^2; all partial
            // sums are integers below 2^24 for SIFT data, hence exact);
            // I8: t = |b'|^2 - 2 q'.b' with the norms from LDS
            key_t tv[QW][16];
            if constexpr (NORM_IN_K) {
#pragma unroll
                for (int qb = 0; qb < QW; qb++)
#pragma unroll
                    for (int j = 0; j < 16; j++) tv[qb][j] = acc[qb][j];
                if (tile_row0 + 32u > a.r_n) {  // last tile: pad rows never rank
#pragma unroll
                    for (int j = 0; j < 16; j++) {
                        const bool pad = rowbase + 8u * (j >> 2) + (j & 3) >= a.r_n;
#pragma unroll
                        for (int qb = 0; qb < QW; qb++) tv[qb][j] = pad ? Key<KIND>::none() : tv[qb][j];
                    }
                }
            } else {
                const float4* nlp = reinterpret_cast<const float4*>(&s_nl[buf][tt * 32 + h * 16]);
                int nl[16];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const float4 v = nlp[c];
                    nl[4 * c] = __float_as_int(v.x); nl[4 * c + 1] = __float_as_int(v.y);
                    nl[4 * c + 2] = __float_as_int(v.z); nl[4 * c + 3] = __float_as_int(v.w);
                }
#pragma unroll
                for (int qb = 0; qb < QW; qb++)
#pragma unroll
                    for (int j = 0; j < 16; j++) tv[qb][j] = nl[j] - 2 * acc[qb][j];
            }
            // admission: registers 4c..4c+3 hold 4 consecutive rows (8c + 4h + e);
            // one wave-wide branch per tile, per-group tests only inside it
            key_t gm[QW][4];
            bool any = false;
#pragma unroll
            for (int qb = 0; qb < QW; qb++)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    gm[qb][c] = min(min(tv[qb][4 * c], tv[qb][4 * c + 1]), min(tv[qb][4 * c + 2], tv[qb][4 * c + 3]));
                    any |= gm[qb][c] < m2[qb];
                }
            if (any) {  // rare once the bound is tight
#pragma unroll
                for (int qb = 0; qb < QW; qb++)
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        if (gm[qb][c] < m2[qb]) {
#pragma unroll
                            for (int e = 0; e < 4; e++) {
                                const key_t v = tv[qb][4 * c + e];
                                const uint32_t row = rowbase + 8u * c + e;
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
        if (more) store_stage(buf ^ 1);  // the other buffer was released by the last barrier
        __syncthreads();
        buf ^= 1;
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

template <int KIND, int QW, int WAVES, int TPS>
void fused_launch(const FusedArgs& a, unsigned grid, hipStream_t s) {
    static_assert(QPAD % (QW * WAVES * 32) == 0, "group size must divide QPAD");
    fused_top2<KIND, QW, WAVES, TPS><<<grid, WAVES * 64, 0, s>>>(a);
}

const FusedShape kShapes[] = {
    // default per kind comes first
    {KIND_BF16, 2, 4, 2, fused_launch<KIND_BF16, 2, 4, 2>},
    {KIND_BF16, 2, 4, 4, fused_launch<KIND_BF16, 2, 4, 4>},
    {KIND_BF16, 2, 8, 2, fused_launch<KIND_BF16, 2, 8, 2>},
    {KIND_BF16, 4, 4, 2, fused_launch<KIND_BF16, 4, 4, 2>},
    {KIND_BF16, 1, 8, 4, fused_launch<KIND_BF16, 1, 8, 4>},
    {KIND_F16, 2, 4, 2, fused_launch<KIND_F16, 2, 4, 2>},
    {KIND_F32, 2, 4, 1, fused_launch<KIND_F32, 2, 4, 1>},
    {KIND_I8, 4, 4, 4, fused_launch<KIND_I8, 4, 4, 4>},
    {KIND_I8, 2, 4, 4, fused_launch<KIND_I8, 2, 4, 4>},
    {KIND_I8, 4, 8, 4, fused_launch<KIND_I8, 4, 8, 4>},
};

}  // namespace

const FusedShape* pick_shape(int kind) {
    const FusedShape* first = nullptr;
    int qw = 0, waves = 0, tps = 0;
    const char* env = std::getenv("SIFT_SHAPE");
    if (env) std::sscanf(env, "%d,%d,%d", &qw, &waves, &tps);
    for (const auto& s : kShapes) {
        if (s.kind != kind) continue;
        if (!first) first = &s;
        if (s.qw == qw && s.waves == waves && s.tps == tps) return &s;
    }
    return first;
}

}  // namespace sift
