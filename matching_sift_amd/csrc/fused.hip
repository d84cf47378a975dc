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

// Global -> LDS copies that bypass VGPRs (gfx950 global_load_lds_*): lane l
// writes its 16 (or 4) bytes to lds + 16*l (4*l).  Written as inline asm so
// the compiler does not pair them with a conservative vmcnt(0) before every
// later LDS read; the kernel waits for them explicitly (wait_vm_lgkm0) before
// the barrier that hands a stage to the other waves.  M0 carries the LDS
// address and is restored.  Device pass only: the host pass never runs these.
__device__ inline void dma16(const void* g, const void* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t la = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)lds);
    uint32_t save;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %1\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dwordx4 %2, off\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(save)
        : "s"(la), "v"(g)
        : "memory");
#else
    (void)g;
    (void)lds;
#endif
}
__device__ inline void dma4(const void* g, const void* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t la = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)lds);
    uint32_t save;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %1\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dword %2, off\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(save)
        : "s"(la), "v"(g)
        : "memory");
#else
    (void)g;
    (void)lds;
#endif
}
// The same copies with a uniform base address in SGPRs plus a per-lane 32-bit
// offset (the saddr form): no 64-bit address arithmetic per lane.
__device__ inline void dma16_s(const void* sbase, uint32_t voff, uint32_t lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t save;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %1\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dwordx4 %2, %3\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(save)
        : "s"(lds), "v"(voff), "s"(sbase)
        : "memory");
#else
    (void)sbase;
    (void)voff;
    (void)lds;
#endif
}
__device__ inline void dma4_s(const void* sbase, uint32_t voff, uint32_t lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t save;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %1\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dword %2, %3\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(save)
        : "s"(lds), "v"(voff), "s"(sbase)
        : "memory");
#else
    (void)sbase;
    (void)voff;
    (void)lds;
#endif
}
// s_waitcnt vmcnt(N) lgkmcnt(0) (gfx9 encoding: vmcnt[3:0] and [15:14],
// expcnt[6:4] left at its maximum, lgkmcnt[11:8]).
template <int N>
__device__ inline void wait_vm_lgkm0() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4));
}
// Work-group barrier without the implicit memory fence: every wait it needs
// is explicit above it.
__device__ inline void barrier_only() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Stage hand-off counters in LDS (see fused_slab): wait until *c reaches
// target, sleeping between polls.  The poll count is capped (`polls`, about 64
// cycles each), so a wave can never wait forever; a wave that gives up ORs 1
// into *status, and the host reports the launch as failed.
__device__ inline void wait_count(const uint32_t* c, uint32_t target, int polls, int* status) {
    bool done = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= target;
    for (int it = 0; !done && it < polls; it++) {
        __builtin_amdgcn_s_sleep(1);
        done = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= target;
    }
    if (!done && (threadIdx.x & 63) == 0) atomicOr(status, 1);
    asm volatile("" ::: "memory");
}
// One increment per wave (lane 0), after everything before it in program order.
__device__ inline void signal_count(uint32_t* c) {
    asm volatile("" ::: "memory");
    if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// s_waitcnt lgkmcnt(0) alone (vmcnt and expcnt left at their maxima).
__device__ inline void wait_lgkm0() {
    __builtin_amdgcn_s_waitcnt(0xF | (7 << 4) | (0 << 8) | (3 << 14));
    asm volatile("" ::: "memory");
}

// One (query group x base tile range) unit of work: the body of both the
// grid-mapped kernel (one unit per block) and the persistent pairs kernel
// (units from a task list).  Query index space: packed query rows; a query is
// valid when (row - q_row0) < q_count.  Base rows >= r_lim are padding.
struct SlabTask {
    int t_begin, t_end;  // base tiles [t_begin, t_end)
    int qblk0;           // first query block of the work-group
    uint32_t r_lim;      // base row limit (absolute)
    uint32_t row_off;    // absolute row of base row 0 (subtracted from output indices)
    uint32_t q_row0;     // packed query row of query 0
    uint32_t q_count;    // valid queries
    uint32_t out_off;    // PAIRS: output slot of query 0
    int slab;            // !PAIRS: partial-result slab
};

template <typename T>
__device__ inline T min3k(T a, T b, T c) { return min(min(a, b), c); }

// v_med3_i32 / v_min3_i32 written out: min/max trees over shared operands do
// not reliably combine into them.  Plain VALU ops (not volatile), so the
// compiler still schedules, combines and removes them.
__device__ inline int med3_i32(int a, int b, int c) {
#if defined(__HIP_DEVICE_COMPILE__)
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return max(min(a, b), min(max(a, b), c));
#endif
}
// (m & a) | (~m & b) in one v_bfi_b32 (the compiler turns the expression back
// into a compare and a select when m is a sign-extended bit).
__device__ inline uint32_t bfi_b32(uint32_t m, uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
#else
    return (m & a) | (~m & b);
#endif
}
__device__ inline int min3_i32(int a, int b, int c) {
#if defined(__HIP_DEVICE_COMPILE__)
    int r;
    asm("v_min3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return min(min(a, b), c);
#endif
}

template <int KIND, int QW, int WAVES, int TPS, bool PAIRS>
__device__ __forceinline__ void fused_slab(const FusedArgs& a, const SlabTask& task) {
    using KTk = KT<KIND>;
    using acc_t = typename KTk::acc_t;
    using key_t = typename Key<KIND>::T;
    constexpr int G = KTk::G;    // groups multiplied per tile (K = 128)
    constexpr int TG = KTk::TG;  // groups stored per packed tile (float kinds: + norm group)
    constexpr bool FLOAT_KEY = KIND != KIND_I8;
    constexpr int UPT = G * 64;  // 16-byte units per tile in LDS

    // Three LDS stages: stage k is consumed while k+1 and k+2 are in flight.
    // A stage is CH 1-KB chunks, padded to CHP so every wave issues the same
    // number of copies (NPW) per full stage, plus NLC 256-byte copies of the
    // stage's row norms (accumulator-lane order) into the wave's own slot.
    constexpr int CH = TPS * G;
    constexpr int NPW = (CH + WAVES - 1) / WAVES;
    constexpr int CHP = NPW * WAVES;
    constexpr int NLC = (TPS * 32 + 63) / 64;
    // I8 brute force, 4-wave blocks: the rows' c (exact keys on the slow
    // path) comes into LDS with the stage as well, instead of a global read
    // per flagged tile (8-wave blocks free a slot before their slow path, so
    // they keep the global read)
    constexpr bool NC = KIND == KIND_I8 && !PAIRS && WAVES < 8;
    constexpr int NPW_T = NPW + NLC * (NC ? 2 : 1);
    __shared__ __align__(16) uint4 s_tiles[3][CHP * 64];
    __shared__ __align__(16) float s_nl[3][WAVES][NLC * 64];
    __shared__ __align__(16) int s_nc[NC ? 3 : 1][NC ? WAVES : 1][NC ? NLC * 64 : 4];

    const int t_begin = task.t_begin, t_end = task.t_end;
    const int tid = threadIdx.x;
    const int lane = tid & 63, h = lane >> 5;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar address arithmetic
    const int qblk0 = task.qblk0 + wave * QW;  // this wave's first query block
    // A lane's new 2nd best lowers the shared bound, except for coarse keys
    // (FP8, or float kinds on data that is not all integers in [0,255]): an
    // exact re-rank decides there, so only the widened seed bounds apply.
    const bool tighten = KIND != KIND_FP8 && a.bound != nullptr && !(a.coarse != nullptr && *a.coarse != 0);

    uint4 qop[QW][G];
#pragma unroll
    for (int qb = 0; qb < QW; qb++)
#pragma unroll
        for (int g = 0; g < G; g++) qop[qb][g] = a.q_tiles[((size_t)(qblk0 + qb) * TG + g) * 64 + lane];

    // Per-lane top-2 plus the query's shared bound B: an inclusive upper bound
    // on its 2nd-best key over the whole base (first from a base sample, then
    // lowered with atomicMin by any block that finds a better 2nd best; a
    // block reads B when it starts).  A row with key > B cannot be among the
    // two nearest, so the admission test is key < thr = min(m2, next_up(B)).
    // Both slots start at (next_up(B), none), so every true top-2 row
    // qualifies.
    key_t m1[QW], m2[QW], thr[QW];
    // I8 tests the raw accumulator q''.b' + (c >> 1), half the key rounded
    // down (key = 2 q''.b' + c): any key below thr has accumulator below
    // ceil(thr / 2), so the test on thr_h = ceil(thr / 2) misses no row
    key_t thr_h[QW];
    uint32_t i1[QW], i2[QW];
    // Packed mode (pairs kernel, integer kinds): each lane keeps its top-2 as
    // int32 "composites" key*F + slot so one v_min + one v_med3 per row
    // maintain it with no branch (see the insert below).  With two tiles per
    // stage the slot also carries the tile (RB = 1), and new entries are
    // recorded once per stage instead of once per tile.
    constexpr bool COMP = PAIRS && KIND != KIND_F32;
    constexpr int RB = TPS == 2 ? 1 : 0;     // tile bits in a slot
    constexpr int SB = 6 + RB;               // slot bits: F = 2^SB
    constexpr int NEWB = 1 << (SB - 1);      // set in the slot of an entry not yet recorded
    constexpr int CNONE = 0x7FFFFFFF & ~((1 << SB) - 1);  // above every real composite, slot bits 0
    static_assert(!(PAIRS && KIND == KIND_I8) || TPS == 2, "I8 pair composites carry one tile bit");
    int c1[QW], c2[QW];
    const uint32_t lq = (uint32_t)(lane & 31);
#pragma unroll
    for (int qb = 0; qb < QW; qb++) {
        const uint32_t q = (uint32_t)(qblk0 + qb) * 32u + lq;
        const bool valid = q - task.q_row0 < task.q_count;
        key_t init = Key<KIND>::none();
        if (a.bound && valid) init = next_up(Key<KIND>::from_ord(a.bound[q]));
        // padding queries admit nothing, so they never flag a tile
        m1[qb] = m2[qb] = init;
        thr[qb] = valid ? init : Key<KIND>::lowest();
        thr_h[qb] = half_up(thr[qb]);
        i1[qb] = i2[qb] = 0xFFFFFFFFu;
        c1[qb] = c2[qb] = CNONE;
    }

    auto dma_stage = [&](int ts, int bsel) -> bool {  // true: full stage (NPW_T copies issued)
        const int nt = min(TPS, t_end - ts);
        const bool full = nt == TPS;
        // chunk c = tile c / G, group c % G (the packed norm group is skipped)
        const uint4* src = a.b_tiles + (size_t)ts * TG * 64;
        auto gsrc = [&](int c) { return src + (size_t)((c / G) * TG + c % G) * 64; };  // chunk base (uniform)
        auto lds_of = [&](const void* p) { return __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)p); };
        if (full) {  // the common case: every copy, no per-copy branch
#pragma unroll
            for (int i = 0; i < NPW; i++) {
                const int c = wave + i * WAVES;  // padding slots reload chunk 0 (never read back)
                dma16_s(gsrc(c < CH ? c : 0), (uint32_t)lane * 16u, lds_of(&s_tiles[bsel][c * 64]));
            }
            if constexpr (NLC * 64 <= TPS * 32) {  // every norm copy in range
#pragma unroll
                for (int c = 0; c < NLC; c++)
                    dma4_s(a.b_nl + (size_t)ts * 32 + c * 64, (uint32_t)lane * 4u, lds_of(&s_nl[bsel][wave][c * 64]));
                if constexpr (NC) {
#pragma unroll
                    for (int c = 0; c < NLC; c++)
                        dma4_s(a.b_nlc + (size_t)ts * 32 + c * 64, (uint32_t)lane * 4u, lds_of(&s_nc[bsel][wave][c * 64]));
                }
                return true;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NPW; i++) {
                const int c = wave + i * WAVES;
                if (c < nt * G) dma16_s(gsrc(c), (uint32_t)lane * 16u, lds_of(&s_tiles[bsel][c * 64]));
            }
        }
#pragma unroll
        for (int c = 0; c < NLC; c++)
            dma4(a.b_nl + (size_t)ts * 32 + min(c * 64 + lane, nt * 32 - 1), &s_nl[bsel][wave][c * 64]);
        if constexpr (NC) {
#pragma unroll
            for (int c = 0; c < NLC; c++)
                dma4(a.b_nlc + (size_t)ts * 32 + min(c * 64 + lane, nt * 32 - 1), &s_nc[bsel][wave][c * 64]);
        }
        return full;
    };

    // Stage hand-off without block barriers.  Stage k lives in LDS slot
    // k % 3.  s_ready[slot] counts the waves whose copies of that slot's
    // current stage have landed; s_free[slot] the waves that have read it
    // into registers.  A wave waits only for the stage it is about to read
    // (every wave's copies landed) and for the slot it is about to refill
    // (every wave has read the stage that used it), so a wave on the slow
    // path holds the others back by at most one stage instead of stopping
    // them at every stage.  Counters only grow within a unit of work.
    // Blocks of fewer than 8 waves keep one barrier per stage instead: their
    // waves stay in step anyway, and the barrier is cheaper than the polls.
    constexpr bool HANDOFF = WAVES >= 8;
    __shared__ uint32_t s_ready[3], s_free[3];
    if constexpr (HANDOFF) {
        if (threadIdx.x < 3) {
            s_ready[threadIdx.x] = 0;
            s_free[threadIdx.x] = 0;
        }
        barrier_only();
    }

    const int nst = (t_end - t_begin + TPS - 1) / TPS;  // stages of this slab
    if (nst > 0) dma_stage(t_begin, 0);
    // The query fragments are complete from here on, as far as the compiler
    // can tell: it does not see the stage copies (inline asm), so without
    // this it waits for the fragment loads at their first use inside the
    // stage loop, with vmcnt values that the copies in flight then make wait
    // for the copies issued two stages ahead.  The wait it places here
    // covers the fragments and stage 0's copies, which the first stage needs
    // anyway; stage 1's copies are issued after it.
#pragma unroll
    for (int qb = 0; qb < QW; qb++)
#pragma unroll
        for (int g = 0; g < G; g++)
            asm volatile("" : "+v"(qop[qb][g].x), "+v"(qop[qb][g].y), "+v"(qop[qb][g].z), "+v"(qop[qb][g].w));
    bool next_full = nst > 1 ? dma_stage(t_begin + TPS, 1) : false;
    if (nst > 0) {
        if (next_full) wait_vm_lgkm0<NPW_T>(); else wait_vm_lgkm0<0>();
        if constexpr (HANDOFF) signal_count(&s_ready[0]); else barrier_only();
    }
    for (int k = 0; k < nst; k++) {
        const int ts = t_begin + k * TPS;
        const int ntile = min(TPS, t_end - ts);
        const int buf = k % 3;
        if constexpr (HANDOFF) wait_count(&s_ready[buf], (uint32_t)(WAVES * (k / 3 + 1)), a.wait_polls, a.status);
        // stage k+2 goes into the slot of stage k-1 once every wave has read
        // it (hand-off: its counter; otherwise the barrier that closed stage
        // k-1, so 4-wave blocks issue the copies before this stage's LDS reads)
        const int b2 = (k + 2) % 3;
        const bool issued = k + 2 < nst;
        bool issued_full = false;
        if constexpr (!HANDOFF) issued_full = issued ? dma_stage(ts + 2 * TPS, b2) : false;
        uint4 afr[TPS][G];
        int nlr[TPS][16];
        // (tiles past the slab end read stale LDS: in bounds, never flagged)
#pragma unroll
        for (int tt = 0; tt < TPS; tt++) {
#pragma unroll
            for (int g = 0; g < G; g++) afr[tt][g] = s_tiles[buf][tt * UPT + g * 64 + lane];
            {
                const float4* nlp = reinterpret_cast<const float4*>(&s_nl[buf][wave][tt * 32 + h * 16]);
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const float4 v = nlp[c];
                    nlr[tt][4 * c] = __float_as_int(v.x); nlr[tt][4 * c + 1] = __float_as_int(v.y);
                    nlr[tt][4 * c + 2] = __float_as_int(v.z); nlr[tt][4 * c + 3] = __float_as_int(v.w);
                }
            }
        }
        if constexpr (HANDOFF) {
            wait_lgkm0();  // this stage's fragments are in registers
            signal_count(&s_free[buf]);
        }
        if constexpr (HANDOFF) {
            if (issued && k >= 1)
                wait_count(&s_free[b2], (uint32_t)(WAVES * ((k - 1) / 3 + 1)), a.wait_polls, a.status);
            issued_full = issued ? dma_stage(ts + 2 * TPS, b2) : false;
        }
        if constexpr (HANDOFF) {
            if (k + 1 < nst) {  // this wave's copies of stage k+1 have landed
                if (issued_full) wait_vm_lgkm0<NPW_T>(); else wait_vm_lgkm0<0>();
                signal_count(&s_ready[(k + 1) % 3]);
            }
        }

        // Float kinds: the accumulator starts at the rows' |b|^2 (accumulator
        // lane order) and the queries are packed as -2q, so the MFMA chain
        // ends at t = |b|^2 - 2 q.b (every partial sum is an integer of
        // magnitude below 2^24 for SIFT data, hence exact).  I8 (brute force)
        // starts at c >> 1 and ends at q''.b' + (c >> 1) = floor(t / 2) with
        // t = |b'|^2 - 2 q'.b' = 2 q''.b' + c; the pairs kernel starts at 0.
        auto tile_acc = [&](int tt, acc_t (&acc)[QW]) {
            acc_t init{};
            if constexpr (FLOAT_KEY) {
#pragma unroll
                for (int j = 0; j < 16; j++) init[j] = __int_as_float(nlr[tt][j]);
            } else if constexpr (!PAIRS) {
#pragma unroll
                for (int j = 0; j < 16; j++) init[j] = nlr[tt][j];
            }
#pragma unroll
            for (int qb = 0; qb < QW; qb++) acc[qb] = KTk::mma_at(afr[tt], qop[qb], 0, init);
#pragma unroll
            for (int m = 1; m < KTk::NM; m++) {
#pragma unroll
                for (int qb = 0; qb < QW; qb++) acc[qb] = KTk::mma_at(afr[tt], qop[qb], m, acc[qb]);
            }
        };
        auto tile_keys = [&](int tt, key_t (&tv)[QW][16]) {
            acc_t acc[QW];
            tile_acc(tt, acc);
#pragma unroll
            for (int qb = 0; qb < QW; qb++)
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    tv[qb][j] = acc[qb][j];  // I8: floor(key / 2), tested against thr_h
                }
        };

        if constexpr (COMP) {
            // Every row is inserted, branch-free.  A row's composite is
            // key*F + NEWB + 32*tt + r (r = the row within its 32-row tile, tt
            // = the tile within the stage when RB = 1); entries kept from
            // earlier stages are stored as key*F (slot bits cleared).  Order
            // by composite = order by (key, row): a new row loses a tie
            // against a kept one (earlier stage, lower index) and wins or
            // loses within its stage by 32*tt + r.  Keys are exact integers
            // with |key| < 2^24 for SIFT data (base pad rows: 65504 * 256), so
            // key*128 + 127 fits in int32.  I8: the packed base norms arrive as
            // (c << SB) | (NEWB + 32 * parity + r) (compose_slots; base images
            // start on even tiles and stages on even tiles of their image, so
            // the parity is tt), and a composite is one v_mad_i32_i24:
            // 2F q''.b' + that.
#pragma unroll
            for (int tt = 0; tt < TPS; tt++) {
                if (tt >= ntile) break;  // uniform: slab end
                acc_t acc[QW];
                tile_acc(tt, acc);
                // base pad rows carry PAD_KEY, so they need no mask here: they
                // lose to every real row and are dropped after the loop
#pragma unroll
                for (int qb = 0; qb < QW; qb++) {
#pragma unroll
                    for (int j = 0; j < 16; j += 2) {
                        int cv[2];
#pragma unroll
                        for (int e = 0; e < 2; e++) {
                            const int jj = j + e;
                            if constexpr (FLOAT_KEY)
                                cv[e] = ((int)acc[qb][jj] << SB) |
                                        (NEWB + 32 * (RB ? tt : 0) + 8 * (jj >> 2) + 4 * h + (jj & 3));
                            else
                                cv[e] = __mul24((int)acc[qb][jj], a.comp_mul) + nlr[tt][jj];
                        }
                        // two rows at once (c1 <= c2): the 2nd smallest of
                        // {c1, x, y} is their median, and c2 only enters if it is
                        // below that; the smallest is min3.  3 VALU per two rows.
                        c2[qb] = min(c2[qb], med3_i32(c1[qb], cv[0], cv[1]));
                        c1[qb] = min3_i32(c1[qb], cv[0], cv[1]);
                    }
                }
                // after the stage's last tile (or the slab's): new entries
                // carry slot bits; record their rows, then clear the bits
                if (RB == 0 || tt == TPS - 1 || tt + 1 >= ntile) {
                    const uint32_t row0 = (uint32_t)(ts + (RB ? 0 : tt)) * 32u;
#pragma unroll
                    for (int qb = 0; qb < QW; qb++) {
                        // all ones where the entry is new (v_bfe_i32), and the
                        // selects as bit-field inserts (v_bfi_b32)
                        const uint32_t n1 = (uint32_t)((c1[qb] << (32 - SB)) >> 31);
                        const uint32_t n2 = (uint32_t)((c2[qb] << (32 - SB)) >> 31);
                        const uint32_t s1 = (uint32_t)(c1[qb] & (NEWB - 1)), s2 = (uint32_t)(c2[qb] & (NEWB - 1));
                        // stage rows start on multiples of NEWB except in FP8's
                        // image quarters: there the slot is added, else OR-ed in
                        const uint32_t r1 = KIND == KIND_FP8 ? row0 + s1 : (row0 | s1);
                        const uint32_t r2 = KIND == KIND_FP8 ? row0 + s2 : (row0 | s2);
                        // a kept 2nd entry is the old 1st if the 1st was displaced
                        i2[qb] = bfi_b32(n2, r2, bfi_b32(n1, i1[qb], i2[qb]));
                        i1[qb] = bfi_b32(n1, r1, i1[qb]);
                        c1[qb] &= ~((1 << SB) - 1);
                        c2[qb] &= ~((1 << SB) - 1);
                    }
                }
            }
        } else {
        // Common path, branch-free over the whole stage so the compiler can
        // overlap one tile's key tests with the next tile's MFMAs: per tile
        // and query block a wave-wide flag "some row may enter some lane's
        // top-2".  The last tile (pad rows) always takes the slow path, which
        // masks them.
        uint32_t fmask[TPS];
        bool any_flag = false;
        // Float kinds keep every key of the stage for the slow path (their
        // accumulators are live together anyway, and the fragments die after
        // the MFMAs); I8 recomputes instead (4 MFMAs per tile and query
        // block), since holding its keys would cost a wave per SIMD.
        constexpr bool KEEP = KIND != KIND_I8;
        key_t tv[KEEP ? TPS : 1][QW][16];
#pragma unroll
        for (int tt = 0; tt < TPS; tt++) {
            auto& tk = tv[KEEP ? tt : 0];
            tile_keys(tt, tk);
            const bool last = (uint32_t)(ts + tt + 1) * 32u > task.r_lim;
            fmask[tt] = 0;
#pragma unroll
            for (int qb = 0; qb < QW; qb++) {
                // minimum of the 16 keys in 8 min/min3 ops, one compare, and
                // the wave-wide test on the scalar side
                const key_t* v = tk[qb];
                const key_t a0 = min3k(v[0], v[1], v[2]), a1 = min3k(v[3], v[4], v[5]);
                const key_t a2 = min3k(v[6], v[7], v[8]), a3 = min3k(v[9], v[10], v[11]);
                const key_t a4 = min3k(v[12], v[13], v[14]);
                const key_t mn = min(min3k(a0, a1, a2), min3k(a3, a4, v[15]));
                const bool hit = __builtin_amdgcn_ballot_w64(mn < (FLOAT_KEY ? thr[qb] : thr_h[qb])) != 0;
                fmask[tt] |= ((hit || last) ? 1u : 0u) << qb;
            }
            if (tt >= ntile) fmask[tt] = 0;
            any_flag |= fmask[tt] != 0;
        }
        // Slow path (rare once the bound is tight): insert the flagged query
        // blocks of the flagged tiles (kept keys, or recomputed from the
        // resident fragments), 4 rows at a time.
        if (any_flag) {
#pragma unroll
            for (int tt = 0; tt < TPS; tt++) {
#pragma unroll
                for (int qb = 0; qb < QW; qb++) {
                    if (!((fmask[tt] >> qb) & 1u)) continue;
                    key_t t[16];
                    if constexpr (KEEP) {
#pragma unroll
                        for (int j = 0; j < 16; j++) t[j] = tv[tt][qb][j];
                    } else {
                        // the exact keys: 2 q''.b' + c (the MFMAs again from 0)
                        acc_t acc;
                        acc_t init{};
                        if constexpr (FLOAT_KEY) {
#pragma unroll
                            for (int j = 0; j < 16; j++) init[j] = __int_as_float(nlr[tt][j]);
                        }
                        acc = KTk::mma_at(afr[tt], qop[qb], 0, init);
#pragma unroll
                        for (int m = 1; m < KTk::NM; m++) acc = KTk::mma_at(afr[tt], qop[qb], m, acc);
                        if constexpr (FLOAT_KEY) {
#pragma unroll
                            for (int j = 0; j < 16; j++) t[j] = acc[j];
                        } else {
                            const int4* cl = NC ? reinterpret_cast<const int4*>(&s_nc[buf][wave][tt * 32 + h * 16])
                                                : reinterpret_cast<const int4*>(a.b_nlc + (size_t)(ts + tt) * 32 + h * 16);
#pragma unroll
                            for (int c4 = 0; c4 < 4; c4++) {
                                const int4 cv = cl[c4];
                                t[4 * c4] = 2 * (int)acc[4 * c4] + cv.x;
                                t[4 * c4 + 1] = 2 * (int)acc[4 * c4 + 1] + cv.y;
                                t[4 * c4 + 2] = 2 * (int)acc[4 * c4 + 2] + cv.z;
                                t[4 * c4 + 3] = 2 * (int)acc[4 * c4 + 3] + cv.w;
                            }
                        }
                    }
                    const uint32_t tile_row0 = (uint32_t)(ts + tt) * 32u;
                    const uint32_t rowbase = tile_row0 + 4u * h;
                    if (tile_row0 + 32u > task.r_lim) {  // last tile: pad rows never rank
#pragma unroll
                        for (int j = 0; j < 16; j++) {
                            const bool pad = rowbase + 8u * (j >> 2) + (j & 3) >= task.r_lim;
                            t[j] = pad ? Key<KIND>::none() : t[j];
                        }
                    }
                    const uint32_t old2 = i2[qb];
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const key_t gm = min(min(t[4 * c], t[4 * c + 1]), min(t[4 * c + 2], t[4 * c + 3]));
                        if (gm < thr[qb]) {
#pragma unroll
                            for (int e = 0; e < 4; e++) {
                                const key_t v = t[4 * c + e];
                                const uint32_t row = rowbase + 8u * c + e;
                                const bool c1 = v < m1[qb], c2 = v < m2[qb];
                                m2[qb] = c1 ? m1[qb] : (c2 ? v : m2[qb]);
                                i2[qb] = c1 ? i1[qb] : (c2 ? row : i2[qb]);
                                m1[qb] = c1 ? v : m1[qb];
                                i1[qb] = c1 ? row : i1[qb];
                            }
                            thr[qb] = min(thr[qb], m2[qb]);
                            thr_h[qb] = half_up(thr[qb]);
                        }
                    }
                    // a real 2nd best found here bounds the query everywhere (not
                    // for coarse keys: they may rank the true 2-NN above a coarse
                    // 2nd best, and the candidates must not depend on block timing)
                    if (tighten && i2[qb] != old2 && i2[qb] != 0xFFFFFFFFu) {
                        const uint32_t q = (uint32_t)(qblk0 + qb) * 32u + lq;
                        if (q - task.q_row0 < task.q_count) atomicMin(&a.bound[q], Key<KIND>::ord(m2[qb]));
                    }
                }
            }
        }
        }  // !COMP
        if constexpr (!HANDOFF) {
            if (k + 1 < nst) {  // stage k+1 complete for every wave; this stage's slot free
                if (issued_full) wait_vm_lgkm0<NPW_T>(); else wait_vm_lgkm0<0>();
                barrier_only();
            }
        }
    }

    if constexpr (COMP) {  // composites back to keys; pad rows (>= r_lim) are absent
#pragma unroll
        for (int qb = 0; qb < QW; qb++) {
            if (i1[qb] >= task.r_lim) i1[qb] = 0xFFFFFFFFu;
            if (i2[qb] >= task.r_lim) i2[qb] = 0xFFFFFFFFu;
            m1[qb] = i1[qb] != 0xFFFFFFFFu ? (key_t)(c1[qb] >> SB) : Key<KIND>::none();
            m2[qb] = i2[qb] != 0xFFFFFFFFu ? (key_t)(c2[qb] >> SB) : Key<KIND>::none();
        }
    }
    if constexpr (PAIRS && KIND == KIND_FP8) {
        // coarse pass: each lane half's top-2 is a candidate list of its own
        // (slot 4 * cand_slot + 2 * h); the exact re-rank merges them
#pragma unroll
        for (int qb = 0; qb < QW; qb++) {
            const uint32_t ql = (uint32_t)(qblk0 + qb) * 32u + lq - task.q_row0;
            if (ql < task.q_count) {
                const size_t o = ((size_t)task.out_off + ql) * 16 + 4 * (size_t)task.slab + 2 * h;
                a.out_cand[o] = i1[qb] != 0xFFFFFFFFu ? i1[qb] - task.row_off : 0xFFFFFFFFu;
                a.out_cand[o + 1] = i2[qb] != 0xFFFFFFFFu ? i2[qb] - task.row_off : 0xFFFFFFFFu;
            }
        }
        return;
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
        const uint32_t ql = q - task.q_row0;
        if (h == 0 && ql < task.q_count) {
            if constexpr (PAIRS) {
                // single slab per unit: final outputs (d = key + |q|^2, ratio test)
                const bool h1 = i1[qb] != 0xFFFFFFFFu, h2 = i2[qb] != 0xFFFFFFFFu;
                float d1 = 3.402823466e+38f, d2 = 3.402823466e+38f;
                if constexpr (KIND == KIND_I8) {
                    const int nq = __float_as_int(a.q_norm[q]);
                    if (h1) d1 = (float)(m1[qb] + nq);
                    if (h2) d2 = (float)(m2[qb] + nq);
                } else {
                    const float nq = a.q_norm[q];
                    if (h1) d1 = fmaxf(m1[qb] + nq, 0.f);
                    if (h2) d2 = fmaxf(m2[qb] + nq, 0.f);
                }
                const uint32_t o = task.out_off + ql;
                const uint32_t j1 = h1 ? i1[qb] - task.row_off : 0xFFFFFFFFu;
                const uint32_t j2 = h2 ? i2[qb] - task.row_off : 0xFFFFFFFFu;
                if (a.out_idx) {
                    a.out_idx[2 * (size_t)o] = j1;
                    a.out_idx[2 * (size_t)o + 1] = j2;
                }
                if (a.out_dist) {
                    a.out_dist[2 * (size_t)o] = d1;
                    a.out_dist[2 * (size_t)o + 1] = d2;
                }
                if (a.out_matches) {
                    // IEEE sqrt and division; d1 = d2 = 0 gives NaN -> no match
                    const float r = sqrtf(d1) / sqrtf(d2);
                    a.out_matches[o] = (h1 && r < a.ratio) ? j1 : 0xFFFFFFFFu;
                }
            } else {
                uint4 o;
                o.x = __builtin_bit_cast(uint32_t, m1[qb]);
                o.y = i1[qb];
                o.z = __builtin_bit_cast(uint32_t, m2[qb]);
                o.w = i2[qb];
                a.part[(size_t)task.slab * a.q_stride + q] = o;
            }
        }
    }
}

template <int KIND, int QW, int WAVES, int TPS>
__global__ __launch_bounds__(WAVES * 64) void fused_top2(FusedArgs a) {
    // block -> (group, slab), XCD-aware: block b runs on XCD b % 8, and all
    // groups of one slab are placed on the same XCD to share its L2 copy
    const int b = blockIdx.x;
    const int xcd = b & 7, r = b >> 3;
    const int k = r / a.n_groups;  // the XCD's k-th slab
    const int group = r % a.n_groups;
    SlabTask t;
    int slab;
    if (a.n_chunks > 0) {
        // guided: XCD x covers region x in chunks that shrink with k; the
        // dispatcher starts blocks in index order, so each XCD ends on short
        // blocks instead of a last round of full slabs
        if (k >= a.n_chunks) return;  // uniform per block
        const int rbase = a.tile0 + xcd * a.region;
        t.t_begin = min(rbase + a.chunk_start[k], a.end_tile);
        t.t_end = min(rbase + a.chunk_start[k + 1], a.end_tile);
        slab = xcd * a.n_chunks + k;
    } else {
        slab = k * 8 + xcd;
        if (slab >= a.n_slabs) return;  // uniform per block
        t.t_begin = a.tile0 + slab * a.tiles_per_slab;
        t.t_end = min(t.t_begin + a.tiles_per_slab, a.end_tile);
    }
    t.qblk0 = group * WAVES * QW;
    t.r_lim = a.r_n;
    t.row_off = 0;
    t.q_row0 = 0;
    t.q_count = a.q_n;
    t.out_off = 0;
    t.slab = slab;
    fused_slab<KIND, QW, WAVES, TPS, false>(a, t);
}

// Persistent kernel over a task list (image pairs): block b runs tasks
// [task_off[b], task_off[b + 1]), a run the host balanced across blocks;
// LDS is reused between tasks behind a barrier.
// The I8 (2, 4, 2) pairs kernel fits 128 VGPRs when asked (122, no spill):
// 4 waves per SIMD instead of 3 cover more of each tile's wait for its MFMA
// results before the inserts (2-3% on the 512-pair workload).
template <int KIND, int QW, int WAVES>
constexpr int pairs_waves_per_eu() { return (KIND == KIND_I8 && QW == 2 && WAVES == 4) ? 4 : 1; }
template <int KIND, int QW, int WAVES, int TPS>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(pairs_waves_per_eu<KIND, QW, WAVES>())))
void pairs_top2(FusedArgs a) {
    const int it_end = a.task_off[blockIdx.x + 1];
    for (int it = a.task_off[blockIdx.x]; it < it_end; it++) {
        const PairTask pt = a.tasks[it];
        SlabTask t;
        t.t_begin = pt.t_begin;
        t.t_end = pt.t_end;
        t.qblk0 = pt.qblk0;
        t.r_lim = pt.r_lim;
        t.row_off = pt.row_off;
        t.q_row0 = pt.q_row0;
        t.q_count = pt.q_count;
        t.out_off = pt.out_off;
        t.slab = pt.slot;
        fused_slab<KIND, QW, WAVES, TPS, true>(a, t);
        wait_vm_lgkm0<0>();
        barrier_only();  // the next task refills the LDS stages
    }
}

template <int KIND, int QW, int WAVES, int TPS>
void fused_launch(const FusedArgs& a, unsigned grid, hipStream_t s) {
    static_assert(QPAD % (QW * WAVES * 32) == 0, "group size must divide QPAD");
    fused_top2<KIND, QW, WAVES, TPS><<<grid, WAVES * 64, 0, s>>>(a);
}
template <int KIND, int QW, int WAVES, int TPS>
void pairs_launch(const FusedArgs& a, unsigned grid, hipStream_t s) {
    pairs_top2<KIND, QW, WAVES, TPS><<<grid, WAVES * 64, 0, s>>>(a);
}
template <int KIND, int QW, int WAVES, int TPS>
int pairs_occupancy() {
    static int n = -1;
    if (n < 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, pairs_top2<KIND, QW, WAVES, TPS>, WAVES * 64, 0) !=
                hipSuccess ||
            v < 1)
            v = 1;
        n = v;
    }
    return n;
}

template <int KIND, int QW, int WAVES, int TPS>
int fused_occupancy() {
    static int n = -1;
    if (n < 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, fused_top2<KIND, QW, WAVES, TPS>, WAVES * 64, 0) !=
                hipSuccess ||
            v < 1)
            v = 1;
        n = v;
    }
    return n;
}
#define SHAPE(K, Q, W, T) \
    {K, Q, W, T, fused_launch<K, Q, W, T>, fused_occupancy<K, Q, W, T>, pairs_launch<K, Q, W, T>, pairs_occupancy<K, Q, W, T>}

const FusedShape kShapes[] = {
    // default per kind comes first (the brute force times the first two of
    // BF16 and F16 on a large base and keeps the faster)
    SHAPE(KIND_BF16, 2, 4, 2),
    SHAPE(KIND_BF16, 2, 8, 2),
    SHAPE(KIND_BF16, 2, 4, 4),
    SHAPE(KIND_BF16, 4, 4, 2),
    SHAPE(KIND_BF16, 1, 4, 2),
    SHAPE(KIND_BF16, 1, 8, 2),
    SHAPE(KIND_F16, 2, 4, 2),
    SHAPE(KIND_F16, 2, 8, 2),
    SHAPE(KIND_F32, 2, 4, 1),
    SHAPE(KIND_FP8, 4, 4, 2),
    SHAPE(KIND_FP8, 2, 4, 2),
    SHAPE(KIND_FP8, 2, 8, 2),
    SHAPE(KIND_I8, 4, 4, 2),
    SHAPE(KIND_I8, 2, 4, 2),
    SHAPE(KIND_I8, 2, 8, 2),
};

}  // namespace

int device_cu_count() {
    static int n = -1;
    if (n < 0) {
        int dev = 0, v = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v < 1) v = 256;
        n = v;
    }
    return n;
}

int shapes_of_kind(int kind, const FusedShape** out, int max) {
    int n = 0;
    for (const auto& s : kShapes)
        if (s.kind == kind && n < max) out[n++] = &s;
    return n;
}

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
