// pairs.hip — batched image-pair 2-NN in one persistent launch.
//
// The reference matches one image pair per cublas_2nn_sift call
// (src/test.cu drives include/knn_brute.h:22 once per pair); here a plan
// takes every pair of a batch at once:
//   * both sides are CSR descriptor sets; the pack kernel gathers them through
//     a row map so every image starts on a 32-row tile boundary (pad rows get
//     norm +inf and are masked by the image's row limit);
//   * work items are (pair, query group): one group of QW*WAVES*32 queries of
//     the pair's query image against all tiles of its base image, sorted by
//     decreasing cost and taken round-robin by a persistent grid;
//   * the kernel writes final outputs (d = key + |q|^2, indices relative to
//     the base image, ratio test), so there is no merge pass;
//   * with BF16/F16/I8 compute every row is inserted branch-free into a
//     packed int32 top-2 (small bases never let an admission bound settle,
//     so the bounded path of the 1M-row kernel would take its slow path on
//     nearly every tile); this needs integer keys, i.e. SIFT data in [0,255];
//     F32 compute keeps the float path and accepts any data;
//   * the execute sequence (pack queries, pack bases, pairs kernel) allocates
//     nothing and can be captured once in a HIP graph and replayed;
//   * an optional autotune times the kind's tile shapes on the first execute.
#include <algorithm>
#include <queue>
#include <cfloat>
#include <vector>

#include "fused.h"
#include "internal.h"

using namespace sift;

struct sift_pairs_plan {
    sift_context* h = nullptr;
    int type = 0, kind = 0;
    unsigned flags = 0;
    uint32_t n_qimg = 0, n_rimg = 0, n_pairs = 0;
    std::vector<uint32_t> q_off, r_off, pairs;
    std::vector<uint32_t> q_row0, r_row0;  // packed first row of each image
    std::vector<uint32_t> out_off;         // first output slot of each pair
    uint64_t total_q = 0;
    uint32_t q_rows = 0, r_rows = 0;  // input rows of each side
    PackedSet qs, bs;
    DevBuf qmap, rmap, tasks, flag;
    // FP8 coarse pass + exact re-rank: candidates, uint8 copies of both sides
    // with int32 norms, and per output query its input row and base offset
    DevBuf cand, q_u8, q_n2, r_u8, r_n2, qsrc, boff;
    const FusedShape* shape = nullptr;
    int n_tasks = 0;
    size_t off_bytes = 0;  // the per-block item offsets follow the items in `tasks`
    unsigned grid = 0;
    bool tune_pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // what the captured graph is bound to
    const void *g_q = nullptr, *g_r = nullptr;
    uint32_t* g_idx = nullptr;
    float* g_dist = nullptr;
    uint32_t* g_m = nullptr;
    float g_ratio = 0.f;
};

namespace {

inline uint32_t ceil32(uint32_t x) { return (x + 31u) & ~31u; }

// FP8 coarse pass: 4 base quarters x 2 lane halves x top-2 = 16 candidates.
constexpr int FP8_SLOTS = 4;

// Exact re-rank of the FP8 candidates: 16 lanes per query, one candidate each
// (uint8 rows, int32 |q - b|^2 = |q|^2 + |b|^2 - 2 q.b with v_dot4_u32_u8),
// then a butterfly over the 16 lanes keeps the two smallest (d, index) pairs,
// packed as d * 2^32 + index so one integer order is the lexicographic one.
// I8 base norms of the pairs kernel, in accumulator-lane order: c becomes
// (c << 6) | (32 + r), r the row within its 32-row tile, so the kernel forms
// each row's composite key*64 + 32 + r with a single multiply-add.
__global__ void compose_slots(int* __restrict__ nl, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = i & 31, hl = p >> 4, j = p & 15;
    const uint32_t r = 8 * (j >> 2) + 4 * hl + (j & 3);
    // slot = 64 (new) + 32 * (tile parity) + row in tile; images start on
    // even tiles, so the parity is the tile's position within a 2-tile stage
    const uint32_t par = (i >> 5) & 1u;
    nl[i] = (int)(((uint32_t)nl[i] << 7) | (64u + 32u * par + r));
}

__global__ __launch_bounds__(256) void fp8_rerank(const uint32_t* __restrict__ cand,
                                                  const uint4* __restrict__ q_u8, const int* __restrict__ q_n2,
                                                  const uint4* __restrict__ r_u8, const int* __restrict__ r_n2,
                                                  const uint32_t* __restrict__ qsrc,
                                                  const uint32_t* __restrict__ boff, uint32_t total,
                                                  uint32_t* __restrict__ idx, float* __restrict__ dist,
                                                  sift_float4* __restrict__ sorted,
                                                  uint32_t* __restrict__ matches, float ratio) {
    const uint64_t NONE = ~0ull;
    const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
    const uint32_t o = gid >> 4;
    const int s = gid & 15;
    if (o >= total) return;  // whole 16-lane groups exit together
    const uint32_t c = cand[(size_t)o * 16 + s];
    uint64_t m1 = NONE, m2 = NONE;
    if (c != 0xFFFFFFFFu) {
        const uint32_t qr = qsrc ? qsrc[o] : o;
        const size_t br = (size_t)(boff ? boff[o] : 0u) + c;
        const uint4* qp = q_u8 + (size_t)qr * 8;
        const uint4* bp = r_u8 + br * 8;
        uint32_t dot = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint4 a = qp[i], b = bp[i];
            dot = __builtin_amdgcn_udot4(a.x, b.x, dot, false);
            dot = __builtin_amdgcn_udot4(a.y, b.y, dot, false);
            dot = __builtin_amdgcn_udot4(a.z, b.z, dot, false);
            dot = __builtin_amdgcn_udot4(a.w, b.w, dot, false);
        }
        const uint32_t d = (uint32_t)(q_n2[qr] + r_n2[br]) - 2u * dot;
        m1 = ((uint64_t)d << 32) | c;
    }
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
        const uint64_t b1 = __shfl_xor(m1, off, 16), b2 = __shfl_xor(m2, off, 16);
        const uint64_t n2 = min(max(m1, b1), min(m2, b2));  // 2nd smallest of two sorted pairs
        m1 = min(m1, b1);
        m2 = n2;
    }
    if (s == 0) {
        const bool h1 = m1 != NONE, h2 = m2 != NONE;
        const float d1 = h1 ? (float)(uint32_t)(m1 >> 32) : 3.402823466e+38f;
        const float d2 = h2 ? (float)(uint32_t)(m2 >> 32) : 3.402823466e+38f;
        const uint32_t j1 = h1 ? (uint32_t)m1 : 0xFFFFFFFFu, j2 = h2 ? (uint32_t)m2 : 0xFFFFFFFFu;
        if (idx) {
            idx[2 * (size_t)o] = j1;
            idx[2 * (size_t)o + 1] = j2;
        }
        if (dist) {
            dist[2 * (size_t)o] = d1;
            dist[2 * (size_t)o + 1] = d2;
        }
        if (sorted) {
            sift_float4 f;
            f.x = d1;
            f.y = d2;
            f.z = h1 ? (float)j1 : 3.402823466e+38f;
            f.w = h2 ? (float)j2 : 3.402823466e+38f;
            sorted[o] = f;
        }
        if (matches) matches[o] = (h1 && sqrtf(d1) / sqrtf(d2) < ratio) ? j1 : 0xFFFFFFFFu;
    }
}

void drop_graph(sift_pairs_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

void free_plan(sift_pairs_plan* p) {
    drop_graph(p);
    for (DevBuf* b : {&p->qs.tiles, &p->qs.norm, &p->qs.norm_lane, &p->bs.tiles, &p->bs.norm,
                      &p->bs.norm_lane, &p->qmap, &p->rmap, &p->tasks, &p->flag, &p->cand, &p->q_u8,
                      &p->q_n2, &p->r_u8, &p->r_n2, &p->qsrc, &p->boff})
        release(*b);
    delete p;
}

// Work items for the plan's current shape, uploaded to p->tasks.
int build_tasks(sift_pairs_plan* p) {
    sift_context* h = p->h;
    const FusedShape& sh = *p->shape;
    const uint32_t qpg = (uint32_t)(sh.qw * sh.waves * 32);
    std::vector<PairTask> t;
    for (uint32_t i = 0; i < p->n_pairs; i++) {
        const uint32_t qi = p->pairs[2 * i], ri = p->pairs[2 * i + 1];
        const uint32_t nq = p->q_off[qi + 1] - p->q_off[qi];
        const uint32_t nr = p->r_off[ri + 1] - p->r_off[ri];
        const int32_t t0 = (int32_t)(p->r_row0[ri] / 32), nt = (int32_t)(ceil32(nr) / 32);
        // FP8: the base image in FP8_SLOTS quarters, each a candidate slot
        const int slots = p->kind == KIND_FP8 ? FP8_SLOTS : 1;
        const int32_t per = (nt + slots - 1) / slots;
        for (uint32_t g = 0; g * qpg < nq; g++) {
            for (int sl = 0; sl < slots; sl++) {
                PairTask w;
                w.t_begin = std::min(t0 + sl * per, t0 + nt);
                w.t_end = std::min(t0 + (sl + 1) * per, t0 + nt);
                w.qblk0 = (int32_t)((p->q_row0[qi] + g * qpg) / 32);
                w.r_lim = p->r_row0[ri] + nr;
                w.row_off = p->r_row0[ri];
                w.q_row0 = p->q_row0[qi];
                w.q_count = nq;
                w.out_off = p->out_off[i];
                w.slot = sl;
                t.push_back(w);
            }
        }
    }
    // Longest first, then each item to the block with the least work so far
    // (LPT): block b runs items [off[b], off[b + 1]) of the uploaded list.
    // An item's time is close to linear in its base tiles (measured on the
    // 512-pair workload: 1.04 us per tile plus 0.11 us); the earlier
    // round-robin of the sorted list (block b: items b, b + grid, ...) gave
    // the first blocks the largest item of every round and left 21% of the
    // grid's slot time idle.  (Grouping a pair's items on one XCD for L2
    // reuse measured slower: the kernel is bound by its top-2 epilogue.)
    std::stable_sort(t.begin(), t.end(), [](const PairTask& a, const PairTask& b) {
        return a.t_end - a.t_begin > b.t_end - b.t_begin;
    });
    p->n_tasks = (int)t.size();
    const long cap = (long)device_cu_count() * sh.pairs_blocks_per_cu();
    p->grid = (unsigned)std::max(1L, std::min<long>(cap, p->n_tasks));
    std::vector<std::vector<uint32_t>> bins(p->grid);
    {
        using Load = std::pair<double, uint32_t>;  // (work so far, block)
        std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
        for (uint32_t b = 0; b < p->grid; b++) heap.push({0.0, b});
        for (uint32_t i = 0; i < (uint32_t)t.size(); i++) {
            Load l = heap.top();
            heap.pop();
            bins[l.second].push_back(i);
            l.first += (double)(t[i].t_end - t[i].t_begin) + 0.1;
            heap.push(l);
        }
    }
    std::vector<PairTask> order;
    order.reserve(t.size());
    std::vector<int32_t> off(p->grid + 1, 0);
    for (uint32_t b = 0; b < p->grid; b++) {
        off[b] = (int32_t)order.size();
        for (uint32_t i : bins[b]) order.push_back(t[i]);
    }
    off[p->grid] = (int32_t)order.size();
    const size_t task_bytes = std::max<size_t>(1, order.size()) * sizeof(PairTask);
    p->off_bytes = task_bytes;
    SIFT_TRY(reserve(h, p->tasks, task_bytes + off.size() * sizeof(int32_t)));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));  // no queued launch still reads the old list
    if (!order.empty())
        SIFT_HIP_TRY(h, hipMemcpy(p->tasks.p, order.data(), order.size() * sizeof(PairTask), hipMemcpyHostToDevice));
    SIFT_HIP_TRY(h, hipMemcpy(static_cast<char*>(p->tasks.p) + task_bytes, off.data(), off.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice));
    return SIFT_OK;
}

FusedArgs make_args(sift_pairs_plan* p, uint32_t* idx, float* dist, uint32_t* matches, float ratio) {
    FusedArgs a{};
    a.q_tiles = static_cast<const uint4*>(p->qs.tiles.p);
    a.b_tiles = static_cast<const uint4*>(p->bs.tiles.p);
    a.b_nl = static_cast<const float*>(p->bs.norm_lane.p);
    a.part = nullptr;
    a.bound = nullptr;
    a.q_n = 0;
    a.q_stride = 0;
    a.r_n = 0;
    a.tile0 = 0;
    a.end_tile = (int)p->bs.ntiles;
    a.tiles_per_slab = 0;
    a.n_slabs = 0;
    a.n_groups = 0;
    a.tasks = static_cast<const PairTask*>(p->tasks.p);
    a.task_off = reinterpret_cast<const int32_t*>(static_cast<const char*>(p->tasks.p) + p->off_bytes);
    a.n_tasks = p->n_tasks;
    a.q_norm = static_cast<const float*>(p->qs.norm.p);
    a.out_idx = idx;
    a.out_dist = dist;
    a.out_matches = matches;
    a.ratio = ratio;
    a.comp_mul = 256;  // I8 composite = 256 q''.b' + ((c << 7) | slot) = key * 128 + slot
    a.out_cand = static_cast<uint32_t*>(p->cand.p);
    a.wait_polls = p->h->wait_polls;
    a.status = wait_status(p->h);
    return a;
}

// The execute sequence; prof=false under stream capture (no event records).
int run_sequence(sift_pairs_plan* p, const void* q, const void* r, uint32_t* idx, float* dist,
                 uint32_t* matches, float ratio, bool prof) {
    sift_context* h = p->h;
    int* flag = static_cast<int*>(p->flag.p);
    {
        ProfScope ps(h, prof ? PROF_PACK : -1);
        SIFT_TRY(pack_launch(h, q, p->type, p->q_rows, static_cast<const int32_t*>(p->qmap.p), p->kind,
                             PACK_QUERY, p->qs, flag));
        SIFT_TRY(pack_launch(h, r, p->type, p->r_rows, static_cast<const int32_t*>(p->rmap.p), p->kind,
                             PACK_BASE, p->bs, flag));
        if (p->kind == KIND_I8 && p->bs.rows_pad) {
            compose_slots<<<(p->bs.rows_pad + 255) / 256, 256, 0, h->stream>>>(static_cast<int*>(p->bs.norm_lane.p),
                                                                                p->bs.rows_pad);
            SIFT_HIP_TRY(h, hipGetLastError());
        }
    }
    if (p->kind == KIND_FP8) {
        ProfScope ps(h, prof ? PROF_PACK : -1);
        SIFT_TRY(to_u8(h, q, p->type, p->q_rows, p->q_u8, p->q_n2, flag));
        SIFT_TRY(to_u8(h, r, p->type, p->r_rows, p->r_u8, p->r_n2, flag));
    }
    if (p->n_tasks == 0) return SIFT_OK;
    {
        ProfScope ps(h, prof ? PROF_PAIRS : -1);
        p->shape->pairs_fn(make_args(p, idx, dist, matches, ratio), p->grid, h->stream);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    if (p->kind == KIND_FP8) {
        ProfScope ps(h, prof ? PROF_FP8 : -1);
        const unsigned grid = (unsigned)((p->total_q * 16 + 255) / 256);
        fp8_rerank<<<grid, 256, 0, h->stream>>>(
            static_cast<const uint32_t*>(p->cand.p), static_cast<const uint4*>(p->q_u8.p),
            static_cast<const int*>(p->q_n2.p), static_cast<const uint4*>(p->r_u8.p),
            static_cast<const int*>(p->r_n2.p), static_cast<const uint32_t*>(p->qsrc.p),
            static_cast<const uint32_t*>(p->boff.p), (uint32_t)p->total_q, idx, dist, nullptr, matches, ratio);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    return SIFT_OK;
}

// Time every listed shape of the kind on the real data and keep the one with
// the lowest total: the shapes in listed order (an untimed launch, then a
// timed one each), then in reverse order (one timed launch each), so that a
// clock still rising on the first execute favours no shape by its position.
int autotune(sift_pairs_plan* p, const void* q, const void* r, uint32_t* idx, float* dist,
             uint32_t* matches, float ratio) {
    sift_context* h = p->h;
    const FusedShape* cand[16];
    const int nc = shapes_of_kind(p->kind, cand, 16);
    if (nc <= 1) return SIFT_OK;
    SIFT_TRY(run_sequence(p, q, r, idx, dist, matches, ratio, false));  // packs stay valid
    hipEvent_t e0, e1;
    SIFT_HIP_TRY(h, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(h, SIFT_ERR_HIP, "hipEventCreate");
    }
    float sum_ms[16];
    for (int i = 0; i < nc; i++) sum_ms[i] = 0.f;
    int rc = SIFT_OK;
    for (int pass = 0; pass < 2 && rc == SIFT_OK; pass++) {
        for (int k = 0; k < nc && rc == SIFT_OK; k++) {
            const int i = pass ? nc - 1 - k : k;
            p->shape = cand[i];
            rc = build_tasks(p);
            if (rc != SIFT_OK || p->n_tasks == 0) break;
            const FusedArgs a = make_args(p, idx, dist, matches, ratio);
            if (pass == 0) p->shape->pairs_fn(a, p->grid, h->stream);
            (void)hipEventRecord(e0, h->stream);
            p->shape->pairs_fn(a, p->grid, h->stream);
            (void)hipEventRecord(e1, h->stream);
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = set_error(h, SIFT_ERR_HIP, "pairs autotune launch failed");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            sum_ms[i] += ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SIFT_TRY(rc);
    int best = 0;
    for (int i = 1; i < nc; i++)
        if (sum_ms[i] < sum_ms[best]) best = i;
    p->shape = cand[best];
    return build_tasks(p);
}

int capture(sift_pairs_plan* p, const void* q, const void* r, uint32_t* idx, float* dist,
            uint32_t* matches, float ratio) {
    sift_context* h = p->h;
    drop_graph(p);
    SIFT_HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = run_sequence(p, q, r, idx, dist, matches, ratio, false);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->stream, &g);
    if (rc != SIFT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return set_error(h, SIFT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    p->graph = g;
    const hipError_t ei = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        drop_graph(p);
        return set_error(h, SIFT_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    p->g_q = q;
    p->g_r = r;
    p->g_idx = idx;
    p->g_dist = dist;
    p->g_m = matches;
    p->g_ratio = ratio;
    return SIFT_OK;
}

int check_offsets(sift_context* h, uint32_t n, const uint32_t* off, const char* what) {
    if (n > 0 && !off) return set_error(h, SIFT_ERR_INVALID, std::string(what) + ": null offsets");
    if (n == 0) return SIFT_OK;
    if (off[0] != 0) return set_error(h, SIFT_ERR_INVALID, std::string(what) + ": offsets must start at 0");
    for (uint32_t i = 0; i < n; i++) {
        if (off[i + 1] < off[i])
            return set_error(h, SIFT_ERR_INVALID, std::string(what) + ": offsets must be non-decreasing");
        if (off[i + 1] - off[i] >= (1u << 24))
            return set_error(h, SIFT_ERR_UNSUPPORTED, std::string(what) + ": image of 2^24 rows or more");
    }
    return SIFT_OK;
}

int plan_create(sift_context* h, int type, int kind, uint32_t n_qimg, const uint32_t* q_off,
                uint32_t n_rimg, const uint32_t* r_off, uint32_t n_pairs, const uint32_t* pairs,
                unsigned flags, sift_pairs_plan** out) {
    SIFT_TRY(check_offsets(h, n_qimg, q_off, "query images"));
    SIFT_TRY(check_offsets(h, n_rimg, r_off, "base images"));
    if (n_pairs > 0 && !pairs) return set_error(h, SIFT_ERR_INVALID, "null pair list");
    for (uint32_t i = 0; i < n_pairs; i++)
        if (pairs[2 * i] >= n_qimg || pairs[2 * i + 1] >= n_rimg)
            return set_error(h, SIFT_ERR_INVALID, "pair " + std::to_string(i) + " names an image out of range");
    auto* p = new sift_pairs_plan();
    p->h = h;
    p->type = type;
    p->kind = kind;
    p->flags = flags;
    p->n_qimg = n_qimg;
    p->n_rimg = n_rimg;
    p->n_pairs = n_pairs;
    p->q_off.assign(q_off ? q_off : nullptr, q_off ? q_off + n_qimg + 1 : nullptr);
    p->r_off.assign(r_off ? r_off : nullptr, r_off ? r_off + n_rimg + 1 : nullptr);
    if (p->q_off.empty()) p->q_off.push_back(0);
    if (p->r_off.empty()) p->r_off.push_back(0);
    p->pairs.assign(pairs, pairs + 2 * (size_t)n_pairs);
    p->q_rows = p->q_off.back();
    p->r_rows = p->r_off.back();

    // tile-aligned image layout on both sides
    uint64_t qr = 0, rr = 0;
    p->q_row0.resize(n_qimg);
    p->r_row0.resize(n_rimg);
    for (uint32_t i = 0; i < n_qimg; i++) {
        p->q_row0[i] = (uint32_t)qr;
        qr += ceil32(p->q_off[i + 1] - p->q_off[i]);
    }
    // base images start on even tiles: the pairs kernel's composites carry the
    // tile's parity within its image (see compose_slots)
    for (uint32_t i = 0; i < n_rimg; i++) {
        p->r_row0[i] = (uint32_t)rr;
        rr += (p->r_off[i + 1] - p->r_off[i] + 63u) & ~63u;
    }
    // query groups may run up to one group (< QPAD rows) past their image
    const uint64_t q_pad = qr + QPAD;
    const uint64_t r_pad = std::max<uint64_t>(32, rr);
    p->out_off.resize(n_pairs);
    uint64_t o = 0;
    for (uint32_t i = 0; i < n_pairs; i++) {
        p->out_off[i] = (uint32_t)o;
        o += p->q_off[p->pairs[2 * i] + 1] - p->q_off[p->pairs[2 * i]];
    }
    p->total_q = o;
    if (q_pad >= (1ull << 31) || r_pad >= (1ull << 31) || o >= (1ull << 31)) {
        free_plan(p);
        return set_error(h, SIFT_ERR_UNSUPPORTED, "batch too large (2^31 packed rows or outputs)");
    }

    std::vector<int32_t> qm(q_pad, -1), rm(r_pad, -1);
    for (uint32_t i = 0; i < n_qimg; i++)
        for (uint32_t k = 0; k < p->q_off[i + 1] - p->q_off[i]; k++) qm[p->q_row0[i] + k] = (int32_t)(p->q_off[i] + k);
    for (uint32_t i = 0; i < n_rimg; i++)
        for (uint32_t k = 0; k < p->r_off[i + 1] - p->r_off[i]; k++) rm[p->r_row0[i] + k] = (int32_t)(p->r_off[i] + k);

    auto fail = [&](int rc) {
        free_plan(p);
        return rc;
    };
    int rc = SIFT_OK;
    auto setup = [&]() -> int {
        SIFT_TRY(reserve(h, p->qmap, qm.size() * 4));
        SIFT_TRY(reserve(h, p->rmap, rm.size() * 4));
        SIFT_HIP_TRY(h, hipMemcpy(p->qmap.p, qm.data(), qm.size() * 4, hipMemcpyHostToDevice));
        SIFT_HIP_TRY(h, hipMemcpy(p->rmap.p, rm.data(), rm.size() * 4, hipMemcpyHostToDevice));
        for (PackedSet* s : {&p->qs, &p->bs}) {
            const uint32_t rows = s == &p->qs ? (uint32_t)q_pad : (uint32_t)r_pad;
            s->rows = s == &p->qs ? p->q_rows : p->r_rows;
            s->rows_pad = rows;
            s->ntiles = rows / TILE_ROWS;
            s->kind = kind;
            SIFT_TRY(reserve(h, s->tiles, (size_t)s->ntiles * tile_bytes(kind)));
            SIFT_TRY(reserve(h, s->norm, (size_t)rows * 4));
        }
        SIFT_TRY(reserve(h, p->bs.norm_lane, (size_t)r_pad * 4));
        SIFT_TRY(reserve(h, p->flag, 16));
        SIFT_HIP_TRY(h, hipMemset(p->flag.p, 0, 16));
        if (kind == KIND_FP8) {
            SIFT_TRY(reserve(h, p->cand, (size_t)std::max<uint64_t>(1, p->total_q) * 16 * 4));
            SIFT_TRY(reserve(h, p->q_u8, (size_t)p->q_rows * 128));
            SIFT_TRY(reserve(h, p->q_n2, (size_t)p->q_rows * 4));
            SIFT_TRY(reserve(h, p->r_u8, (size_t)p->r_rows * 128));
            SIFT_TRY(reserve(h, p->r_n2, (size_t)p->r_rows * 4));
            std::vector<uint32_t> qs_(std::max<uint64_t>(1, p->total_q)), bo(std::max<uint64_t>(1, p->total_q));
            for (uint32_t i = 0; i < n_pairs; i++) {
                const uint32_t qi = p->pairs[2 * i], ri = p->pairs[2 * i + 1];
                for (uint32_t k = 0; k < p->q_off[qi + 1] - p->q_off[qi]; k++) {
                    qs_[p->out_off[i] + k] = p->q_off[qi] + k;
                    bo[p->out_off[i] + k] = p->r_off[ri];
                }
            }
            SIFT_TRY(reserve(h, p->qsrc, qs_.size() * 4));
            SIFT_TRY(reserve(h, p->boff, bo.size() * 4));
            SIFT_HIP_TRY(h, hipMemcpy(p->qsrc.p, qs_.data(), qs_.size() * 4, hipMemcpyHostToDevice));
            SIFT_HIP_TRY(h, hipMemcpy(p->boff.p, bo.data(), bo.size() * 4, hipMemcpyHostToDevice));
        }
        p->shape = pick_shape(kind);
        if (!p->shape) return set_error(h, SIFT_ERR_INVALID, "no tile shape for the compute kind");
        SIFT_TRY(build_tasks(p));
        p->tune_pending = (flags & SIFT_PAIRS_AUTOTUNE) != 0;
        return SIFT_OK;
    };
    rc = setup();
    if (rc != SIFT_OK) return fail(rc);
    *out = p;
    return SIFT_OK;
}

int kind_of(sift_context* h, int compute, int* kind) {
    switch (compute) {
        case SIFT_COMPUTE_BF16: *kind = KIND_BF16; return SIFT_OK;
        case SIFT_COMPUTE_F16: *kind = KIND_F16; return SIFT_OK;
        case SIFT_COMPUTE_F32: *kind = KIND_F32; return SIFT_OK;
        case SIFT_COMPUTE_I8: *kind = KIND_I8; return SIFT_OK;
        case SIFT_COMPUTE_FP8: *kind = KIND_FP8; return SIFT_OK;
        default:
            return set_error(h, SIFT_ERR_INVALID,
                             "a pairs plan needs compute BF16, F16, F32, I8 or FP8 (got " + std::to_string(compute) + ")");
    }
}

bool out_ok(const void* p) { return !p || device_accessible(p); }

}  // namespace

namespace sift {
int fp8_rerank_launch(sift_context* h, const uint32_t* cand, const void* q_u8, const int* q_n2,
                      const void* r_u8, const int* r_n2, const uint32_t* qsrc, const uint32_t* boff,
                      uint32_t total, uint32_t* idx, float* dist, sift_float4* sorted, uint32_t* matches,
                      float ratio) {
    if (total == 0) return SIFT_OK;
    const unsigned grid = (unsigned)(((size_t)total * 16 + 255) / 256);
    fp8_rerank<<<grid, 256, 0, h->stream>>>(cand, static_cast<const uint4*>(q_u8), q_n2,
                                            static_cast<const uint4*>(r_u8), r_n2, qsrc, boff, total, idx,
                                            dist, sorted, matches, ratio);
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}
}  // namespace sift

extern "C" {

int sift_pairs_plan_create(sift_handle_t h, int type, int compute, uint32_t n_qimg, const uint32_t* q_off,
                           uint32_t n_rimg, const uint32_t* r_off, uint32_t n_pairs, const uint32_t* pairs,
                           unsigned flags, sift_pairs_plan_t* plan) {
    if (!h || !plan) return set_error(h, SIFT_ERR_INVALID, "null handle or plan pointer");
    *plan = nullptr;
    if (type < SIFT_FLOAT_DEVICE || type > SIFT_U8_DEVICE)
        return set_error(h, SIFT_ERR_INVALID, "a pairs plan takes device descriptors (type 1..4)");
    if (flags & ~(SIFT_PAIRS_GRAPH | SIFT_PAIRS_AUTOTUNE)) return set_error(h, SIFT_ERR_INVALID, "unknown plan flags");
    int kind = 0;
    SIFT_TRY(kind_of(h, compute, &kind));
    SIFT_HIP_TRY(h, hipSetDevice(h->device));
    return plan_create(h, type, kind, n_qimg, q_off, n_rimg, r_off, n_pairs, pairs, flags, plan);
}

int sift_pairs_plan_execute(sift_handle_t h, sift_pairs_plan_t p, const void* q, const void* r, uint32_t* idx,
                            float* dist, uint32_t* matches, float ratio, int blocking) {
    if (!h || !p) return set_error(h, SIFT_ERR_INVALID, "null handle or plan");
    if (p->h != h) return set_error(h, SIFT_ERR_INVALID, "plan belongs to another handle");
    if ((p->q_rows && !q) || (p->r_rows && !r)) return set_error(h, SIFT_ERR_INVALID, "null descriptor pointer");
    if (!out_ok(idx) || !out_ok(dist) || !out_ok(matches))
        return set_error(h, SIFT_ERR_INVALID, "outputs must be device-accessible memory");
    if ((p->q_rows && !device_accessible(q)) || (p->r_rows && !device_accessible(r)))
        return set_error(h, SIFT_ERR_INVALID, "descriptors must be device-accessible memory");
    SIFT_HIP_TRY(h, hipSetDevice(h->device));
    if (p->total_q == 0) return SIFT_OK;
    if (p->tune_pending) {
        p->tune_pending = false;
        drop_graph(p);
        SIFT_TRY(autotune(p, q, r, idx, dist, matches, ratio));
    }
    if (p->flags & SIFT_PAIRS_GRAPH) {
        if (!p->exec || p->g_q != q || p->g_r != r || p->g_idx != idx || p->g_dist != dist || p->g_m != matches ||
            p->g_ratio != ratio)
            SIFT_TRY(capture(p, q, r, idx, dist, matches, ratio));
        ProfScope ps(h, PROF_PAIRS);
        SIFT_HIP_TRY(h, hipGraphLaunch(p->exec, h->stream));
    } else {
        SIFT_TRY(run_sequence(p, q, r, idx, dist, matches, ratio, true));
    }
    if (blocking) return check_wait(h, "sift_pairs_plan_execute");
    return SIFT_OK;
}

int sift_pairs_plan_set_shape(sift_pairs_plan_t p, int qw, int waves, int tps) {
    if (!p) return SIFT_ERR_INVALID;
    const FusedShape* cand[16];
    const int nc = shapes_of_kind(p->kind, cand, 16);
    for (int i = 0; i < nc; i++) {
        if (cand[i]->qw == qw && cand[i]->waves == waves && cand[i]->tps == tps) {
            drop_graph(p);
            p->tune_pending = false;
            p->shape = cand[i];
            return build_tasks(p);
        }
    }
    return set_error(p->h, SIFT_ERR_INVALID, "shape not listed for this compute kind");
}

int sift_pairs_plan_info(sift_pairs_plan_t p, int64_t* o) {
    if (!p || !o) return SIFT_ERR_INVALID;
    o[0] = (int64_t)p->total_q;
    o[1] = p->n_tasks;
    o[2] = p->grid;
    o[3] = p->shape ? p->shape->qw : 0;
    o[4] = p->shape ? p->shape->waves : 0;
    o[5] = p->shape ? p->shape->tps : 0;
    o[6] = p->exec ? 1 : 0;
    o[7] = p->kind;
    // sticky: some packed value was not an integer in [0,255] (reads a device flag)
    int bad = 0;
    if (p->h) (void)hipStreamSynchronize(p->h->stream);
    if (p->flag.p && hipMemcpy(&bad, p->flag.p, 4, hipMemcpyDeviceToHost) != hipSuccess) bad = -1;
    o[8] = bad;
    return SIFT_OK;
}

int sift_pairs_plan_destroy(sift_pairs_plan_t p) {
    if (!p) return SIFT_OK;
    if (p->h) (void)hipStreamSynchronize(p->h->stream);
    free_plan(p);
    return SIFT_OK;
}

int sift_match_pairs(sift_handle_t h, const void* q, const void* r, int type, uint32_t n_qimg,
                     const uint32_t* q_off, uint32_t n_rimg, const uint32_t* r_off, uint32_t n_pairs,
                     const uint32_t* pairs, int compute, uint32_t* idx, float* dist, uint32_t* matches,
                     float ratio) {
    if (!h) return set_error(nullptr, SIFT_ERR_INVALID, "null handle");
    if (!valid_type(type)) return set_error(h, SIFT_ERR_INVALID, "unknown input type " + std::to_string(type));
    SIFT_HIP_TRY(h, hipSetDevice(h->device));
    const uint32_t nq = (n_qimg && q_off) ? q_off[n_qimg] : 0;
    const uint32_t nr = (n_rimg && r_off) ? r_off[n_rimg] : 0;
    if (type == SIFT_FLOAT_HOST) {
        if ((nq && !q) || (nr && !r)) return set_error(h, SIFT_ERR_INVALID, "null descriptor pointer");
        SIFT_TRY(stage_host_input(h, q, nq, h->stage_q, &q));
        SIFT_TRY(stage_host_input(h, r, nr, h->stage_r, &r));
        type = SIFT_FLOAT_DEVICE;
    }
    int kind = 0;
    if (compute == SIFT_COMPUTE_AUTO) {
        SIFT_TRY(resolve_kind(h, q, r, type, nq, nr, compute, KIND_F32, &kind));
    } else {
        SIFT_TRY(kind_of(h, compute, &kind));
    }
    sift_pairs_plan* p = nullptr;
    SIFT_TRY(plan_create(h, type, kind, n_qimg, q_off, n_rimg, r_off, n_pairs, pairs, 0u, &p));
    int rc = sift_pairs_plan_execute(h, p, q, r, idx, dist, matches, ratio, 1);
    int bad = 0;
    if (rc == SIFT_OK && kind != KIND_F32 &&
        hipMemcpy(&bad, p->flag.p, 4, hipMemcpyDeviceToHost) == hipSuccess && bad)
        rc = set_error(h, SIFT_ERR_NOT_EXACT, "pairs with BF16/F16/I8 compute need integer data in [0,255]");
    sift_pairs_plan_destroy(p);
    return rc;
}

}  // extern "C"
