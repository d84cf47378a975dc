// noalloc.hip — brute-force index queries that allocate nothing.
//
// sift_bf_index_query needs device scratch beyond its inputs and outputs: the
// packed query tiles and norms, the per-(slab, query) top-2 partials of the
// fused kernel and the per-query seed bounds.  It keeps them in grow-only
// buffers of the handle, which may call hipMalloc.  Callers that cannot
// allocate (code under HIP graph capture, tight serving loops) use
// sift_bf_index_query_noalloc instead: the same kernels with that scratch in
// one file-scope device array of SIFT_NOALLOC_SCRATCH_BYTES per GPU.
#include <algorithm>
#include <mutex>

#include "internal.h"

// Device scratch shared by every no-allocation query on a GPU (each GPU the
// library runs on holds its own copy until the process ends).  Host code
// never uses this name as an address: the per-device address comes from
// hipGetSymbolAddress at handle creation.
static __device__ __align__(256) unsigned char scratch[SIFT_NOALLOC_SCRATCH_BYTES];

namespace sift {
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};

__global__ void set_word(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

int noalloc_init(sift_context* h) {
    const int dev = h->device;
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(h, SIFT_ERR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[dev]) return SIFT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the device is current
    if (e != hipSuccess || !p)
        return set_error(h, SIFT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
    return SIFT_OK;
}

}  // namespace sift

using namespace sift;

extern "C" {

int sift_bf_index_query_noalloc(sift_handle_t h, sift_bf_index_t ix, const void* q, int type, uint32_t q_n,
                                uint32_t* idx, float* dist, sift_float4* sorted, uint32_t* matches, float ratio,
                                int blocking) {
    if (!h || !ix) return set_error(h, SIFT_ERR_INVALID, "null handle/index");
    if (ix->h != h) return set_error(h, SIFT_ERR_INVALID, "the index belongs to another handle");
    if (type < SIFT_FLOAT_DEVICE || type > SIFT_U8_DEVICE)
        return set_error(h, SIFT_ERR_INVALID, "sift_bf_index_query_noalloc takes device rows (type 1..4)");
    if (ix->kind == KIND_FP8)
        return set_error(h, SIFT_ERR_UNSUPPORTED, "sift_bf_index_query_noalloc: FP8 indexes need more scratch");
    if (q_n == 0) return SIFT_OK;
    if (!q) return set_error(h, SIFT_ERR_INVALID, "null queries");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = (h->device >= 0 && h->device < MAX_DEVICES) ? g_scratch[h->device] : nullptr;
    }
    if (!base) return set_error(h, SIFT_ERR_HIP, "no-allocation scratch not available on this device");

    // [query tiles | query norms | partials | bounds] in the device array
    const uint32_t rows_pad = query_rows_pad(ix->kind, q_n);
    size_t bound_bytes = 0;
    const size_t part_bytes = knn2_work_bytes(h, ix->kind, ix->base.ntiles, q_n, ix->r_n, &bound_bytes);
    const size_t tiles_bytes = (size_t)(rows_pad / TILE_ROWS) * tile_bytes(ix->kind);
    const size_t o_norm = align256(tiles_bytes);
    const size_t o_part = o_norm + align256((size_t)rows_pad * 4);
    const size_t o_bound = o_part + align256(part_bytes);
    const size_t need = o_bound + align256(bound_bytes);
    if (part_bytes == 0 || need > NOALLOC_SCRATCH_BYTES)
        return set_error(h, SIFT_ERR_INVALID,
                         "sift_bf_index_query_noalloc: " + std::to_string(q_n) + " queries against " +
                             std::to_string(ix->r_n) + " rows need " + std::to_string(need) +
                             " bytes of scratch, above the " + std::to_string((size_t)NOALLOC_SCRATCH_BYTES) +
                             "-byte maximum (split the queries)");
    auto* s = static_cast<unsigned char*>(base);
    PackedSet qs;  // views into the scratch array, never released
    qs.tiles.p = s;
    qs.tiles.cap = tiles_bytes;
    qs.norm.p = s + o_norm;
    qs.norm.cap = (size_t)rows_pad * 4;
    qs.rows = q_n;
    qs.rows_pad = rows_pad;
    qs.ntiles = rows_pad / TILE_ROWS;
    qs.kind = ix->kind;
    Knn2Work w;
    w.part = s + o_part;
    w.part_bytes = part_bytes;
    w.bound = reinterpret_cast<int*>(s + o_bound);
    w.bound_bytes = bound_bytes;

    int* flag = static_cast<int*>(h->flag.p);  // allocated with the handle
    Knn2Out o{idx, dist, sorted, matches, ratio};
    SIFT_TRY(bf_index_run(h, ix, q, type, q_n, o, flag, qs, &w));
    if (blocking) {
        int fl[3] = {0, 0, 0};
        SIFT_HIP_TRY(h, hipMemcpyAsync(fl, flag, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
        SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
        SIFT_TRY(wait_error(h, fl[2], "sift_bf_index_query_noalloc"));
        if (fl[0] && ix->kind == KIND_I8)
            return set_error(h, SIFT_ERR_NOT_EXACT, "I8 indexes need integer queries in [0,255]");
    }
    return SIFT_OK;
}

}  // extern "C"

namespace sift {
// The exactness flag at the start of a query, written by a kernel (no
// memset node, so the sequence may be captured in a graph).
int set_flag_word(sift_context* h, int* p, int v) {
    set_word<<<1, 64, 0, h->stream>>>(p, v);
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}
}  // namespace sift
