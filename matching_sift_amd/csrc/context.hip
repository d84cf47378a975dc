// context.hip — handle lifetime, errors, device-memory plumbing, profiling.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "internal.h"

namespace {
thread_local std::string tl_error;
thread_local int tl_status = SIFT_OK;
}  // namespace

namespace sift {

int set_error(sift_context* h, int code, const std::string& msg) {
    if (h) h->last_error = msg;
    tl_error = msg;
    tl_status = code;
    if (std::getenv("SIFT_VERBOSE")) std::fprintf(stderr, "[sift2nn] error %d: %s\n", code, msg.c_str());
    return code;
}

int reserve(sift_context* h, DevBuf& b, size_t bytes) {
    if (bytes <= b.cap && b.p) return SIFT_OK;
    if (b.p) {
        // the buffer may still be read by queued work on the handle's stream
        hipStreamSynchronize(h->stream);
        hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return set_error(h, SIFT_ERR_OOM, "hipMalloc(" + std::to_string(want) + ") failed: " +
                                              hipGetErrorString(e));
    }
    b.cap = want;
    return SIFT_OK;
}

void release(DevBuf& b) {
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

int* wait_status(sift_context* h) { return static_cast<int*>(h->flag.p) + 2; }

int wait_error(sift_context* h, int status, const char* what) {
    if (!status) return SIFT_OK;
    SIFT_HIP_TRY(h, hipMemsetAsync(wait_status(h), 0, 4, h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return set_error(h, SIFT_ERR_HIP,
                     std::string(what) + ": a stage hand-off wait of the fused kernel gave up after " +
                         std::to_string(h->wait_polls) + " polls; the results of that launch are not valid");
}

int check_wait(sift_context* h, const char* what) {
    int st = 0;
    SIFT_HIP_TRY(h, hipMemcpyAsync(&st, wait_status(h), 4, hipMemcpyDeviceToHost, h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return wait_error(h, st, what);
}

const char* prof_name(int k) {
    static const char* names[PROF_COUNT] = {"pack", "fused", "merge", "seed", "lsh_proj", "lsh_sort",
                                            "lsh_query", "fp8", "pairs", "mlp", "lsh_aux"};
    return (k >= 0 && k < PROF_COUNT) ? names[k] : "?";
}

ProfScope::ProfScope(sift_context* h_, int k_) : h(h_), k(k_) {
    if (!h->profile || k < 0) return;  // k < 0: no timing (e.g. under stream capture)
    if (h->npend == h->cap_pend) prof_flush(h);
    if (h->npend == h->cap_pend) return;
    auto& p = h->pend[h->npend];
    hipEventRecord(p.a, h->stream);
}

ProfScope::~ProfScope() {
    if (!h->profile || k < 0 || h->npend == h->cap_pend) return;
    auto& p = h->pend[h->npend];
    hipEventRecord(p.b, h->stream);
    p.k = k;
    h->npend++;
}

void prof_flush(sift_context* h) {
    for (int i = 0; i < h->npend; i++) {
        auto& p = h->pend[i];
        hipEventSynchronize(p.b);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            h->prof_ms[p.k] += ms;
            h->prof_n[p.k] += 1;
        }
    }
    h->npend = 0;
}

bool device_accessible(const void* p) {
    hipPointerAttribute_t attr;
    const bool ok = hipPointerGetAttributes(&attr, p) == hipSuccess &&
                    (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
                     attr.type == hipMemoryTypeHost);
    (void)hipGetLastError();
    return ok;
}

size_t input_row_bytes(int type) {
    switch (type) {
        case SIFT_FLOAT_HOST:
        case SIFT_FLOAT_DEVICE: return 512;
        case SIFT_HALF_DEVICE:
        case SIFT_BF16_DEVICE: return 256;
        case SIFT_U8_DEVICE: return 128;
        default: return 0;
    }
}

int stage_host_input(sift_context* h, const void* host, uint32_t n, DevBuf& buf, const void** dev) {
    size_t bytes = (size_t)n * 512;
    SIFT_TRY(reserve(h, buf, bytes));
    if (bytes) SIFT_HIP_TRY(h, hipMemcpyAsync(buf.p, host, bytes, hipMemcpyHostToDevice, h->stream));
    *dev = buf.p;
    return SIFT_OK;
}

}  // namespace sift

using namespace sift;

extern "C" {

int sift_create(sift_handle_t* handle) {
    if (!handle) return set_error(nullptr, SIFT_ERR_INVALID, "sift_create: null handle pointer");
    *handle = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return set_error(nullptr, SIFT_ERR_HIP, "sift_create: no HIP device available");
    auto* h = new sift_context();
    hipGetDevice(&h->device);
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return set_error(nullptr, SIFT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    h->cap_pend = 256;
    h->pend = new sift_context::Pending[h->cap_pend];
    for (int i = 0; i < h->cap_pend; i++) {
        hipEventCreate(&h->pend[i].a);
        hipEventCreate(&h->pend[i].b);
    }
    // flags and the sticky wait status exist from the start, zeroed
    if (reserve(h, h->flag, 16) != SIFT_OK || hipMemset(h->flag.p, 0, 16) != hipSuccess) {
        sift_destroy(h);
        return set_error(nullptr, SIFT_ERR_HIP, "sift_create: status buffer");
    }
    if (const char* env = std::getenv("SIFT_WAIT_POLLS")) {
        const long v = std::strtol(env, nullptr, 10);
        h->wait_polls = (int)std::max(0L, std::min(v, 1L << 24));
    }
    (void)noalloc_init(h);  // on failure the no-allocation entry point reports it
    *handle = h;
    tl_status = SIFT_OK;
    return SIFT_OK;
}

void sift_lsh_state_free(sift_context* h);  // lsh.hip

int sift_destroy(sift_handle_t h) {
    if (!h) return SIFT_OK;
    hipStreamSynchronize(h->stream);
    sift_lsh_state_free(h);
    release_set(h->q_set);
    release_set(h->b_set);
    release(h->part); release(h->thresh); release(h->cand); release(h->flag); release(h->stage_q); release(h->stage_r);
    for (int i = 0; i < h->cap_pend; i++) {
        hipEventDestroy(h->pend[i].a);
        hipEventDestroy(h->pend[i].b);
    }
    delete[] h->pend;
    hipStreamDestroy(h->stream);
    delete h;
    return SIFT_OK;
}

const char* sift_last_error(sift_handle_t h) { return h ? h->last_error.c_str() : tl_error.c_str(); }
int sift_last_status(void) { return tl_status; }

int sift_synchronize(sift_handle_t h) {
    if (!h) return set_error(nullptr, SIFT_ERR_INVALID, "null handle");
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return check_wait(h, "sift_synchronize");
}

void* sift_get_stream(sift_handle_t h) { return h ? (void*)h->stream : nullptr; }

int sift_set_seed(sift_handle_t h, uint64_t seed) {
    if (!h) return set_error(nullptr, SIFT_ERR_INVALID, "null handle");
    h->seed = seed;
    return SIFT_OK;
}

int sift_device_name(char* buf, int len) {
    hipDeviceProp_t p;
    int dev = 0;
    hipGetDevice(&dev);
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return SIFT_ERR_HIP;
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return SIFT_OK;
}

int sift_malloc(void** ptr, size_t bytes) {
    if (!ptr) return set_error(nullptr, SIFT_ERR_INVALID, "sift_malloc: null");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return set_error(nullptr, SIFT_ERR_OOM, hipGetErrorString(e));
    return SIFT_OK;
}
int sift_free(void* ptr) { return hipFree(ptr) == hipSuccess ? SIFT_OK : SIFT_ERR_HIP; }
int sift_host_alloc(void** ptr, size_t bytes) {
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocMapped);
    if (e != hipSuccess) return set_error(nullptr, SIFT_ERR_OOM, hipGetErrorString(e));
    return SIFT_OK;
}
int sift_host_free(void* ptr) { return hipHostFree(ptr) == hipSuccess ? SIFT_OK : SIFT_ERR_HIP; }
int sift_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? SIFT_OK : set_error(nullptr, SIFT_ERR_HIP, hipGetErrorString(e));
}
int sift_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? SIFT_OK : set_error(nullptr, SIFT_ERR_HIP, hipGetErrorString(e));
}
int sift_memset(void* dst, int value, size_t bytes) {
    hipError_t e = hipMemset(dst, value, bytes);
    return e == hipSuccess ? SIFT_OK : set_error(nullptr, SIFT_ERR_HIP, hipGetErrorString(e));
}

int sift_profile_enable(sift_handle_t h, int on) {
    if (!h) return SIFT_ERR_INVALID;
    if (!on) prof_flush(h);
    h->profile = on != 0;
    return SIFT_OK;
}
int sift_profile_reset(sift_handle_t h) {
    if (!h) return SIFT_ERR_INVALID;
    prof_flush(h);
    for (int i = 0; i < PROF_COUNT; i++) {
        h->prof_ms[i] = 0;
        h->prof_n[i] = 0;
    }
    return SIFT_OK;
}
int sift_profile_read(sift_handle_t h, const char* kernel, double* total_ms, long* launches) {
    if (!h || !kernel) return SIFT_ERR_INVALID;
    prof_flush(h);
    for (int i = 0; i < PROF_COUNT; i++) {
        if (std::strcmp(kernel, prof_name(i)) == 0) {
            if (total_ms) *total_ms = h->prof_ms[i];
            if (launches) *launches = h->prof_n[i];
            return SIFT_OK;
        }
    }
    return set_error(h, SIFT_ERR_INVALID, std::string("unknown kernel name ") + kernel);
}

}  // extern "C"
