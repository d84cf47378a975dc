// internal.h — shared declarations of the sift2nn HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/sift2nn.h"

namespace sift {

// ------------------------------------------------------------- errors ----
int set_error(sift_context* h, int code, const std::string& msg);
#define SIFT_HIP_TRY(h, expr)                                                              \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return ::sift::set_error((h), SIFT_ERR_HIP,                                    \
                                     std::string(#expr ": ") + hipGetErrorString(_e));     \
    } while (0)
#define SIFT_TRY(expr)               \
    do {                             \
        int _rc = (expr);            \
        if (_rc != SIFT_OK) return _rc; \
    } while (0)

// ------------------------------------------------------ device scratch ----
// Grow-only device buffer owned by a handle: after the first call at a given
// size no further allocation happens.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};
int reserve(sift_context* h, DevBuf& b, size_t bytes);
void release(DevBuf& b);
// Device status word of the handle (h->flag word 2): a fused-kernel wave whose
// stage hand-off wait ran out of polls ORs 1 into it.  check_wait syncs the
// handle's stream, reads it, clears it and returns SIFT_ERR_HIP if it was set.
int* wait_status(sift_context* h);
int check_wait(sift_context* h, const char* what);
// The same check on a status value already copied to the host.
int wait_error(sift_context* h, int status, const char* what);

// ---------------------------------------------------------- profiling ----
enum ProfKernel { PROF_PACK = 0, PROF_FUSED, PROF_MERGE, PROF_SEED, PROF_LSH_PROJ, PROF_LSH_SORT,
                  PROF_LSH_QUERY, PROF_FP8, PROF_PAIRS, PROF_MLP, PROF_LSH_AUX, PROF_COUNT };
const char* prof_name(int k);
// RAII pair of events around one launch on the handle's stream.
struct ProfScope {
    sift_context* h;
    int k;
    ProfScope(sift_context* h_, int k_);
    ~ProfScope();
};
void prof_flush(sift_context* h);  // accumulate completed event pairs (syncs)

// Packed operand layout: a "tile" is 32 rows of 128 dims stored in MFMA
// operand order, [group][64 lanes][16 bytes]; lane l holds row l%32.  The
// float kinds append one "norm group" (K = 128 + 16): base rows carry their
// squared norm there and query rows the matching multiplier, so the MFMA
// itself adds |b|^2 to -2 q.b.
enum Kind { KIND_BF16 = 1, KIND_F16 = 2, KIND_F32 = 3, KIND_I8 = 4, KIND_FP8 = 5 };
// Key of a base pad row (zero data, this squared norm).  Real keys of SIFT
// data t = |b|^2 - 2 q.b are at most 128 * 255^2 = 8,323,200 (float kinds) or
// 128 * 128^2 + 2 * 128^3 = 6,291,456 (I8, values shifted by -128), so a pad
// row never outranks a real one.  65504 * 256 is exact in the 3-part BF16
// and FP16 norm split, and key * 32 + 31 still fits in int32.
constexpr int PAD_KEY = 65504 * 256;
constexpr int TILE_ROWS = 32;
inline int groups_per_tile(int kind) {
    // FP8: 4 groups of e4m3 data + 1 group carrying each row's scale byte
    return kind == KIND_F32 ? 17 : (kind == KIND_I8 ? 4 : (kind == KIND_FP8 ? 5 : 9));
}
inline size_t tile_bytes(int kind) { return (size_t)groups_per_tile(kind) * 64 * 16; }

struct PackedSet {
    DevBuf tiles;  // [tiles][groups][64][16B]
    DevBuf norm;   // float[rows_pad] (int32 for KIND_I8), natural row order
    DevBuf norm_lane;  // [tiles][32] norms in accumulator-lane order (base only; I8: see below)
    DevBuf norm_half;  // KIND_I8 base, brute force: norm_lane >> 1 (the MFMA's initial value)
    DevBuf u8, n2;     // KIND_FP8: uint8 rows [rows][128] and int32 |x|^2 for the exact re-rank
    uint32_t rows = 0, rows_pad = 0, ntiles = 0;
    int kind = 0;
};

// Pack n rows of `type` input into `kind` tiles padded to rows_pad rows.
// PACK_BASE: also writes norm_lane; pad rows get norm = +inf (INT32_MAX for
//            I8) so they never rank.  KIND_I8 stores base rows as b' = x - 128
//            and queries as q'' = 127 - x = -(x - 128) - 1; the base's
//            norm_lane holds c = |b'|^2 + 2 sum(b'), so the rank key
//            |b'|^2 - 2 q'.b' equals 2 q''.b' + c.
// PACK_QUERY: float kinds store -2q (the fused kernel's operand).
// PACK_PLAIN: values as given (hyperplanes, projection inputs).
// `bad` (device int) is OR-ed with 1 if a value is not an integer in [0,255].
// PACK_QUANT: plain BF16 tiles of clamp(rint(x * 2^qexp), -255, 255) (LSH
//            hashing of data that is not all integers in [0,255]).
enum PackRole { PACK_PLAIN = 0, PACK_BASE = 1, PACK_QUERY = 2, PACK_QUANT = 3 };
int pack_rows(sift_context* h, const void* in, int type, uint32_t n, uint32_t rows_pad, int kind,
              int role, PackedSet& out, int* bad_dev);
// Launch the pack kernel into an already sized set (o.rows_pad rows; no
// allocation, so it may run under stream capture).  rowmap (device, optional):
// packed row -> input row, -1 for a pad row; without it rows >= n are pads.
int pack_launch(sift_context* h, const void* in, int type, uint32_t n, const int32_t* rowmap, int kind,
                int role, PackedSet& o, int* bad_dev);
// PACK_QUANT tiles (KIND_BF16) of n rows with scale 2^qexp, qexp in [-100, 100].
int pack_quant(sift_context* h, const void* in, int type, uint32_t n, uint32_t rows_pad, int qexp,
               PackedSet& o, int* bad_dev);
// Largest |value| of n rows as float bits, atomicMax-ed into *out_dev (caller zeroes it).
int abs_max(sift_context* h, const void* in, int type, uint32_t n, uint32_t* out_dev);
// float32 copy [n][128] of n input rows.
int to_f32(sift_context* h, const void* in, int type, uint32_t n, DevBuf& out);
// uint8 copy [n][128] and int32 squared norms of n input rows (flags non-integers).
int to_u8(sift_context* h, const void* in, int type, uint32_t n, DevBuf& out, DevBuf& nrm,
          int* bad_dev);
// pack_rows, plus for KIND_FP8 the uint8 copy and norms of the same rows.
int pack_side(sift_context* h, const void* in, int type, uint32_t n, uint32_t rows_pad, int kind,
              int role, PackedSet& o, int* bad_dev);
void release_set(PackedSet& s);
// Exact re-rank of 16 candidates per output (candidate row c of output o is
// base row boff[o] + c; query row qsrc[o]; null qsrc / boff: o and 0):
// int32 distances from the uint8 copies, the two smallest (d, row), and the
// output formats.
int fp8_rerank_launch(sift_context* h, const uint32_t* cand, const void* q_u8, const int* q_n2,
                      const void* r_u8, const int* r_n2, const uint32_t* qsrc, const uint32_t* boff,
                      uint32_t total, uint32_t* idx, float* dist, sift_float4* sorted, uint32_t* matches,
                      float ratio);
// Check that every value is an integer in [0,255]; result in *bad_dev (OR-ed).
int check_exact(sift_context* h, const void* in, int type, uint32_t n, int* bad_dev);

// Rows for the exact FP32 re-rank of non-integer data: the caller's input
// rows (types 1..4), or for the base a packed set's tiles: F32 (any data),
// BF16 / F16 (integer data in [0,255], held exactly).
constexpr int SRC_F32_TILES = 100;
constexpr int SRC_BF16_TILES = 101;
constexpr int SRC_F16_TILES = 102;
struct ExactSrc {
    const void* q = nullptr;
    int q_type = 0;
    const void* r = nullptr;
    int r_type = 0;
    const int32_t* q_map = nullptr;  // query row of output i: q_map[i] (null: i)
};
struct Knn2Out {
    uint32_t* idx = nullptr;     // [q][2]
    float* dist = nullptr;       // [q][2]
    sift_float4* sorted = nullptr;
    uint32_t* matches = nullptr;
    float ratio = 0.8f;
    // Float kinds: device flag, nonzero when some input value is not an
    // integer in [0,255] (the pack kernels set it).  The fused pass then only
    // ranks coarsely: seed bounds are widened by the kind's rounding error,
    // the 16 best coarse candidates of each query are re-ranked with the
    // sequential FP32 sum of host_brute over `src`, and those results are the
    // outputs.  Null: integer data assumed (exact keys).
    const int* coarse = nullptr;
    ExactSrc src;
};
// Caller-provided device scratch for knn2_packed: the slab partials and the
// per-query bounds.  With it knn2_packed allocates nothing and never times
// launch shapes (it uses the shape already picked on the handle, or the
// kind's default).
struct Knn2Work {
    void* part = nullptr;
    size_t part_bytes = 0;
    int* bound = nullptr;
    size_t bound_bytes = 0;
};
// Fused distance + top-2 over packed sets, then merge into the outputs
// (work: null = the handle's grow-only buffers).
int knn2_packed(sift_context* h, const PackedSet& qs, const PackedSet& bs, uint32_t q_n,
                uint32_t r_n, const Knn2Out& out, const Knn2Work* work = nullptr);
// Scratch bytes knn2_packed with a Knn2Work needs: partials (returned) and
// bounds (*bound_bytes), for the shape it would use on this handle.
size_t knn2_work_bytes(sift_context* h, int kind, uint32_t n_tiles, uint32_t q_n, uint32_t r_n,
                       size_t* bound_bytes);
// Packed query rows for q_n queries (whole launch groups).
uint32_t query_rows_pad(int kind, uint32_t q_n);

bool valid_type(int type);
// Resolve SIFT_COMPUTE_* to a packed kind.  AUTO: integer data in [0,255] ->
// I8 (exact), otherwise `fallback`.  AUTO reads one device flag (syncs).
int resolve_kind(sift_context* h, const void* q, const void* r, int type, uint32_t q_n,
                 uint32_t r_n, int compute, int fallback, int* kind);

// Host pointer for SIFT_FLOAT_HOST input: stage it into a device buffer.
int stage_host_input(sift_context* h, const void* host, uint32_t n, DevBuf& buf,
                     const void** dev);
size_t input_row_bytes(int type);
// True for device, managed and pinned (mapped) host memory.
bool device_accessible(const void* p);

// Learned hash (mlp.hip): forward over packed BF16 plain tiles into codes
// of table `table` (combine: prefix the table index above the b code bits).
int mlp_forward(sift_mlp* m, const PackedSet& xs, uint32_t n, float* h_out, float* z_out, uint32_t* codes,
                uint32_t stride, int table, bool combine);
// The same forward for L networks of equal width in one launch per 16: codes
// (and z, [table][stride][32], if non-null) of every table.
int mlp_forward_tables(sift_mlp* const* ms, int L, const PackedSet& xs, uint32_t n, float* z_out,
                       uint32_t* codes, uint32_t stride, bool combine);
int mlp_nbits(const sift_mlp* m);

}  // namespace sift

// Brute-force index (knn2.hip): the base packed once.
struct sift_bf_index {
    sift_context* h;
    sift::PackedSet base;
    uint32_t r_n;
    int kind;
    bool coarse = false;  // F32 over data that is not all integers in [0,255]
};
namespace sift {
// No-allocation queries (noalloc.hip): the device scratch array's size, the
// per-device address lookup done at handle creation, and the kernel that
// sets the exactness flag where a memset node would not be capturable.
constexpr size_t NOALLOC_SCRATCH_BYTES = SIFT_NOALLOC_SCRATCH_BYTES;
int noalloc_init(sift_context* h);
int set_flag_word(sift_context* h, int* p, int v);
// Pack q_n queries into qset (work: sized in advance, no allocation) and run
// the index's fused pass; flag = device int for the non-integer test.
int bf_index_run(sift_context* h, sift_bf_index* ix, const void* q, int type, uint32_t q_n, const Knn2Out& out,
                 int* flag, PackedSet& qset, const Knn2Work* work);
}  // namespace sift

struct sift_context {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t seed = 0x5eed2022ull;
    std::string last_error;

    sift::PackedSet q_set, b_set;
    sift::DevBuf part;      // fused-kernel partial top-2
    sift::DevBuf thresh;    // per-query admission bounds from the seeding pass
    sift::DevBuf cand;      // FP8 coarse pass: 16 candidate rows per query
    sift::DevBuf flag;      // int[4]: [0] input outside the exact domain (cleared per call),
                            // [2] a kernel's bounded stage wait gave up (sticky until read)
    sift::DevBuf stage_q, stage_r;  // host-input staging

    // profiling
    bool profile = false;
    double prof_ms[sift::PROF_COUNT] = {};
    long prof_n[sift::PROF_COUNT] = {};
    struct Pending {
        hipEvent_t a, b;
        int k;
    };
    Pending* pend = nullptr;
    int npend = 0, cap_pend = 0;

    // poll limit of the fused kernel's stage hand-off waits (SIFT_WAIT_POLLS at
    // creation; about 64 cycles per poll)
    int wait_polls = 1 << 16;

    // launch shape picked by timing, per compute kind (index among the
    // kind's listed shapes; -1 = not yet timed)
    int tuned_shape[8] = {-1, -1, -1, -1, -1, -1, -1, -1};

    // LSH / learned-hash state lives in their own translation units
    void* lsh_state = nullptr;
};
