// pack.hip — convert descriptors into the MFMA operand layout (HBM-bound).
//
// Replaces the reference's float2half (src/knn_brute.cu:315-338): one pass
// reads the input rows once and writes
//   tiles     : 32-row tiles in MFMA operand order, [tile][group][lane][16 B],
//               so the fused kernel's operand loads are 1 KB contiguous per
//               wave instruction;
//   norm      : per-row squared norm (natural order);
//   norm_lane : per-tile norms in the order the 32x32 accumulator holds rows
//               (lane half h, register j -> row 8*(j/4) + 4*h + j%4).
// Conversions are round-to-nearest-even (exact for integers in [0,255]).
#include "internal.h"

namespace sift {
namespace {

template <int TYPE>
__device__ inline void load8(const void* in, size_t row, int blk, float v[8]) {
    if constexpr (TYPE == SIFT_FLOAT_DEVICE || TYPE == SIFT_FLOAT_HOST) {
        const float4* p = reinterpret_cast<const float4*>(static_cast<const float*>(in) + row * 128 + blk * 8);
        float4 a = p[0], b = p[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else if constexpr (TYPE == SIFT_HALF_DEVICE) {
        uint4 r = *reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(in) + row * 128 + blk * 8);
        uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            _Float16 lo, hi;
            uint16_t l16 = (uint16_t)(w[i] & 0xFFFF), h16 = (uint16_t)(w[i] >> 16);
            __builtin_memcpy(&lo, &l16, 2);
            __builtin_memcpy(&hi, &h16, 2);
            v[2 * i] = (float)lo;
            v[2 * i + 1] = (float)hi;
        }
    } else if constexpr (TYPE == SIFT_BF16_DEVICE) {
        uint4 r = *reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(in) + row * 128 + blk * 8);
        uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[2 * i] = __uint_as_float(w[i] << 16);
            v[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
        }
    } else {  // SIFT_U8_DEVICE
        uint2 r = *reinterpret_cast<const uint2*>(static_cast<const uint8_t*>(in) + row * 128 + blk * 8);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[i] = (float)((r.x >> (8 * i)) & 0xFF);
            v[4 + i] = (float)((r.y >> (8 * i)) & 0xFF);
        }
    }
}

__device__ inline uint32_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ inline uint32_t f2h(float f) {
    _Float16 h = (_Float16)f;
    uint16_t b;
    __builtin_memcpy(&b, &h, 2);
    return b;
}
__device__ inline bool exact_u8(float x) { return x >= 0.f && x <= 255.f && x == rintf(x); }

// OCP FP8 e4m3 (bias 7, no infinities, largest finite 448), round to nearest
// even, saturating at 448; v must be finite.  Subnormals step by 2^-9; a
// normal value 2^E (1 + m/8) is ((E + 7) << 3) | m, so rounding the scaled
// mantissa up to 8 carries into the exponent by plain addition.
__device__ inline uint32_t f32_to_e4m3(float v) {
    const uint32_t sign = (__float_as_uint(v) >> 24) & 0x80u;
    const float a = fabsf(v);
    uint32_t code;
    if (a < 0.015625f) {  // below 2^-6: subnormal range (the top rounds up to code 8 = 2^-6)
        code = (uint32_t)rintf(a * 512.f);
    } else {
        const int e = (int)((__float_as_uint(a) >> 23) & 0xFF) - 127;  // floor(log2 a)
        const float mant = __uint_as_float(__float_as_uint(a) & 0x807FFFFFu | 0x3F800000u);  // a / 2^e
        code = (uint32_t)((e + 7) << 3) + (uint32_t)rintf((mant - 1.f) * 8.f);
    }
    return sign | min(code, 0x7Eu);
}
// Per-row power-of-two scale: the largest k with max * 2^k <= 448, clamped to
// [-100, 100] (k = 0 for an all-zero row).
__device__ inline int fp8_exponent(float mx) {
    if (!(mx > 0.f)) return 0;
    const int e = (int)((__float_as_uint(mx) >> 23) & 0xFF) - 127;  // mx = m * 2^e, m in [1, 2)
    const float m = __uint_as_float(__float_as_uint(mx) & 0x807FFFFFu | 0x3F800000u);
    const int k = (m <= 1.75f ? 8 : 7) - e;
    return max(-100, min(100, k));
}
// One atomic per wave: the lowest active lane with the condition set reports.
__device__ inline void flag_wave(bool cond, int* flag) {
    const unsigned long long m = __ballot(cond);
    if (m && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicOr(flag, 1);
}

// 16 threads per row, thread b handles dims [8b, 8b+8).
template <int TYPE, int KIND>
__global__ __launch_bounds__(256) void pack_kernel(const void* __restrict__ in, uint32_t n,
                                                   uint32_t rows_pad, uint4* __restrict__ tiles,
                                                   float* __restrict__ norm,
                                                   float* __restrict__ norm_lane,
                                                   int* __restrict__ norm_half, int role,
                                                   int* __restrict__ bad,
                                                   const int32_t* __restrict__ rowmap, int qexp) {
    const bool is_base = role == PACK_BASE;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t row = gid >> 4;
    const int blk = gid & 15;
    if (row >= rows_pad) return;  // rows_pad is a multiple of 32; whole 16-lane groups exit
    // rowmap (optional): packed row -> input row, -1 for a pad row
    const int32_t src = rowmap ? rowmap[row] : (row < n ? (int32_t)row : -1);
    const bool real = src >= 0;
    float v[8];
    if (real) {
        load8<TYPE>(in, (size_t)src, blk, v);
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 8; i++) ok &= exact_u8(v[i]);
        flag_wave(!ok, bad);
        if (role == PACK_QUANT) {  // integers clamp(rint(x * 2^qexp), -255, 255): exact in BF16
            const float sc = __uint_as_float((uint32_t)(127 + qexp) << 23);
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = fminf(fmaxf(rintf(v[i] * sc), -255.f), 255.f);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = (KIND == KIND_I8) ? 128.f : 0.f;
    }
    // Query operands of the float kinds are stored as -2q (exact: a power-of-two
    // scale); the norm below is computed from q itself.
    const float opscale = (role == PACK_QUERY && KIND != KIND_I8) ? -2.f : 1.f;
    const uint32_t t = row >> 5, i32 = row & 31;
    int sum_b = 0;  // KIND_I8 base: sum of b' over the row
    constexpr int G = (KIND == KIND_F32) ? 17 : (KIND == KIND_I8 ? 4 : (KIND == KIND_FP8 ? 5 : 9));
    uint4* tile = tiles + (size_t)t * G * 64;
    float nrm;
    if constexpr (KIND == KIND_BF16 || KIND == KIND_F16) {
        uint32_t w[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint32_t lo = (KIND == KIND_BF16) ? f2bf(v[2 * i]) : f2h(v[2 * i]);
            uint32_t hi = (KIND == KIND_BF16) ? f2bf(v[2 * i + 1]) : f2h(v[2 * i + 1]);
            w[i] = lo | (hi << 16);
        }
        const int g = blk >> 1, h = blk & 1;
        if (opscale != 1.f) {
            uint32_t o[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float lo, hi;
                if constexpr (KIND == KIND_BF16) {
                    lo = __uint_as_float(w[i] << 16);
                    hi = __uint_as_float(w[i] & 0xFFFF0000u);
                    o[i] = f2bf(opscale * lo) | (f2bf(opscale * hi) << 16);
                } else {
                    _Float16 a, b;
                    uint16_t la = (uint16_t)(w[i] & 0xFFFF), hb = (uint16_t)(w[i] >> 16);
                    __builtin_memcpy(&a, &la, 2);
                    __builtin_memcpy(&b, &hb, 2);
                    o[i] = f2h(opscale * (float)a) | (f2h(opscale * (float)b) << 16);
                }
            }
            tile[g * 64 + i32 + 32 * h] = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
            tile[g * 64 + i32 + 32 * h] = make_uint4(w[0], w[1], w[2], w[3]);
        }
        // norm of the rounded values the MFMA sees (identical for integer data)
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float lo, hi;
            if constexpr (KIND == KIND_BF16) {
                lo = __uint_as_float(w[i] << 16);
                hi = __uint_as_float(w[i] & 0xFFFF0000u);
            } else {
                _Float16 a, b;
                uint16_t la = (uint16_t)(w[i] & 0xFFFF), hb = (uint16_t)(w[i] >> 16);
                __builtin_memcpy(&a, &la, 2);
                __builtin_memcpy(&b, &hb, 2);
                lo = (float)a;
                hi = (float)b;
            }
            s = fmaf(lo, lo, s);
            s = fmaf(hi, hi, s);
        }
        nrm = s;
    } else if constexpr (KIND == KIND_F32) {
        // group g = blk holds k-steps 4g..4g+3; lane half h holds dims 8g + 2i + h
        tile[blk * 64 + i32] = make_uint4(__float_as_uint(opscale * v[0]), __float_as_uint(opscale * v[2]),
                                          __float_as_uint(opscale * v[4]), __float_as_uint(opscale * v[6]));
        tile[blk * 64 + i32 + 32] = make_uint4(__float_as_uint(opscale * v[1]), __float_as_uint(opscale * v[3]),
                                               __float_as_uint(opscale * v[5]), __float_as_uint(opscale * v[7]));
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < 8; i++) s = fmaf(v[i], v[i], s);
        nrm = s;
    } else if constexpr (KIND == KIND_FP8) {
        // e4m3 of x * 2^k (queries: -x * 2^k, with the factor 2 of -2 q.b in
        // the scale), k per row so its largest |value| maps to at most 448;
        // the MFMA applies the scale byte 127 - k (queries 128 - k) itself.
        // Data group c, lane half h, byte p holds dim 64 (c/2) + 32 h + 16 (c%2) + p.
        float mx = 0.f, s = 0.f;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            mx = fmaxf(mx, fabsf(v[i]));
            s = fmaf(v[i], v[i], s);
        }
        mx = fmaxf(mx, __shfl_xor(mx, 8, 16));
        mx = fmaxf(mx, __shfl_xor(mx, 4, 16));
        mx = fmaxf(mx, __shfl_xor(mx, 2, 16));
        mx = fmaxf(mx, __shfl_xor(mx, 1, 16));
        const int k = fp8_exponent(mx);
        const float sc = __uint_as_float((uint32_t)(127 + k) << 23) * (role == PACK_QUERY ? -1.f : 1.f);
        uint32_t w[2] = {0u, 0u};
#pragma unroll
        for (int i = 0; i < 8; i++) w[i >> 2] |= f32_to_e4m3(v[i] * sc) << (8 * (i & 3));
        const int m = blk >> 3, h = (blk >> 2) & 1, c = 2 * m + ((blk >> 1) & 1);
        uint2* dst = reinterpret_cast<uint2*>(tile + c * 64 + i32 + 32 * h) + (blk & 1);
        *dst = make_uint2(w[0], w[1]);
        if (blk < 2) {  // scale group: lane halves 0/1 of the row
            const uint32_t sbyte = (uint32_t)(127 - k + (role == PACK_QUERY ? 1 : 0));
            tile[4 * 64 + i32 + 32 * blk] = make_uint4(sbyte, 0u, 0u, 0u);
        }
        nrm = s;
    } else {  // KIND_I8: base b' = x - 128, queries q'' = 127 - x (both in int8); lane half h
              // holds dims 32g + 16h + 0..15.  Base rows also sum b' (for c = |b'|^2 + 2 sum b').
        int si = 0, sb = 0;
        uint32_t w[2] = {0u, 0u};
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int x = (int)v[i] - 128;
            si += x * x;
            sb += x;
            // base pad rows are zero vectors (b' = 0): their key is exactly
            // the PAD_KEY norm, so a pairs composite key * 128 + slot of a pad
            // row stays below 2^31 for every query
            const int op = role == PACK_QUERY ? -x - 1 : (real || !is_base ? x : 0);
            w[i >> 2] |= ((uint32_t)(op & 0xFF)) << (8 * (i & 3));
        }
        const int g = blk >> 2, h = (blk >> 1) & 1, half = blk & 1;
        uint2* dst = reinterpret_cast<uint2*>(tile + g * 64 + i32 + 32 * h) + half;
        *dst = make_uint2(w[0], w[1]);
        nrm = __int_as_float(si);
        if (is_base) {
            sb += __shfl_xor(sb, 8, 16);
            sb += __shfl_xor(sb, 4, 16);
            sb += __shfl_xor(sb, 2, 16);
            sb += __shfl_xor(sb, 1, 16);
            sum_b = sb;
        }
    }
    // reduce the 16 partial norms of this row (lanes 16k..16k+15)
    if constexpr (KIND == KIND_I8) {
        int s = __float_as_int(nrm);
        s += __shfl_xor(s, 8, 16);
        s += __shfl_xor(s, 4, 16);
        s += __shfl_xor(s, 2, 16);
        s += __shfl_xor(s, 1, 16);
        if (!real && is_base) s = 0x7FFFFFFF;
        nrm = __int_as_float(s);
    } else {
        float s = nrm;
        s += __shfl_xor(s, 8, 16);
        s += __shfl_xor(s, 4, 16);
        s += __shfl_xor(s, 2, 16);
        s += __shfl_xor(s, 1, 16);
        if (!real && is_base) s = __builtin_huge_valf();
        nrm = s;
    }
    if (blk == 0) {
        norm[row] = nrm;
        if (is_base) {
            const int hl = (i32 >> 2) & 1, j = 4 * (i32 >> 3) + (i32 & 3);
            const size_t li = (size_t)t * 32 + hl * 16 + j;
            if constexpr (KIND == KIND_I8) {
                const int c = real ? __float_as_int(nrm) + 2 * sum_b : PAD_KEY;
                norm_lane[li] = __int_as_float(c);
                if (norm_half) norm_half[li] = c >> 1;
            } else {
                norm_lane[li] = real ? nrm : (float)PAD_KEY;
            }
        }
    }
    // Norm group (float kinds): threads 0/1 of the row write lane halves 0/1.
    //   base  : |b|^2 (F32), or |b|^2 / 256 split into three parts (BF16/F16,
    //           exact for integer norms below 2^24)
    //   query : the multiplier (1, or 256 per part) so the MFMA adds |b|^2
    //   base pad rows: PAD_KEY, above every real key, so a pad row can only
    //   reach a top-2 when the base has fewer than two rows
    //   plain rows: zero
    if constexpr (KIND == KIND_BF16 || KIND == KIND_F16 || KIND == KIND_F32) {
        if (blk < 2) {
            uint4 u = make_uint4(0u, 0u, 0u, 0u);
            if (blk == 0 && (role == PACK_QUERY || role == PACK_BASE)) {
                const float bn = real ? nrm : (float)PAD_KEY;
                if constexpr (KIND == KIND_F32) {
                    u.x = __float_as_uint(role == PACK_QUERY ? 1.f : bn);
                } else if (role == PACK_QUERY) {
                    const uint32_t m = (KIND == KIND_BF16) ? f2bf(256.f) : f2h(256.f);
                    u.x = m | (m << 16);
                    u.y = m;
                } else {
                    float v = bn * (1.f / 256.f);
                    uint32_t p[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        if constexpr (KIND == KIND_BF16) {
                            p[i] = f2bf(v);
                            v -= __uint_as_float(p[i] << 16);
                        } else {
                            p[i] = f2h(v);
                            _Float16 hv;
                            uint16_t hb = (uint16_t)p[i];
                            __builtin_memcpy(&hv, &hb, 2);
                            v -= (float)hv;
                        }
                    }
                    u.x = p[0] | (p[1] << 16);
                    u.y = p[2];
                }
            }
            tile[(G - 1) * 64 + i32 + 32 * blk] = u;
        }
    }
}

template <int TYPE>
__global__ __launch_bounds__(256) void check_kernel(const void* __restrict__ in, uint32_t n,
                                                    int* __restrict__ bad) {
    const size_t total = (size_t)n * 16;
    bool ok = true;
    for (size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x; gid < total;
         gid += (size_t)gridDim.x * blockDim.x) {
        float v[8];
        load8<TYPE>(in, gid >> 4, (int)(gid & 15), v);
#pragma unroll
        for (int i = 0; i < 8; i++) ok &= exact_u8(v[i]);
    }
    flag_wave(!ok, bad);
}

// Compact copy for gathers: uint8 rows and int32 squared norms (16 threads
// per row, 8 dims each).  Values must be integers in [0,255] (flagged).
template <int TYPE>
__global__ __launch_bounds__(256) void to_u8_kernel(const void* __restrict__ in, uint32_t n,
                                                    uint2* __restrict__ out, int* __restrict__ nrm,
                                                    int* __restrict__ bad) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t row = gid >> 4;
    const int blk = gid & 15;
    if (row >= n) return;  // n rows of 16 lanes: whole groups exit together
    float v[8];
    load8<TYPE>(in, row, blk, v);
    bool ok = true;
    uint32_t w[2] = {0u, 0u};
    int s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        ok &= exact_u8(v[i]);
        const int x = (int)fminf(fmaxf(v[i], 0.f), 255.f);
        s += x * x;
        w[i >> 2] |= (uint32_t)x << (8 * (i & 3));
    }
    flag_wave(!ok, bad);
    out[(size_t)row * 16 + blk] = make_uint2(w[0], w[1]);
    s += __shfl_xor(s, 8, 16);
    s += __shfl_xor(s, 4, 16);
    s += __shfl_xor(s, 2, 16);
    s += __shfl_xor(s, 1, 16);
    if (blk == 0) nrm[row] = s;
}

// Largest |value| of n rows (float bit patterns of non-negative values order
// as unsigned integers): one atomicMax per wave.
template <int TYPE>
__global__ __launch_bounds__(256) void absmax_kernel(const void* __restrict__ in, uint32_t n,
                                                     uint32_t* __restrict__ out) {
    const size_t total = (size_t)n * 16;
    float m = 0.f;
    for (size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x; gid < total;
         gid += (size_t)gridDim.x * blockDim.x) {
        float v[8];
        load8<TYPE>(in, gid >> 4, (int)(gid & 15), v);
#pragma unroll
        for (int i = 0; i < 8; i++) m = fmaxf(m, fabsf(v[i]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(out, __float_as_uint(m));
}

// float32 copy [n][128] of n input rows (exact conversions).
template <int TYPE>
__global__ __launch_bounds__(256) void to_f32_kernel(const void* __restrict__ in, uint32_t n,
                                                     float4* __restrict__ out) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t row = gid >> 4;
    const int blk = gid & 15;
    if (row >= n) return;
    float v[8];
    load8<TYPE>(in, row, blk, v);
    out[((size_t)row * 16 + blk) * 2] = make_float4(v[0], v[1], v[2], v[3]);
    out[((size_t)row * 16 + blk) * 2 + 1] = make_float4(v[4], v[5], v[6], v[7]);
}

template <int TYPE>
int launch_pack_t(sift_context* h, const void* in, uint32_t n, uint32_t rows_pad, int kind,
                  int role, PackedSet& o, int* bad, const int32_t* rowmap, int qexp = 0) {
    const size_t threads = (size_t)rows_pad * 16;
    const unsigned grid = (unsigned)((threads + 255) / 256);
    if (grid == 0) return SIFT_OK;
    auto* tiles = static_cast<uint4*>(o.tiles.p);
    auto* norm = static_cast<float*>(o.norm.p);
    auto* nl = static_cast<float*>(o.norm_lane.p);
    auto* nh = static_cast<int*>(o.norm_half.p);
    switch (kind) {
        case KIND_BF16:
            pack_kernel<TYPE, KIND_BF16><<<grid, 256, 0, h->stream>>>(in, n, rows_pad, tiles, norm, nl, nh, role, bad, rowmap, qexp);
            break;
        case KIND_F16:
            pack_kernel<TYPE, KIND_F16><<<grid, 256, 0, h->stream>>>(in, n, rows_pad, tiles, norm, nl, nh, role, bad, rowmap, qexp);
            break;
        case KIND_F32:
            pack_kernel<TYPE, KIND_F32><<<grid, 256, 0, h->stream>>>(in, n, rows_pad, tiles, norm, nl, nh, role, bad, rowmap, qexp);
            break;
        case KIND_I8:
            pack_kernel<TYPE, KIND_I8><<<grid, 256, 0, h->stream>>>(in, n, rows_pad, tiles, norm, nl, nh, role, bad, rowmap, qexp);
            break;
        case KIND_FP8:
            pack_kernel<TYPE, KIND_FP8><<<grid, 256, 0, h->stream>>>(in, n, rows_pad, tiles, norm, nl, nh, role, bad, rowmap, qexp);
            break;
        default: return set_error(h, SIFT_ERR_INVALID, "pack: bad kind");
    }
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

}  // namespace

int pack_rows(sift_context* h, const void* in, int type, uint32_t n, uint32_t rows_pad, int kind,
              int role, PackedSet& o, int* bad_dev) {
    const bool is_base = role == PACK_BASE;
    if (rows_pad % TILE_ROWS) return set_error(h, SIFT_ERR_INVALID, "pack: rows_pad not a multiple of 32");
    const uint32_t ntiles = rows_pad / TILE_ROWS;
    SIFT_TRY(reserve(h, o.tiles, (size_t)ntiles * tile_bytes(kind)));
    SIFT_TRY(reserve(h, o.norm, (size_t)rows_pad * 4));
    if (is_base) SIFT_TRY(reserve(h, o.norm_lane, (size_t)rows_pad * 4));
    if (is_base && kind == KIND_I8) SIFT_TRY(reserve(h, o.norm_half, (size_t)rows_pad * 4));
    o.rows = n;
    o.rows_pad = rows_pad;
    o.ntiles = ntiles;
    o.kind = kind;
    ProfScope ps(h, PROF_PACK);
    return pack_launch(h, in, type, n, nullptr, kind, role, o, bad_dev);
}

int pack_launch(sift_context* h, const void* in, int type, uint32_t n, const int32_t* rowmap, int kind,
                int role, PackedSet& o, int* bad_dev) {
    const uint32_t rows_pad = o.rows_pad;
    switch (type) {
        case SIFT_FLOAT_HOST:
        case SIFT_FLOAT_DEVICE:
            return launch_pack_t<SIFT_FLOAT_DEVICE>(h, in, n, rows_pad, kind, role, o, bad_dev, rowmap);
        case SIFT_HALF_DEVICE:
            return launch_pack_t<SIFT_HALF_DEVICE>(h, in, n, rows_pad, kind, role, o, bad_dev, rowmap);
        case SIFT_BF16_DEVICE:
            return launch_pack_t<SIFT_BF16_DEVICE>(h, in, n, rows_pad, kind, role, o, bad_dev, rowmap);
        case SIFT_U8_DEVICE:
            return launch_pack_t<SIFT_U8_DEVICE>(h, in, n, rows_pad, kind, role, o, bad_dev, rowmap);
        default: return set_error(h, SIFT_ERR_INVALID, "pack: unknown input type " + std::to_string(type));
    }
}

int check_exact(sift_context* h, const void* in, int type, uint32_t n, int* bad_dev) {
    if (n == 0) return SIFT_OK;
    const size_t threads = (size_t)n * 16;
    unsigned grid = (unsigned)std::min<size_t>((threads + 255) / 256, 4096);
    switch (type) {
        case SIFT_FLOAT_HOST:
        case SIFT_FLOAT_DEVICE: check_kernel<SIFT_FLOAT_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, bad_dev); break;
        case SIFT_HALF_DEVICE: check_kernel<SIFT_HALF_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, bad_dev); break;
        case SIFT_BF16_DEVICE: check_kernel<SIFT_BF16_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, bad_dev); break;
        case SIFT_U8_DEVICE: return SIFT_OK;  // always exact
        default: return set_error(h, SIFT_ERR_INVALID, "check: unknown input type");
    }
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

int pack_side(sift_context* h, const void* in, int type, uint32_t n, uint32_t rows_pad, int kind,
              int role, PackedSet& o, int* bad_dev) {
    SIFT_TRY(pack_rows(h, in, type, n, rows_pad, kind, role, o, bad_dev));
    if (kind == KIND_FP8) SIFT_TRY(to_u8(h, in, type, n, o.u8, o.n2, bad_dev));
    return SIFT_OK;
}

void release_set(PackedSet& s) {
    for (DevBuf* b : {&s.tiles, &s.norm, &s.norm_lane, &s.norm_half, &s.u8, &s.n2}) release(*b);
}

int pack_quant(sift_context* h, const void* in, int type, uint32_t n, uint32_t rows_pad, int qexp,
               PackedSet& o, int* bad_dev) {
    if (rows_pad % TILE_ROWS) return set_error(h, SIFT_ERR_INVALID, "pack: rows_pad not a multiple of 32");
    if (qexp < -100 || qexp > 100) return set_error(h, SIFT_ERR_INVALID, "pack: scale exponent out of range");
    const uint32_t ntiles = rows_pad / TILE_ROWS;
    SIFT_TRY(reserve(h, o.tiles, (size_t)ntiles * tile_bytes(KIND_BF16)));
    SIFT_TRY(reserve(h, o.norm, (size_t)rows_pad * 4));
    o.rows = n;
    o.rows_pad = rows_pad;
    o.ntiles = ntiles;
    o.kind = KIND_BF16;
    ProfScope ps(h, PROF_PACK);
    switch (type) {
        case SIFT_FLOAT_HOST:
        case SIFT_FLOAT_DEVICE:
            return launch_pack_t<SIFT_FLOAT_DEVICE>(h, in, n, rows_pad, KIND_BF16, PACK_QUANT, o, bad_dev, nullptr, qexp);
        case SIFT_HALF_DEVICE:
            return launch_pack_t<SIFT_HALF_DEVICE>(h, in, n, rows_pad, KIND_BF16, PACK_QUANT, o, bad_dev, nullptr, qexp);
        case SIFT_BF16_DEVICE:
            return launch_pack_t<SIFT_BF16_DEVICE>(h, in, n, rows_pad, KIND_BF16, PACK_QUANT, o, bad_dev, nullptr, qexp);
        case SIFT_U8_DEVICE:
            return launch_pack_t<SIFT_U8_DEVICE>(h, in, n, rows_pad, KIND_BF16, PACK_QUANT, o, bad_dev, nullptr, qexp);
        default: return set_error(h, SIFT_ERR_INVALID, "pack: unknown input type " + std::to_string(type));
    }
}

int abs_max(sift_context* h, const void* in, int type, uint32_t n, uint32_t* out_dev) {
    if (n == 0) return SIFT_OK;
    const unsigned grid = (unsigned)std::min<size_t>(((size_t)n * 16 + 255) / 256, 4096);
    switch (type) {
        case SIFT_FLOAT_HOST:
        case SIFT_FLOAT_DEVICE: absmax_kernel<SIFT_FLOAT_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, out_dev); break;
        case SIFT_HALF_DEVICE: absmax_kernel<SIFT_HALF_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, out_dev); break;
        case SIFT_BF16_DEVICE: absmax_kernel<SIFT_BF16_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, out_dev); break;
        case SIFT_U8_DEVICE: absmax_kernel<SIFT_U8_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, out_dev); break;
        default: return set_error(h, SIFT_ERR_INVALID, "abs_max: unknown input type");
    }
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

int to_f32(sift_context* h, const void* in, int type, uint32_t n, DevBuf& out) {
    SIFT_TRY(reserve(h, out, std::max<size_t>((size_t)n * 512, 1)));
    if (n == 0) return SIFT_OK;
    const unsigned grid = (unsigned)(((size_t)n * 16 + 255) / 256);
    auto* o = static_cast<float4*>(out.p);
    switch (type) {
        case SIFT_FLOAT_HOST:
        case SIFT_FLOAT_DEVICE: to_f32_kernel<SIFT_FLOAT_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, o); break;
        case SIFT_HALF_DEVICE: to_f32_kernel<SIFT_HALF_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, o); break;
        case SIFT_BF16_DEVICE: to_f32_kernel<SIFT_BF16_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, o); break;
        case SIFT_U8_DEVICE: to_f32_kernel<SIFT_U8_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, o); break;
        default: return set_error(h, SIFT_ERR_INVALID, "to_f32: unknown input type");
    }
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

int to_u8(sift_context* h, const void* in, int type, uint32_t n, DevBuf& out, DevBuf& nrm,
          int* bad_dev) {
    SIFT_TRY(reserve(h, out, (size_t)n * 128));
    SIFT_TRY(reserve(h, nrm, (size_t)n * 4));
    if (n == 0) return SIFT_OK;
    const unsigned grid = (unsigned)(((size_t)n * 16 + 255) / 256);
    auto* o = static_cast<uint2*>(out.p);
    auto* r = static_cast<int*>(nrm.p);
    switch (type) {
        case SIFT_FLOAT_HOST:
        case SIFT_FLOAT_DEVICE: to_u8_kernel<SIFT_FLOAT_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, o, r, bad_dev); break;
        case SIFT_HALF_DEVICE: to_u8_kernel<SIFT_HALF_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, o, r, bad_dev); break;
        case SIFT_BF16_DEVICE: to_u8_kernel<SIFT_BF16_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, o, r, bad_dev); break;
        case SIFT_U8_DEVICE: to_u8_kernel<SIFT_U8_DEVICE><<<grid, 256, 0, h->stream>>>(in, n, o, r, bad_dev); break;
        default: return set_error(h, SIFT_ERR_INVALID, "to_u8: unknown input type");
    }
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

}  // namespace sift
