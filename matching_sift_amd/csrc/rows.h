// rows.h — descriptor rows as floats on the device, and the reference's
// sequential squared distance over them (exact re-ranks of float data).
#pragma once

#include "internal.h"

namespace sift {

// Eight consecutive dims [8c, 8c + 8) of row `row` as floats (exact
// conversions): input rows of type 1..4, or an F32 packed set (group c,
// lane half h holds dims 8c + 2i + h, see pack.hip).
__device__ inline void load8_src(const void* p, int type, size_t row, int c, float v[8]) {
    switch (type) {
        case SIFT_FLOAT_DEVICE: {
            const float4* s = reinterpret_cast<const float4*>(static_cast<const float*>(p) + row * 128 + c * 8);
            const float4 a = s[0], b = s[1];
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
            break;
        }
        case SIFT_HALF_DEVICE: {
            const uint4 r = *reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(p) + row * 128 + c * 8);
            const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                _Float16 lo, hi;
                const uint16_t l16 = (uint16_t)(w[i] & 0xFFFF), h16 = (uint16_t)(w[i] >> 16);
                __builtin_memcpy(&lo, &l16, 2);
                __builtin_memcpy(&hi, &h16, 2);
                v[2 * i] = (float)lo;
                v[2 * i + 1] = (float)hi;
            }
            break;
        }
        case SIFT_BF16_DEVICE: {
            const uint4 r = *reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(p) + row * 128 + c * 8);
            const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                v[2 * i] = __uint_as_float(w[i] << 16);
                v[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
            }
            break;
        }
        case SIFT_U8_DEVICE: {
            const uint2 r = *reinterpret_cast<const uint2*>(static_cast<const uint8_t*>(p) + row * 128 + c * 8);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                v[i] = (float)((r.x >> (8 * i)) & 0xFF);
                v[4 + i] = (float)((r.y >> (8 * i)) & 0xFF);
            }
            break;
        }
        case SRC_BF16_TILES:
        case SRC_F16_TILES: {  // 9 groups of 64 x 16 B per 32-row tile; dims 8c.. in group c/2, half c%2
            const uint4 r = static_cast<const uint4*>(p)[(row >> 5) * 9 * 64 + (c >> 1) * 64 + (row & 31) + 32 * (c & 1)];
            const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (type == SRC_BF16_TILES) {
                    v[2 * i] = __uint_as_float(w[i] << 16);
                    v[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
                } else {
                    _Float16 lo, hi;
                    const uint16_t l16 = (uint16_t)(w[i] & 0xFFFF), h16 = (uint16_t)(w[i] >> 16);
                    __builtin_memcpy(&lo, &l16, 2);
                    __builtin_memcpy(&hi, &h16, 2);
                    v[2 * i] = (float)lo;
                    v[2 * i + 1] = (float)hi;
                }
            }
            break;
        }
        default: {  // SRC_F32_TILES: 17 groups of 64 x 16 B per 32-row tile
            const uint4* t = static_cast<const uint4*>(p) + (row >> 5) * 17 * 64 + c * 64 + (row & 31);
            const uint4 e = t[0], o = t[32];
            v[0] = __uint_as_float(e.x); v[2] = __uint_as_float(e.y);
            v[4] = __uint_as_float(e.z); v[6] = __uint_as_float(e.w);
            v[1] = __uint_as_float(o.x); v[3] = __uint_as_float(o.y);
            v[5] = __uint_as_float(o.z); v[7] = __uint_as_float(o.w);
            break;
        }
    }
}

// host_lenght (src/knn_brute.cu:873-885): the squared distance as a
// sequential FP32 sum over dims 0..127, one rounding per subtraction,
// product and addition (no contraction).
__device__ inline float seq_dist(const void* q, int qt, size_t qrow, const void* r, int rt, size_t rrow) {
#pragma clang fp contract(off)
    float d = 0.f;
    for (int c = 0; c < 16; c++) {
        float a[8], b[8];
        load8_src(q, qt, qrow, c, a);
        load8_src(r, rt, rrow, c, b);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float e = a[i] - b[i];
            d = d + e * e;
        }
    }
    return d;
}

}  // namespace sift
