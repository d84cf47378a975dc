// mlp.hip — learned hash: a 128 -> 256 (ReLU) -> b MLP whose output signs are
// the hash bits (bit k set when z_k < 0, the sign convention of the
// hyperplane codes, src/lsh_h.cu:302-309).  No counterpart in the reference.
//
// Forward (training and inference): one wave per 32 input rows, both layers
// on BF16 MFMAs with FP32 accumulation.  Layer 1 computes H^T = W1 X^T so its
// accumulator holds, per lane (= input row), 16 hidden units; those registers
// are used directly as the B operand of layer 2 (Z^T = W2 H^T) after BF16
// rounding, with W2's columns packed in the matching order, so no transpose
// or LDS round trip is needed.
//
// Training: sign binarisation with a straight-through gradient (dL/dz =
// dL/ds: the sign passes the gradient unchanged, so saturated outputs still
// feel the balance terms).  Loss over a batch of B positive pairs (x_i, y_i),
// N = 2B coded rows, s = sign(z) in {-1, +1}:
//   L = sum_ik (s_x - s_y)^2 / (4 B b)            Hamming fraction of pairs
//     + alpha * sum_k mean_k(s)^2 / b             balanced bits
//     + beta  * sum_{k != l} C_kl^2 / b^2         decorrelated bits, C = S^T S / N
//     + gamma * sum_ik (|z| - 1)^2 / (N b)        outputs near +-1
// Positive pairs come from the exact 2-NN of each training row, computed by
// the fused brute-force kernel.  Backward and Adam are plain FP32 kernels
// (the matrices are small); weights are FP32 masters with BF16 copies packed
// as MFMA operands after every step.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fused.h"
#include "internal.h"

using namespace sift;

namespace {

constexpr int HID = 256;
constexpr int BMAX = 32;
// flat parameter layout: W1 [HID][128], b1 [HID], W2 [BMAX][HID], b2 [BMAX]
constexpr int OFF_W1 = 0, OFF_B1 = HID * 128, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + BMAX * HID;
constexpr int NPARAM = OFF_B2 + BMAX;
// Layer 1 uses W1 = w / 64 for stored weights w: inputs are in [0, 255], and
// this keeps the stored weights near unit scale, where Adam's step size is
// meaningful.  A power of two, so the BF16 copy is exact scaling.
constexpr float W1_SCALE = 1.f / 64.f;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ inline uint32_t f2bf_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ inline float bf_round(float f) { return __uint_as_float(f2bf_rne(f) << 16); }

// Accumulator register j of lane half h holds row 8 (j / 4) + 4 h + j % 4.
__device__ inline int acc_row(int j, int h) { return 8 * (j >> 2) + 4 * h + (j & 3); }

// ------------------------------------------------------------- forward ----
struct FwdArgs {
    const uint4* x_tiles;  // packed inputs (BF16 tile layout, tg groups per tile)
    int tg;
    uint32_t n_rows, n_tiles;
    const uint4* w1p;  // [8 hidden blocks][8 groups][64]
    const float* b1;   // [HID]
    const uint4* w2p;  // [8 hidden blocks][2][64]
    const float* b2;   // [BMAX]
    int nbits;
    float* h_out;  // [n][HID] BF16-rounded ReLU outputs (training) or null
    float* z_out;  // [n][BMAX] or null
    uint32_t* codes;
    uint32_t stride;
    int table, combine;
};

__device__ __forceinline__ void mlp_fwd_body(const FwdArgs& a) {
    const int lane = threadIdx.x & 63, h = lane >> 5;
    const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= a.n_tiles) return;  // whole wave
    const uint32_t row = tile * 32 + (uint32_t)(lane & 31);
    uint4 xa[8];
#pragma unroll
    for (int g = 0; g < 8; g++) xa[g] = a.x_tiles[((size_t)tile * a.tg + g) * 64 + lane];
    f32x16 z;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int k = acc_row(j, h);
        z[j] = k < a.nbits ? a.b2[k] : 0.f;
    }
    for (int hb = 0; hb < HID / 32; hb++) {
        f32x16 acc;
#pragma unroll
        for (int j = 0; j < 16; j++) acc[j] = a.b1[32 * hb + acc_row(j, h)];
#pragma unroll
        for (int g = 0; g < 8; g++)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                __builtin_bit_cast(bf16x8_t, a.w1p[((size_t)hb * 8 + g) * 64 + lane]),
                __builtin_bit_cast(bf16x8_t, xa[g]), acc, 0, 0, 0);
        uint32_t hw[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const float v = fmaxf(acc[j], 0.f);
            hw[j] = f2bf_rne(v);
            if (a.h_out && row < a.n_rows) a.h_out[(size_t)row * HID + 32 * hb + acc_row(j, h)] = __uint_as_float(hw[j] << 16);
        }
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const uint4 frag = make_uint4(hw[8 * m] | (hw[8 * m + 1] << 16), hw[8 * m + 2] | (hw[8 * m + 3] << 16),
                                          hw[8 * m + 4] | (hw[8 * m + 5] << 16), hw[8 * m + 6] | (hw[8 * m + 7] << 16));
            z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                __builtin_bit_cast(bf16x8_t, a.w2p[((size_t)hb * 2 + m) * 64 + lane]),
                __builtin_bit_cast(bf16x8_t, frag), z, 0, 0, 0);
        }
    }
    if (a.z_out && row < a.n_rows) {
#pragma unroll
        for (int j = 0; j < 16; j++) a.z_out[(size_t)row * BMAX + acc_row(j, h)] = z[j];
    }
    if (a.codes) {
        // this lane's 16 bits of its row's code; the other half holds the rest
        uint32_t code = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) code |= (z[j] < 0.f ? 1u : 0u) << acc_row(j, h);
        code |= __shfl_xor(code, 32);
        // bits at or above nbits have z = 0 exactly, so they are clear
        if (lane < 32 && row < a.n_rows)
            a.codes[(size_t)a.table * a.stride + row] =
                (a.combine && a.nbits < 32) ? (((uint32_t)a.table << a.nbits) | code) : code;
    }
}

__global__ __launch_bounds__(256) void mlp_fwd(FwdArgs a) { mlp_fwd_body(a); }

// Several networks in one launch (one per grid row y): the learned index's
// tables hash a small query batch with one launch instead of one each.
constexpr int FWD_NETS = 16;
struct NetPtrs {
    const uint4* w1p;
    const float* b1;
    const uint4* w2p;
    const float* b2;
};
struct FwdMulti {
    FwdArgs a;  // common fields; w1p/b1/w2p/b2/table/z_out taken per network
    NetPtrs nets[FWD_NETS];
    int t0;  // table of grid row 0
};

__global__ __launch_bounds__(256) void mlp_fwd_multi(FwdMulti m) {
    FwdArgs a = m.a;
    const int y = blockIdx.y;
    a.w1p = m.nets[y].w1p;
    a.b1 = m.nets[y].b1;
    a.w2p = m.nets[y].w2p;
    a.b2 = m.nets[y].b2;
    a.table = m.t0 + y;
    if (a.z_out) a.z_out += (size_t)a.table * a.stride * BMAX;
    mlp_fwd_body(a);
}

// W1 [HID][128] -> BF16 A operands: element (u, d) at tile u/32, group d/16,
// lane u%32 + 32 ((d%16)/8), position d%8.  W2 [BMAX][HID] -> A operands of
// layer 2: hidden unit v of block hb is register j = 8 m + 4 a + c of lane
// half hh with v = 16 m + 8 a + 4 hh + c, i.e. position 4 a + c of the
// fragment (hb, m) in lane k + 32 hh.  Also writes the BF16-rounded FP32 W2
// used by the backward pass.
__global__ void mlp_pack_weights(const float* __restrict__ p, uint16_t* __restrict__ w1p, uint16_t* __restrict__ w2p,
                                 float* __restrict__ w2q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < HID * 128) {
        const int u = i / 128, d = i % 128;
        const int t = u / 32, r = u % 32, g = d / 16, hh = (d % 16) / 8, pos = d % 8;
        w1p[(((size_t)t * 8 + g) * 64 + r + 32 * hh) * 8 + pos] = (uint16_t)f2bf_rne(p[OFF_W1 + i] * W1_SCALE);
    }
    if (i < BMAX * HID) {
        const int k = i / HID, u = i % HID;
        const int hb = u / 32, v = u % 32;
        const int m = v / 16, aa = (v % 16) / 8, hh = (v % 8) / 4, c = v % 4;
        const uint32_t q = f2bf_rne(p[OFF_W2 + i]);
        w2p[(((size_t)hb * 2 + m) * 64 + k + 32 * hh) * 8 + 4 * aa + c] = (uint16_t)q;
        w2q[i] = __uint_as_float(q << 16);
    }
}

// ---------------------------------------------------------------- loss ----
// stats[0..BMAX): sum of s per bit; [BMAX .. BMAX + BMAX^2): sum of s_k s_l;
// [BMAX + BMAX^2]: sum of (s_x - s_y)^2 / 4; [+1]: sum of (|z| - 1)^2.
constexpr int ST_SUM = 0, ST_CROSS = BMAX, ST_SIM = BMAX + BMAX * BMAX, ST_Q = ST_SIM + 1, ST_N = ST_Q + 1;

__device__ inline float sgn(float z) { return z >= 0.f ? 1.f : -1.f; }

__global__ __launch_bounds__(256) void mlp_loss_stats(const float* __restrict__ z, uint32_t B, int nbits,
                                                      float* __restrict__ stats) {
    __shared__ float s_sum[BMAX], s_cross[BMAX * BMAX], s_sim, s_q;
    for (int i = threadIdx.x; i < BMAX * BMAX; i += blockDim.x) s_cross[i] = 0.f;
    if (threadIdx.x < BMAX) s_sum[threadIdx.x] = 0.f;
    if (threadIdx.x == 0) s_sim = s_q = 0.f;
    __syncthreads();
    const uint32_t N = 2 * B;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < N) {
        float s[BMAX];
        float sim = 0.f, qq = 0.f;
        for (int k = 0; k < nbits; k++) {
            const float zv = z[(size_t)r * BMAX + k];
            s[k] = sgn(zv);
            qq += (fabsf(zv) - 1.f) * (fabsf(zv) - 1.f);
            if (r < B) {
                const float d = s[k] - sgn(z[(size_t)(r + B) * BMAX + k]);
                sim += d * d * 0.25f;
            }
            atomicAdd(&s_sum[k], s[k]);
        }
        for (int k = 0; k < nbits; k++)
            for (int l = 0; l < nbits; l++) atomicAdd(&s_cross[k * BMAX + l], s[k] * s[l]);
        atomicAdd(&s_sim, sim);
        atomicAdd(&s_q, qq);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < BMAX * BMAX; i += blockDim.x) atomicAdd(&stats[ST_CROSS + i], s_cross[i]);
    if (threadIdx.x < BMAX) atomicAdd(&stats[ST_SUM + threadIdx.x], s_sum[threadIdx.x]);
    if (threadIdx.x == 0) {
        atomicAdd(&stats[ST_SIM], s_sim);
        atomicAdd(&stats[ST_Q], s_q);
    }
}

struct LossW {
    float alpha, beta, gamma;
};

__global__ __launch_bounds__(256) void mlp_loss_grad(const float* __restrict__ z, uint32_t B, int nbits,
                                                     const float* __restrict__ stats, LossW w,
                                                     float* __restrict__ dz) {
    const uint32_t N = 2 * B;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * BMAX) return;
    const uint32_t r = i / BMAX;
    const int k = (int)(i % BMAX);
    float g = 0.f;
    if (k < nbits) {
        const float b = (float)nbits, n = (float)N;
        const float zv = z[i];
        const float s = sgn(zv);
        const uint32_t partner = r < B ? r + B : r - B;
        const float sp = sgn(z[(size_t)partner * BMAX + k]);
        float dls = (s - sp) / (2.f * (float)B * b);  // similarity
        dls += w.alpha * 2.f * (stats[ST_SUM + k] / n) / (b * n);  // balance
        float dec = 0.f;
        for (int l = 0; l < nbits; l++)
            if (l != k) dec += (stats[ST_CROSS + k * BMAX + l] / n) * sgn(z[(size_t)r * BMAX + l]);
        dls += w.beta * 4.f * dec / (b * b * n);  // decorrelation
        g = dls;  // straight-through: sign passes the gradient unchanged
        g += w.gamma * 2.f * (fabsf(zv) - 1.f) * s / (n * b);
    }
    dz[i] = g;
}

// dH[r][u] = (sum_k dz[r][k] W2q[k][u]) where H[r][u] > 0.
__global__ __launch_bounds__(256) void mlp_back_h(const float* __restrict__ dz, const float* __restrict__ w2q,
                                                  const float* __restrict__ hbuf, uint32_t N, int nbits,
                                                  float* __restrict__ dh) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * HID) return;
    const uint32_t r = i / HID, u = i % HID;
    float s = 0.f;
    for (int k = 0; k < nbits; k++) s += dz[(size_t)r * BMAX + k] * w2q[(size_t)k * HID + u];
    dh[i] = hbuf[i] > 0.f ? s : 0.f;
}

// C[m][c] += sum_r A[r][m] * B[r][c] over a chunk of rows (and, with
// bias != null, bias[m] += sum_r A[r][m]); 32 x 32 output tile per block.
__global__ __launch_bounds__(256) void gemm_atb(const float* __restrict__ A, int lda, int M,
                                                const float* __restrict__ Bm, int ldb, int Nc, uint32_t R,
                                                uint32_t rows_per_chunk, float* __restrict__ C, int ldc,
                                                float* __restrict__ bias) {
    __shared__ float sa[32][33], sb[32][33];
    const int m0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const uint32_t r_begin = blockIdx.z * rows_per_chunk;
    const uint32_t r_end = min(R, r_begin + rows_per_chunk);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // ty 0..7: 4 outputs each
    float acc[4] = {0.f, 0.f, 0.f, 0.f}, bacc = 0.f;
    for (uint32_t r0 = r_begin; r0 < r_end; r0 += 32) {
        for (int k = ty; k < 32; k += 8) {
            const uint32_t r = r0 + k;
            sa[k][tx] = (r < r_end && m0 + tx < M) ? A[(size_t)r * lda + m0 + tx] : 0.f;
            sb[k][tx] = (r < r_end && c0 + tx < Nc) ? Bm[(size_t)r * ldb + c0 + tx] : 0.f;
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < 32; k++) {
            const float bv = sb[k][tx];
#pragma unroll
            for (int e = 0; e < 4; e++) acc[e] += sa[k][ty * 4 + e] * bv;
            if (bias && blockIdx.y == 0 && ty == 0) bacc += sa[k][tx];
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int m = m0 + ty * 4 + e, c = c0 + tx;
        if (m < M && c < Nc) atomicAdd(&C[(size_t)m * ldc + c], acc[e]);
    }
    if (bias && blockIdx.y == 0 && ty == 0 && m0 + tx < M) atomicAdd(&bias[m0 + tx], bacc);
}

// Rows of the packed input, unpacked to FP32 [N][128] for the weight gradient.
__global__ void mlp_unpack_x(const uint4* __restrict__ x_tiles, int tg, uint32_t N, float* __restrict__ x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * 16) return;
    const uint32_t r = i / 16, blk = i % 16;  // dims 8 blk .. +7
    const uint32_t t = r / 32, l = (r % 32) + 32 * (blk & 1);
    const uint4 v = x_tiles[((size_t)t * tg + blk / 2) * 64 + l];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; e++) {  // times W1_SCALE: the gradient of the stored weights
        x[(size_t)r * 128 + 8 * blk + 2 * e] = __uint_as_float(w[e] << 16) * W1_SCALE;
        x[(size_t)r * 128 + 8 * blk + 2 * e + 1] = __uint_as_float(w[e] & 0xFFFF0000u) * W1_SCALE;
    }
}

__global__ void mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                         float* __restrict__ m2, int n, float lr, float b1, float b2, float c1, float c2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

__global__ void fill_f32(float* p, size_t n, float v) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ----------------------------------------------------------------- host ----
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u = uniform(), v = uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

}  // namespace

struct sift_mlp {
    sift_context* h = nullptr;
    int nbits = 0;
    DevBuf params, grads, adam_m, adam_v, w1p, w2p, w2q;
    long steps = 0;
    // training scratch
    DevBuf rowmap, hbuf, zbuf, dz, dh, xf, stats, src_u8;
    PackedSet xt;
};

namespace sift {

// Pack the current FP32 parameters into the MFMA operand copies.
int mlp_repack(sift_mlp* m) {
    sift_context* h = m->h;
    mlp_pack_weights<<<(HID * 128 + 255) / 256, 256, 0, h->stream>>>(
        static_cast<const float*>(m->params.p), static_cast<uint16_t*>(m->w1p.p), static_cast<uint16_t*>(m->w2p.p),
        static_cast<float*>(m->w2q.p));
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

// Forward over packed inputs (tg groups per tile) into z / hidden / codes.
int mlp_forward(sift_mlp* m, const PackedSet& xs, uint32_t n, float* h_out, float* z_out, uint32_t* codes,
                uint32_t stride, int table, bool combine) {
    sift_context* h = m->h;
    if (n == 0) return SIFT_OK;
    FwdArgs a;
    a.x_tiles = static_cast<const uint4*>(xs.tiles.p);
    a.tg = groups_per_tile(xs.kind);
    a.n_rows = n;
    a.n_tiles = (n + 31) / 32;
    a.w1p = static_cast<const uint4*>(m->w1p.p);
    a.b1 = static_cast<const float*>(m->params.p) + OFF_B1;
    a.w2p = static_cast<const uint4*>(m->w2p.p);
    a.b2 = static_cast<const float*>(m->params.p) + OFF_B2;
    a.nbits = m->nbits;
    a.h_out = h_out;
    a.z_out = z_out;
    a.codes = codes;
    a.stride = stride;
    a.table = table;
    a.combine = combine ? 1 : 0;
    mlp_fwd<<<(a.n_tiles + 3) / 4, 256, 0, h->stream>>>(a);
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

// Codes (and z, [table][stride][BMAX]) of n rows under L networks of equal
// width, FWD_NETS networks per launch.
int mlp_forward_tables(sift_mlp* const* ms, int L, const PackedSet& xs, uint32_t n, float* z_out,
                       uint32_t* codes, uint32_t stride, bool combine) {
    if (n == 0 || L == 0) return SIFT_OK;
    sift_context* h = ms[0]->h;
    FwdMulti fm{};
    fm.a.x_tiles = static_cast<const uint4*>(xs.tiles.p);
    fm.a.tg = groups_per_tile(xs.kind);
    fm.a.n_rows = n;
    fm.a.n_tiles = (n + 31) / 32;
    fm.a.nbits = ms[0]->nbits;
    fm.a.h_out = nullptr;
    fm.a.z_out = z_out;
    fm.a.codes = codes;
    fm.a.stride = stride;
    fm.a.combine = combine ? 1 : 0;
    for (int t0 = 0; t0 < L; t0 += FWD_NETS) {
        const int nt = std::min(FWD_NETS, L - t0);
        for (int k = 0; k < nt; k++) {
            const sift_mlp* m = ms[t0 + k];
            if (m->nbits != fm.a.nbits) return set_error(h, SIFT_ERR_INVALID, "networks differ in bits");
            fm.nets[k] = NetPtrs{static_cast<const uint4*>(m->w1p.p), static_cast<const float*>(m->params.p) + OFF_B1,
                                 static_cast<const uint4*>(m->w2p.p), static_cast<const float*>(m->params.p) + OFF_B2};
        }
        fm.t0 = t0;
        mlp_fwd_multi<<<dim3((fm.a.n_tiles + 3) / 4, (unsigned)nt), 256, 0, h->stream>>>(fm);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    return SIFT_OK;
}

int mlp_nbits(const sift_mlp* m) { return m->nbits; }

}  // namespace sift

namespace {

int fill(sift_context* h, DevBuf& b, size_t n, float v = 0.f) {
    if (n) fill_f32<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(static_cast<float*>(b.p), n, v);
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

// Loss and gradient of one batch: rows [0, B) of the packed set are x_i,
// rows [B, 2B) the positives y_i.  Leaves dL/dparams in m->grads and the
// batch statistics in m->stats.
int batch_grad(sift_mlp* m, uint32_t B, const LossW& w) {
    sift_context* h = m->h;
    const uint32_t N = 2 * B;
    auto* hb = static_cast<float*>(m->hbuf.p);
    auto* zb = static_cast<float*>(m->zbuf.p);
    SIFT_TRY(fill(h, m->stats, ST_N));
    SIFT_TRY(fill(h, m->grads, NPARAM));
    SIFT_TRY(mlp_forward(m, m->xt, N, hb, zb, nullptr, 0, 0, false));
    mlp_loss_stats<<<(N + 255) / 256, 256, 0, h->stream>>>(zb, B, m->nbits, static_cast<float*>(m->stats.p));
    mlp_loss_grad<<<(N * BMAX + 255) / 256, 256, 0, h->stream>>>(zb, B, m->nbits, static_cast<float*>(m->stats.p),
                                                                w, static_cast<float*>(m->dz.p));
    mlp_back_h<<<(N * HID + 255) / 256, 256, 0, h->stream>>>(static_cast<float*>(m->dz.p),
                                                            static_cast<float*>(m->w2q.p), hb, N, m->nbits,
                                                            static_cast<float*>(m->dh.p));
    mlp_unpack_x<<<(N * 16 + 255) / 256, 256, 0, h->stream>>>(static_cast<const uint4*>(m->xt.tiles.p),
                                                              groups_per_tile(m->xt.kind), N,
                                                              static_cast<float*>(m->xf.p));
    SIFT_HIP_TRY(h, hipGetLastError());
    auto* g = static_cast<float*>(m->grads.p);
    const uint32_t chunk = 1024;
    const unsigned nz = (N + chunk - 1) / chunk;
    // dW2 [BMAX][HID] = dz^T H, db2 = colsum dz
    gemm_atb<<<dim3(1, HID / 32, nz), 256, 0, h->stream>>>(static_cast<float*>(m->dz.p), BMAX, BMAX, hb, HID, HID, N,
                                                          chunk, g + OFF_W2, HID, g + OFF_B2);
    // dW1 [HID][128] = dH^T X, db1 = colsum dH
    gemm_atb<<<dim3(HID / 32, 128 / 32, nz), 256, 0, h->stream>>>(static_cast<float*>(m->dh.p), HID, HID,
                                                                  static_cast<float*>(m->xf.p), 128, 128, N, chunk,
                                                                  g + OFF_W1, 128, g + OFF_B1);
    SIFT_HIP_TRY(h, hipGetLastError());
    return SIFT_OK;
}

int reserve_train(sift_mlp* m, uint32_t B) {
    sift_context* h = m->h;
    const uint32_t N = 2 * B;
    const uint32_t rows_pad = ((N + 31) / 32) * 32;
    SIFT_TRY(reserve(h, m->rowmap, (size_t)rows_pad * 4));
    SIFT_TRY(reserve(h, m->hbuf, (size_t)N * HID * 4));
    SIFT_TRY(reserve(h, m->zbuf, (size_t)rows_pad * BMAX * 4));
    SIFT_TRY(reserve(h, m->dz, (size_t)N * BMAX * 4));
    SIFT_TRY(reserve(h, m->dh, (size_t)N * HID * 4));
    SIFT_TRY(reserve(h, m->xf, (size_t)N * 128 * 4));
    SIFT_TRY(reserve(h, m->stats, ST_N * 4));
    SIFT_TRY(reserve(h, m->xt.tiles, (size_t)(rows_pad / 32) * tile_bytes(KIND_BF16)));
    SIFT_TRY(reserve(h, m->xt.norm, (size_t)rows_pad * 4));
    m->xt.rows = N;
    m->xt.rows_pad = rows_pad;
    m->xt.ntiles = rows_pad / 32;
    m->xt.kind = KIND_BF16;
    return SIFT_OK;
}

// Pack the batch rows src[map[i]] (uint8 [*][128]) into m->xt.
int pack_batch(sift_mlp* m, const std::vector<int32_t>& map) {
    sift_context* h = m->h;
    std::vector<int32_t> mp(map);
    mp.resize(m->xt.rows_pad, -1);
    SIFT_HIP_TRY(h, hipMemcpyAsync(m->rowmap.p, mp.data(), mp.size() * 4, hipMemcpyHostToDevice, h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));  // mp is a host temporary
    int* flag = static_cast<int*>(h->flag.p);
    return pack_launch(h, m->src_u8.p, SIFT_U8_DEVICE, 0, static_cast<const int32_t*>(m->rowmap.p), KIND_BF16,
                       PACK_PLAIN, m->xt, flag);
}

void read_loss(sift_mlp* m, uint32_t B, const LossW& w, float* out) {
    std::vector<float> st(ST_N);
    (void)hipStreamSynchronize(m->h->stream);  // the statistics come from the handle's stream
    (void)hipMemcpy(st.data(), m->stats.p, ST_N * 4, hipMemcpyDeviceToHost);
    const double b = m->nbits, n = 2.0 * B;
    double bal = 0, dec = 0;
    for (int k = 0; k < m->nbits; k++) {
        const double mk = st[ST_SUM + k] / n;
        bal += mk * mk;
        for (int l = 0; l < m->nbits; l++)
            if (l != k) {
                const double c = st[ST_CROSS + k * BMAX + l] / n;
                dec += c * c;
            }
    }
    out[0] = (float)(st[ST_SIM] / (B * b) + w.alpha * bal / b + w.beta * dec / (b * b) + w.gamma * st[ST_Q] / (n * b));
    out[1] = (float)(st[ST_SIM] / (B * b));  // Hamming fraction of the positive pairs
}

int stage_u8(sift_context* h, sift_mlp* m, const void* rows, int type, uint32_t n, uint32_t at, DevBuf& nrm_tmp) {
    // uint8 copy of n rows into src_u8 starting at row `at`
    DevBuf tmp;
    int* flag = static_cast<int*>(h->flag.p);
    SIFT_TRY(to_u8(h, rows, type, n, tmp, nrm_tmp, flag));
    if (n)
        SIFT_HIP_TRY(h, hipMemcpyAsync(static_cast<uint8_t*>(m->src_u8.p) + (size_t)at * 128, tmp.p, (size_t)n * 128,
                                       hipMemcpyDeviceToDevice, h->stream));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    release(tmp);
    return SIFT_OK;
}

}  // namespace

extern "C" {

int sift_mlp_create(sift_handle_t h, int nbits, uint64_t seed, sift_mlp_t* out) {
    if (!h || !out) return set_error(h, SIFT_ERR_INVALID, "null handle/mlp");
    *out = nullptr;
    if (nbits < 1 || nbits > BMAX) return set_error(h, SIFT_ERR_INVALID, "nbits must be in [1,32]");
    auto* m = new sift_mlp();
    m->h = h;
    m->nbits = nbits;
    // He-style init scaled for inputs in [0,255]; rows >= nbits of W2/b2 stay zero
    std::vector<float> p(NPARAM, 0.f);
    SplitMix r{seed ^ 0x6c6561726e656400ull};
    for (int i = 0; i < HID * 128; i++) p[OFF_W1 + i] = (float)(r.normal() * std::sqrt(2.0 / 128.0));
    for (int i = 0; i < HID; i++) p[OFF_B1 + i] = (float)(r.normal() * 0.1);
    for (int k = 0; k < nbits; k++)
        for (int u = 0; u < HID; u++) p[OFF_W2 + k * HID + u] = (float)(r.normal() * std::sqrt(1.0 / HID));
    int rc = SIFT_OK;
    do {
        for (DevBuf* b : {&m->params, &m->grads, &m->adam_m, &m->adam_v})
            if ((rc = reserve(h, *b, NPARAM * 4)) != SIFT_OK) break;
        if (rc != SIFT_OK) break;
        if ((rc = reserve(h, m->w1p, HID * 128 * 2)) != SIFT_OK) break;
        if ((rc = reserve(h, m->w2p, BMAX * HID * 2)) != SIFT_OK) break;
        if ((rc = reserve(h, m->w2q, BMAX * HID * 4)) != SIFT_OK) break;
        if ((rc = reserve(h, h->flag, 16)) != SIFT_OK) break;
        if (hipMemcpy(m->params.p, p.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess) {
            rc = set_error(h, SIFT_ERR_HIP, "copy params");
            break;
        }
        if ((rc = fill(h, m->adam_m, NPARAM)) != SIFT_OK) break;
        if ((rc = fill(h, m->adam_v, NPARAM)) != SIFT_OK) break;
        if ((rc = mlp_repack(m)) != SIFT_OK) break;
        if (hipStreamSynchronize(h->stream) != hipSuccess) rc = set_error(h, SIFT_ERR_HIP, "sync");
    } while (0);
    if (rc != SIFT_OK) {
        sift_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return SIFT_OK;
}

int sift_mlp_destroy(sift_mlp_t m) {
    if (!m) return SIFT_OK;
    if (m->h) (void)hipStreamSynchronize(m->h->stream);
    for (DevBuf* b : {&m->params, &m->grads, &m->adam_m, &m->adam_v, &m->w1p, &m->w2p, &m->w2q, &m->rowmap, &m->hbuf,
                      &m->zbuf, &m->dz, &m->dh, &m->xf, &m->stats, &m->src_u8, &m->xt.tiles, &m->xt.norm,
                      &m->xt.norm_lane})
        release(*b);
    delete m;
    return SIFT_OK;
}

int sift_mlp_get_params(sift_mlp_t m, float* out) {
    if (!m || !out) return SIFT_ERR_INVALID;
    SIFT_HIP_TRY(m->h, hipStreamSynchronize(m->h->stream));
    if (hipMemcpy(out, m->params.p, NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(m->h, SIFT_ERR_HIP, "copy params");
    return SIFT_OK;
}

int sift_mlp_set_params(sift_mlp_t m, const float* in) {
    if (!m || !in) return SIFT_ERR_INVALID;
    SIFT_HIP_TRY(m->h, hipStreamSynchronize(m->h->stream));  // no queued forward still reads them
    if (hipMemcpy(m->params.p, in, NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(m->h, SIFT_ERR_HIP, "copy params");
    SIFT_TRY(mlp_repack(m));
    SIFT_HIP_TRY(m->h, hipStreamSynchronize(m->h->stream));
    return SIFT_OK;
}

int sift_mlp_codes(sift_handle_t h, sift_mlp_t m, const void* x, int type, uint32_t n, uint32_t* codes,
                   float* z) {
    if (!h || !m) return set_error(h, SIFT_ERR_INVALID, "null handle/mlp");
    if (n == 0) return SIFT_OK;
    if (type == SIFT_FLOAT_HOST) SIFT_TRY(stage_host_input(h, x, n, h->stage_q, &x));
    SIFT_TRY(reserve(h, h->flag, 16));
    int* flag = static_cast<int*>(h->flag.p);
    const uint32_t rows_pad = ((n + 31) / 32) * 32;
    SIFT_TRY(pack_rows(h, x, type, n, rows_pad, KIND_BF16, PACK_PLAIN, h->q_set, flag));
    {
        ProfScope ps(h, PROF_MLP);
        SIFT_TRY(mlp_forward(m, h->q_set, n, nullptr, z, codes, n, 0, false));
    }
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SIFT_OK;
}

int sift_mlp_grad(sift_handle_t h, sift_mlp_t m, const void* x, const void* y, int type, uint32_t B,
                  const float* loss_w, float* grads, float* loss, float* z) {
    if (!h || !m || !loss_w) return set_error(h, SIFT_ERR_INVALID, "null argument");
    if (B == 0) return set_error(h, SIFT_ERR_INVALID, "empty batch");
    if (type == SIFT_FLOAT_HOST) return set_error(h, SIFT_ERR_INVALID, "sift_mlp_grad takes device rows");
    const LossW w{loss_w[0], loss_w[1], loss_w[2]};
    SIFT_TRY(reserve(h, h->flag, 16));
    SIFT_TRY(reserve(h, m->src_u8, (size_t)2 * B * 128));
    DevBuf nrm;
    SIFT_TRY(stage_u8(h, m, x, type, B, 0, nrm));
    SIFT_TRY(stage_u8(h, m, y, type, B, B, nrm));
    release(nrm);
    SIFT_TRY(reserve_train(m, B));
    std::vector<int32_t> map(2 * B);
    for (uint32_t i = 0; i < 2 * B; i++) map[i] = (int32_t)i;
    SIFT_TRY(pack_batch(m, map));
    SIFT_TRY(batch_grad(m, B, w));
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (grads) SIFT_HIP_TRY(h, hipMemcpy(grads, m->grads.p, NPARAM * 4, hipMemcpyDeviceToHost));
    if (z) SIFT_HIP_TRY(h, hipMemcpy(z, m->zbuf.p, (size_t)2 * B * BMAX * 4, hipMemcpyDeviceToHost));
    if (loss) read_loss(m, B, w, loss);
    return SIFT_OK;
}

int sift_mlp_train(sift_handle_t h, sift_mlp_t m, const void* x, uint32_t n, const void* base, uint32_t r_n,
                   int type, const sift_mlp_train_opts* o, float* epoch_loss) {
    if (!h || !m || !o) return set_error(h, SIFT_ERR_INVALID, "null argument");
    if (type == SIFT_FLOAT_HOST) return set_error(h, SIFT_ERR_INVALID, "sift_mlp_train takes device rows");
    if (n == 0 || r_n < 2) return set_error(h, SIFT_ERR_INVALID, "need training rows and a base of 2+ rows");
    if (o->batch < 32 || o->epochs < 1) return set_error(h, SIFT_ERR_INVALID, "batch >= 32 and epochs >= 1");
    const LossW w{o->alpha, o->beta, o->gamma};
    // exact 2-NN labels from the brute-force kernel
    DevBuf lab;
    SIFT_TRY(reserve(h, lab, (size_t)n * 16));
    SIFT_TRY(sift_knn2(h, x, base, type, n, r_n, SIFT_COMPUTE_AUTO, static_cast<uint32_t*>(lab.p),
                       reinterpret_cast<float*>(static_cast<uint32_t*>(lab.p) + 2 * (size_t)n), nullptr, nullptr,
                       0.8f));
    std::vector<uint32_t> nn((size_t)n * 2);
    std::vector<float> dd((size_t)n * 2);
    if (hipMemcpy(nn.data(), lab.p, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(dd.data(), static_cast<uint32_t*>(lab.p) + 2 * (size_t)n, (size_t)n * 8, hipMemcpyDeviceToHost) !=
            hipSuccess) {
        release(lab);
        return set_error(h, SIFT_ERR_HIP, "copy labels");
    }
    release(lab);
    // positives: (x, 1st NN) when it passes the ratio test sqrt(d1 / d2) <
    // max_ratio (a distinctive match), or every (x, 1st) and (x, 2nd) when
    // max_ratio is 0.  Pairs of a row with an arbitrary far "nearest" row would
    // reward putting everything in one bucket.
    for (uint32_t i = 0; i < n; i++) {
        if (o->max_ratio > 0.f) {
            const bool keep = nn[2 * i] < r_n && dd[2 * i + 1] > 0.f &&
                              std::sqrt(dd[2 * i]) / std::sqrt(dd[2 * i + 1]) < o->max_ratio;
            if (!keep) nn[2 * i] = 0xFFFFFFFFu;
            nn[2 * i + 1] = 0xFFFFFFFFu;
        }
    }
    // source rows: training rows, then the base rows that are labels
    std::vector<uint32_t> used;
    used.reserve((size_t)n * 2);
    for (uint32_t v : nn)
        if (v < r_n) used.push_back(v);
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    const uint32_t nu = (uint32_t)used.size();
    SIFT_TRY(reserve(h, h->flag, 16));
    SIFT_TRY(reserve(h, m->src_u8, ((size_t)n + nu) * 128));
    DevBuf nrm;
    SIFT_TRY(stage_u8(h, m, x, type, n, 0, nrm));
    {   // the base rows that are labels, in runs of consecutive rows, after the training rows
        DevBuf tmp;
        int rc = to_u8(h, base, type, r_n, tmp, nrm, static_cast<int*>(h->flag.p));
        uint32_t i = 0;
        while (rc == SIFT_OK && i < nu) {
            uint32_t j = i + 1;
            while (j < nu && used[j] == used[j - 1] + 1) j++;
            if (hipMemcpyAsync(static_cast<uint8_t*>(m->src_u8.p) + ((size_t)n + i) * 128,
                               static_cast<uint8_t*>(tmp.p) + (size_t)used[i] * 128, (size_t)(j - i) * 128,
                               hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
                rc = set_error(h, SIFT_ERR_HIP, "gather label rows");
            i = j;
        }
        if (rc == SIFT_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = set_error(h, SIFT_ERR_HIP, "sync");
        release(tmp);
        release(nrm);
        SIFT_TRY(rc);
    }
    // pairs (x_i, base[nn_k(i)]) for k = 1, 2
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    pairs.reserve((size_t)n * 2);
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 2; k++) {
            const uint32_t v = nn[2 * (size_t)i + k];
            if (v >= r_n) continue;
            const uint32_t pos = (uint32_t)(std::lower_bound(used.begin(), used.end(), v) - used.begin());
            pairs.push_back({i, n + pos});
        }
    const uint32_t B = (uint32_t)std::min<size_t>(o->batch, pairs.size());
    if (B == 0) return set_error(h, SIFT_ERR_INVALID, "no labelled pairs");
    SIFT_TRY(reserve_train(m, B));
    if (m->steps == 0) {
        // start balanced: shift each output bias by the median of its output
        // over the first training rows, so every bit splits them in half
        const uint32_t ns = std::min<uint32_t>(n, 2 * B);
        std::vector<int32_t> mp(ns);
        for (uint32_t i = 0; i < ns; i++) mp[i] = (int32_t)i;
        std::vector<int32_t> full(mp);
        full.resize(2 * (size_t)B, -1);
        SIFT_TRY(pack_batch(m, full));
        SIFT_TRY(mlp_forward(m, m->xt, ns, nullptr, static_cast<float*>(m->zbuf.p), nullptr, 0, 0, false));
        std::vector<float> z((size_t)ns * BMAX), p(NPARAM);
        SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
        SIFT_HIP_TRY(h, hipMemcpy(z.data(), m->zbuf.p, z.size() * 4, hipMemcpyDeviceToHost));
        SIFT_HIP_TRY(h, hipMemcpy(p.data(), m->params.p, NPARAM * 4, hipMemcpyDeviceToHost));
        std::vector<float> col(ns);
        for (int k = 0; k < m->nbits; k++) {
            for (uint32_t i = 0; i < ns; i++) col[i] = z[(size_t)i * BMAX + k];
            std::nth_element(col.begin(), col.begin() + ns / 2, col.end());
            p[OFF_B2 + k] -= col[ns / 2];
        }
        SIFT_HIP_TRY(h, hipMemcpy(m->params.p, p.data(), NPARAM * 4, hipMemcpyHostToDevice));
        SIFT_TRY(mlp_repack(m));
    }
    SplitMix rng{o->seed ^ 0x747261696e000000ull};
    std::vector<int32_t> map(2 * (size_t)B);
    const int total = (int)NPARAM;
    for (int ep = 0; ep < o->epochs; ep++) {
        for (size_t i = pairs.size(); i > 1; i--) std::swap(pairs[i - 1], pairs[rng.next() % i]);
        double ep_loss = 0;
        int nb = 0;
        for (size_t s0 = 0; s0 + B <= pairs.size(); s0 += B) {
            for (uint32_t i = 0; i < B; i++) {
                map[i] = (int32_t)pairs[s0 + i].first;
                map[B + i] = (int32_t)pairs[s0 + i].second;
            }
            SIFT_TRY(pack_batch(m, map));
            SIFT_TRY(batch_grad(m, B, w));
            m->steps++;
            const float c1 = 1.f - std::pow(0.9f, (float)m->steps), c2 = 1.f - std::pow(0.999f, (float)m->steps);
            mlp_adam<<<(total + 255) / 256, 256, 0, h->stream>>>(
                static_cast<float*>(m->params.p), static_cast<float*>(m->grads.p), static_cast<float*>(m->adam_m.p),
                static_cast<float*>(m->adam_v.p), total, o->lr, 0.9f, 0.999f, c1, c2, 1e-8f);
            SIFT_HIP_TRY(h, hipGetLastError());
            SIFT_TRY(mlp_repack(m));
            if (epoch_loss) {
                float l[2];
                read_loss(m, B, w, l);
                ep_loss += l[0];
            }
            nb++;
        }
        if (epoch_loss) epoch_loss[ep] = nb ? (float)(ep_loss / nb) : 0.f;
    }
    SIFT_HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SIFT_OK;
}

}  // extern "C"
