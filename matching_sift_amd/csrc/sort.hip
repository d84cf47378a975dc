// sort.hip — stable LSD radix sort of (uint32 key, uint32 value) pairs and
// an exclusive scan, for the LSH bucket build (HBM-bound).
//
// Replaces the reference's thrust::sort with an indirect comparator
// (src/lsh_h.cu:587-591) followed by a host run-length pass (:598-620).  Here
// every table is sorted on the device, each in its own segment of the pair
// arrays, in ceil(bits/8) passes of:
//   hist    : per-block 256-bin digit histogram (LDS atomics), bin-major
//   scan    : exclusive scan of the histogram -> global digit offsets (up to
//             2^20 counts: one launch scans 4096-count blocks and the scatter
//             adds the carries between blocks)
//   scatter : stable scatter; ranks among equal digits come from wave-wide
//             ballots, so elements keep their input order within a digit;
//             each chunk is ordered by digit in LDS and written run by run.
#include "internal.h"

namespace sift {
namespace {

constexpr int RS_T = 512;            // threads per block (at least one per digit)
constexpr int RS_I = 16;             // keys per thread
constexpr int RS_CHUNK = RS_T * RS_I;  // keys per block
constexpr int RS_W = RS_T / 64;      // waves per block
constexpr int RS_SUB = RS_CHUNK / RS_W;  // keys per wave, walked in order
constexpr int RS_STEPS = RS_SUB / 64;  // keys per lane
static_assert(RS_T >= 256, "one thread per digit");

// The pairs form nseg segments of seg_n consecutive elements, each sorted on
// its own (segment s keeps positions [s seg_n, (s + 1) seg_n)); blocks never
// straddle segments: block blk covers chunk blk % nbt of segment blk / nbt,
// nbt = ceil(seg_n / RS_CHUNK).  Digit counts are laid out [s][digit][chunk],
// so one exclusive scan over them yields each (segment, digit, chunk)'s
// output start, inside its segment.
struct SegChunk {
    size_t base;   // first element
    uint32_t m;    // elements in the chunk
    uint32_t hrow; // s * 256 * nbt + chunk: add digit * nbt for the count slot
};

__device__ inline SegChunk seg_chunk(uint32_t blk, uint32_t seg_n, uint32_t nbt) {
    const uint32_t s = blk / nbt, j = blk % nbt;
    SegChunk c;
    c.base = (size_t)s * seg_n + (size_t)j * RS_CHUNK;
    c.m = min((uint32_t)RS_CHUNK, seg_n - j * (uint32_t)RS_CHUNK);
    c.hrow = s * 256u * nbt + j;
    return c;
}

// Per-wave 256-bin counts in LDS (less contention than one shared set), keys
// read 16 bytes per lane where the chunk is whole and 16-byte aligned.
__global__ __launch_bounds__(RS_T) void rs_hist(const uint32_t* __restrict__ keys, uint32_t seg_n, uint32_t nbt,
                                                int shift, uint32_t dmask, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[RS_W][256];
    for (int i = threadIdx.x; i < RS_W * 256; i += RS_T) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const SegChunk c = seg_chunk(blockIdx.x, seg_n, nbt);
    uint32_t* wc = cnt[threadIdx.x >> 6];
    if (c.m == (uint32_t)RS_CHUNK && (c.base & 3) == 0) {
        const uint4* k4 = reinterpret_cast<const uint4*>(keys + c.base);
        uint4 v[RS_I / 4];
#pragma unroll
        for (int i = 0; i < RS_I / 4; i++) v[i] = k4[i * RS_T + threadIdx.x];
#pragma unroll
        for (int i = 0; i < RS_I / 4; i++) {
            atomicAdd(&wc[(v[i].x >> shift) & dmask], 1u);
            atomicAdd(&wc[(v[i].y >> shift) & dmask], 1u);
            atomicAdd(&wc[(v[i].z >> shift) & dmask], 1u);
            atomicAdd(&wc[(v[i].w >> shift) & dmask], 1u);
        }
    } else {
#pragma unroll
        for (int i = 0; i < RS_I; i++) {
            const uint32_t idx = (uint32_t)i * RS_T + threadIdx.x;
            if (idx < c.m) atomicAdd(&wc[(keys[c.base + idx] >> shift) & dmask], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < RS_W; k++) t += cnt[k][threadIdx.x];
        hist[c.hrow + (size_t)threadIdx.x * nbt] = t;
    }
}

// Stable scatter of one chunk: ranks come from wave-wide ballots, the chunk
// is first ordered by digit in LDS, then written out with consecutive
// threads on consecutive positions of each digit's run (whole runs instead
// of scattered 4-byte writes).  vin == nullptr: the value of each element is
// its position in its segment.  offs: the digit counts scanned within
// 4096-entry blocks; sums (n_sums <= 256 block totals, in block order) carry
// them across blocks here, or null when offs is fully scanned.
constexpr int RS_SUMS_MAX = 256;
__global__ __launch_bounds__(RS_T) void rs_scatter(const uint32_t* __restrict__ kin,
                                                   const uint32_t* __restrict__ vin,
                                                   uint32_t* __restrict__ kout,
                                                   uint32_t* __restrict__ vout, uint32_t seg_n, uint32_t nbt,
                                                   int shift, int dbits, const uint32_t* __restrict__ offs,
                                                   const uint32_t* __restrict__ sums, uint32_t n_sums) {
    const uint32_t dmask = (1u << dbits) - 1u;  // the pass's significant digit bits
    __shared__ uint32_t sk[RS_CHUNK], sv[RS_CHUNK];
    __shared__ uint32_t wh[RS_W][256];  // per-wave digit counts, then local cursors
    __shared__ uint32_t delta[256], wsum[4];  // per digit: output position - position in the chunk
    __shared__ uint32_t carry[RS_SUMS_MAX], csum[4];  // exclusive prefix of the block totals
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < RS_W * 256; i += RS_T) (&wh[0][0])[i] = 0;
    uint32_t cx = 0, ct = 0;
    if (sums && threadIdx.x < RS_SUMS_MAX) {  // waves 0..3: inclusive scan of the totals
        ct = threadIdx.x < n_sums ? sums[threadIdx.x] : 0u;
        cx = ct;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(cx, o);
            if (lane >= o) cx += y;
        }
        if (lane == 63) csum[w] = cx;
    }
    __syncthreads();
    if (sums && threadIdx.x < RS_SUMS_MAX) {
        uint32_t pre = 0;
        for (int k = 0; k < w; k++) pre += csum[k];
        carry[threadIdx.x] = pre + cx - ct;
    }
    const SegChunk c = seg_chunk(blockIdx.x, seg_n, nbt);
    const uint32_t sub = (uint32_t)w * RS_SUB;  // the wave's part of the chunk
    uint32_t key[RS_STEPS];  // read once, kept for the ranking pass
#pragma unroll
    for (int step = 0; step < RS_STEPS; step++) {
        const uint32_t idx = sub + (uint32_t)step * 64 + lane;
        key[step] = idx < c.m ? kin[c.base + idx] : 0u;
    }
#pragma unroll
    for (int step = 0; step < RS_STEPS; step++)
        if (sub + (uint32_t)step * 64 + lane < c.m) atomicAdd(&wh[w][(key[step] >> shift) & dmask], 1u);
    __syncthreads();
    // per digit (threads 0..255, waves 0..3): start of its run in the chunk
    // (scan over digits), the block's output offset, and each wave's local cursor
    const int bin = threadIdx.x;
    uint32_t tot = 0, x = 0;
    if (bin < 256) {
#pragma unroll
        for (int k = 0; k < RS_W; k++) tot += wh[k][bin];
        x = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
    }
    __syncthreads();
    if (bin < 256) {
        uint32_t pre = 0;
        for (int k = 0; k < w; k++) pre += wsum[k];
        const uint32_t ls = pre + x - tot;
        const size_t g = c.hrow + (size_t)bin * nbt;
        const uint32_t o = offs[g] + (sums ? carry[g >> 12] : 0u);
        delta[bin] = o - ls;  // modulo 2^32
        uint32_t run = ls;
#pragma unroll
        for (int k = 0; k < RS_W; k++) {
            const uint32_t cw = wh[k][bin];
            wh[k][bin] = run;
            run += cw;
        }
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t vbase = (uint32_t)(c.base % seg_n);
#pragma unroll
    for (int step = 0; step < RS_STEPS; step++) {
        const uint32_t idx = sub + (uint32_t)step * 64 + lane;
        const bool valid = idx < c.m;
        const uint32_t val = !valid ? 0u : vin ? vin[c.base + idx] : vbase + idx;
        const uint32_t d = (key[step] >> shift) & dmask;
        // lanes with the same digit: for each significant bit, the lanes that
        // agree with this lane on it.  s = 0 where this lane's bit is set, all
        // ones where it is clear, so (ballot of set bits) ^ s selects either
        // that ballot or its complement: one bit-field extract, one compare
        // and four 32-bit logic ops per bit
        const uint64_t vb = __builtin_amdgcn_ballot_w64(valid);
        uint32_t p_lo = (uint32_t)vb, p_hi = (uint32_t)(vb >> 32);
        const uint32_t nd = ~d;
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            if (bit >= dbits) break;  // uniform
            const uint32_t sx = (uint32_t)((int)(nd << (31 - bit)) >> 31);
            const uint64_t bb = __builtin_amdgcn_ballot_w64(sx == 0u) & vb;
            p_lo &= (uint32_t)bb ^ sx;
            p_hi &= (uint32_t)(bb >> 32) ^ sx;
        }
        const uint64_t peers = ((uint64_t)p_hi << 32) | p_lo;
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        const uint32_t pos = wh[w][d];
        if (valid) {
            sk[pos + rank] = key[step];
            sv[pos + rank] = val;
        }
        // lowest lane of each digit group advances that digit's cursor
        if (valid && (peers & lt) == 0) wh[w][d] = pos + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < c.m; i += RS_T) {
        const uint32_t k = sk[i];
        const uint32_t pos = i + delta[(k >> shift) & dmask];
        kout[pos] = k;
        vout[pos] = sv[i];
    }
}

// In-place exclusive scan of 4096-element blocks; block totals to sums[].
__global__ __launch_bounds__(256) void scan_blocks(uint32_t* __restrict__ data, uint32_t n,
                                                   uint32_t* __restrict__ sums) {
    __shared__ uint32_t wsum[4];
    const size_t base = (size_t)blockIdx.x * 4096 + (size_t)threadIdx.x * 16;
    const bool whole = base + 16 <= n;
    uint32_t v[16];
    if (whole) {  // four 16-byte loads
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint4 u = reinterpret_cast<const uint4*>(data + base)[i];
            v[4 * i] = u.x; v[4 * i + 1] = u.y; v[4 * i + 2] = u.z; v[4 * i + 3] = u.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = (base + i < n) ? data[base + i] : 0u;
    }
    uint32_t t = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) t += v[i];
    // inclusive scan of thread totals across the wave
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t wpre = 0;
    for (int k = 0; k < w; k++) wpre += wsum[k];
    uint32_t run = wpre + x - t;  // exclusive prefix of this thread
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t c = v[i];
        v[i] = run;
        run += c;
    }
    if (whole) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            reinterpret_cast<uint4*>(data + base)[i] = make_uint4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (base + i < n) data[base + i] = v[i];
    }
    if (threadIdx.x == 255) sums[blockIdx.x] = wpre + x;
}

__global__ __launch_bounds__(256) void scan_add(uint32_t* __restrict__ data, uint32_t n,
                                                const uint32_t* __restrict__ sums) {
    const size_t base = (size_t)blockIdx.x * 4096;
    const uint32_t add = sums[blockIdx.x];
    for (int i = threadIdx.x; i < 4096; i += 256)
        if (base + i < n) data[base + i] += add;
}

}  // namespace

// Scratch of a scan of n entries: one level of block totals per 4096x
// reduction (reserved up front so no allocation falls between launches).
int scan_reserve(sift_context* h, uint32_t n, DevBuf* levels) {
    for (int depth = 0; n > 1 && depth < 4; depth++) {
        const uint32_t nb = (n + 4095) / 4096;
        SIFT_TRY(reserve(h, levels[depth], (size_t)nb * 4));
        n = nb;
    }
    return SIFT_OK;
}

// Everything the sort of nseg segments of seg_n pairs allocates.
int sort_reserve(sift_context* h, uint32_t seg_n, uint32_t nseg, DevBuf& hist, DevBuf* levels) {
    if (seg_n == 0 || nseg == 0) return SIFT_OK;
    const uint64_t nblk = (uint64_t)((seg_n + RS_CHUNK - 1) / RS_CHUNK) * nseg;
    if (nblk * 256 >= (1ull << 32)) return set_error(h, SIFT_ERR_UNSUPPORTED, "sort: too many chunks");
    SIFT_TRY(reserve(h, hist, (size_t)nblk * 256 * 4));
    return scan_reserve(h, (uint32_t)(nblk * 256), levels);
}

// Exclusive scan of n uint32 in place (recursive over 4096-element blocks).
int exclusive_scan(sift_context* h, uint32_t* data, uint32_t n, DevBuf* levels, int depth) {
    if (n == 0) return SIFT_OK;
    if (depth >= 4) return set_error(h, SIFT_ERR_UNSUPPORTED, "scan too deep");
    const uint32_t nb = (n + 4095) / 4096;
    SIFT_TRY(reserve(h, levels[depth], (size_t)nb * 4));
    auto* sums = static_cast<uint32_t*>(levels[depth].p);
    scan_blocks<<<nb, 256, 0, h->stream>>>(data, n, sums);
    SIFT_HIP_TRY(h, hipGetLastError());
    if (nb > 1) {
        SIFT_TRY(exclusive_scan(h, sums, nb, levels, depth + 1));
        scan_add<<<nb, 256, 0, h->stream>>>(data, n, sums);
        SIFT_HIP_TRY(h, hipGetLastError());
    }
    return SIFT_OK;
}

// Stable sort of the (key, value) pairs of each of nseg segments of seg_n
// elements by the low `bits` of key (a segment's pairs stay in its
// positions).  implicit_vals: the values are not read, each is its
// element's position in its segment (*vals is then only scratch).  Sorted
// pairs end in (*keys, *vals); the other buffers are scratch of the same size.
int radix_sort_pairs(sift_context* h, uint32_t** keys, uint32_t** vals, uint32_t** tk, uint32_t** tv,
                     uint32_t seg_n, uint32_t nseg, int bits, DevBuf& hist, DevBuf* levels, bool implicit_vals) {
    if (seg_n == 0 || nseg == 0) return SIFT_OK;
    if ((uint64_t)seg_n * nseg >= (1ull << 32)) return set_error(h, SIFT_ERR_UNSUPPORTED, "sort: too many pairs");
    const uint32_t nbt = (seg_n + RS_CHUNK - 1) / RS_CHUNK;
    const uint64_t nblk = (uint64_t)nbt * nseg;
    if (nblk * 256 >= (1ull << 32)) return set_error(h, SIFT_ERR_UNSUPPORTED, "sort: too many chunks");
    SIFT_TRY(reserve(h, hist, (size_t)nblk * 256 * 4));
    auto* hp = static_cast<uint32_t*>(hist.p);
    const uint32_t n_cnt = (uint32_t)(nblk * 256);
    const uint32_t n_sums = (n_cnt + 4095) / 4096;
    // up to 256 scan blocks (2^20 counts, 32M pairs): one scan launch, and the
    // scatter carries the block totals itself
    const bool carry = n_sums <= (uint32_t)RS_SUMS_MAX;
    if (carry) SIFT_TRY(reserve(h, levels[0], (size_t)n_sums * 4));
    auto* sums = carry ? static_cast<uint32_t*>(levels[0].p) : nullptr;
    for (int shift = 0; shift < std::max(bits, 1); shift += 8) {
        // digit bits at or above `bits` are constant within a segment: masked off
        const int dbits = std::min(8, std::max(bits, 1) - shift);
        rs_hist<<<(unsigned)nblk, RS_T, 0, h->stream>>>(*keys, seg_n, nbt, shift, (1u << dbits) - 1u, hp);
        SIFT_HIP_TRY(h, hipGetLastError());
        if (carry) {
            scan_blocks<<<n_sums, 256, 0, h->stream>>>(hp, n_cnt, sums);
            SIFT_HIP_TRY(h, hipGetLastError());
        } else {
            SIFT_TRY(exclusive_scan(h, hp, n_cnt, levels, 0));
        }
        const uint32_t* vin = implicit_vals && shift == 0 ? nullptr : *vals;
        rs_scatter<<<(unsigned)nblk, RS_T, 0, h->stream>>>(*keys, vin, *tk, *tv, seg_n, nbt, shift, dbits, hp,
                                                            sums, n_sums);
        SIFT_HIP_TRY(h, hipGetLastError());
        std::swap(*keys, *tk);
        std::swap(*vals, *tv);
    }
    return SIFT_OK;
}

}  // namespace sift
