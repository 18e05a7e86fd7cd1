// Label-row helpers shared by the plane kernels: a row of the label bitmap
// (samples (x, y, 0..dim-1)) as W u64 words held in registers, and the
// staging of whole label planes into LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bmf_internal.h"

namespace bmf {

// u64 words of one staged label plane for row-word class W (dim <= 64 W),
// plus the zero pad word.
__host__ __device__ constexpr int plane_words(int W) { return 64 * W * W + 1; }

// ------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Chunk j of the chunk table (j wave-uniform), for kernels that never write
// the table: read through the constant address space, its fields load with
// scalar loads into scalar registers and are not reloaded after the
// kernel's own stores (which the compiler cannot otherwise tell apart from
// the table).
__device__ __forceinline__ ChunkMeta read_chunk(const ChunkMeta* table, int j) {
    static_assert(sizeof(ChunkMeta) % 4 == 0, "ChunkMeta is copied in 32-bit words");
    constexpr int kWords = (int)(sizeof(ChunkMeta) / 4);
    using ConstWords = const __attribute__((address_space(4))) uint32_t*;
    const ConstWords src = (ConstWords)(table + uniform_int(j));
    uint32_t w[kWords];
#pragma unroll
    for (int i = 0; i < kWords; i++) w[i] = src[i];
    ChunkMeta m;
    __builtin_memcpy(&m, w, sizeof(ChunkMeta));
    return m;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Exclusive prefix and total over a block of NT threads (NT / 64 <= 64
// waves) of N values per thread, one value at a time: lane shuffles inside
// each wave, then wave 0 scans the wave totals.  s_w: N x (NT / 64 + 1)
// scratch in LDS (the last column holds the totals).
template <int N, int NT>
__device__ __forceinline__ void block_scan(const int64_t (&v)[N], int64_t (&excl)[N], int64_t (&tot)[N],
                                           int64_t (*s_w)[NT / 64 + 1]) {
    constexpr int NW = NT / 64;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    int64_t inc[N];
    for (int q = 0; q < N; q++) {
        int64_t x = v[q];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int64_t t = __shfl_up(x, off);
            if (lane >= off) x += t;
        }
        inc[q] = x;
        if (lane == 63) s_w[q][w] = x;
    }
    __syncthreads();
    if (w == 0) {
        for (int q = 0; q < N; q++) {
            const int64_t x = lane < NW ? s_w[q][lane] : 0;
            int64_t y = x;
#pragma unroll
            for (int off = 1; off < NW; off <<= 1) {
                const int64_t t = __shfl_up(y, off);
                if (lane >= off) y += t;
            }
            if (lane < NW) s_w[q][lane] = y - x;
            if (lane == NW - 1) s_w[q][NW] = y;
        }
    }
    __syncthreads();
    for (int q = 0; q < N; q++) {
        excl[q] = s_w[q][w] + inc[q] - v[q];
        tot[q] = s_w[q][NW];
    }
    __syncthreads();  // the scratch may be reused
}

// Chunk j whose half-open range [prefix[j], prefix[j+1]) holds g.
__device__ __forceinline__ int find_chunk(const int64_t* __restrict__ prefix, int n, int64_t g) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// q = n / d for n <= 2^24, 0 < d < 2^24 and a quotient below 2^22, with
// rcp = 1/d correctly rounded.  (float)n is exact and fl(n * rcp) =
// (n/d)(1 + e) with |e| <= 2^-23 + 2^-48, so the estimate is within 0.5 of
// n/d and its truncation is floor(n/d) - 1, floor(n/d) or floor(n/d) + 1:
// one correction step each way, as selects, makes it exact.  (The sample
// coordinates divide by dim and dim^2 with quotients below 256.)
__device__ __forceinline__ uint32_t udiv24(uint32_t n, uint32_t d, float rcp) {
    int32_t q = (int32_t)((float)n * rcp);
    int32_t r = (int32_t)n - q * (int32_t)d;
    const bool lo = r < 0;
    q = lo ? q - 1 : q;
    r = lo ? r + (int32_t)d : r;
    q = r >= (int32_t)d ? q + 1 : q;
    return (uint32_t)q;
}

// A label row: bits z = 0..dim-1 of row (x, y) in W u64 words, W = ceil(dim/64).
template <int W>
struct RowBits {
    uint64_t w[W];
};

template <int W>
__device__ __forceinline__ RowBits<W> zero_row() {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) r.w[k] = 0ull;
    return r;
}

template <int W>
__device__ __forceinline__ RowBits<W> full_row(int dim) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) {
        const int nb = dim - 64 * k;
        r.w[k] = nb >= 64 ? ~0ull : ((1ull << nb) - 1ull);
    }
    return r;
}

// Row (x, y) of a chunk's flat label bitmap: bits start at (x*dim + y)*dim.
template <int W>
__device__ __forceinline__ RowBits<W> load_row(const uint64_t* __restrict__ L, int64_t nwords, int dim, int x, int y) {
    RowBits<W> r = zero_row<W>();
    if (x < 0 || y < 0 || x >= dim || y >= dim) return r;
    const int64_t base = ((int64_t)x * dim + y) * dim;
#pragma unroll
    for (int k = 0; k < W; k++) {
        const int nb = dim - 64 * k;  // >= 1 because W = ceil(dim/64)
        const int64_t bit = base + 64 * k;
        const int64_t w = bit >> 6;
        const int sh = (int)(bit & 63);
        uint64_t v = L[w];
        if (sh) {
            const uint64_t hi = (w + 1 < nwords) ? L[w + 1] : 0ull;
            v = (v >> sh) | (hi << (64 - sh));
        }
        if (nb < 64) v &= (1ull << nb) - 1ull;
        r.w[k] = v;
    }
    return r;
}

template <int W>
__device__ __forceinline__ uint64_t word_at(const RowBits<W>& r, int k) {
    uint64_t v = 0ull;
#pragma unroll
    for (int q = 0; q < W; q++) v = k == q ? r.w[q] : v;
    return v;
}

template <int W>
__device__ __forceinline__ uint32_t bit_at(const RowBits<W>& r, int z) {
    return (uint32_t)((word_at<W>(r, z >> 6) >> (z & 63)) & 1ull);
}

// popcount of bits [0, z)
template <int W>
__device__ __forceinline__ uint32_t prefix_pop(const RowBits<W>& r, int z) {
    const int wz = z >> 6;
    const uint64_t m = (1ull << (z & 63)) - 1ull;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint64_t v = r.w[k];
        c += k < wz ? (uint32_t)__popcll(v) : (k == wz ? (uint32_t)__popcll(v & m) : 0u);
    }
    return c;
}

template <int W>
__device__ __forceinline__ uint32_t total_pop(const RowBits<W>& r) {
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < W; k++) c += (uint32_t)__popcll(r.w[k]);
    return c;
}

template <int W>
__device__ __forceinline__ RowBits<W> rxor(const RowBits<W>& a, const RowBits<W>& b, bool valid) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) r.w[k] = valid ? (a.w[k] ^ b.w[k]) : 0ull;
    return r;
}

// z-edges of a row: bit z set when labels z and z+1 differ (z <= dim-2).
template <int W>
__device__ __forceinline__ RowBits<W> zedges(const RowBits<W>& a, int dim) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint64_t nxt = k + 1 < W ? a.w[k + 1] : 0ull;
        r.w[k] = a.w[k] ^ ((a.w[k] >> 1) | (nxt << 63));
    }
    const int last = dim - 1;  // clear bit dim-1 and above
#pragma unroll
    for (int k = 0; k < W; k++) {
        const int lo = 64 * k;
        if (last <= lo) r.w[k] = 0ull;
        else if (last < lo + 64) r.w[k] &= (1ull << (last - lo)) - 1ull;
    }
    return r;
}

template <int W>
__device__ __forceinline__ bool row_equal(const RowBits<W>& a, const RowBits<W>& b) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < W; k++) ok = ok && a.w[k] == b.w[k];
    return ok;
}

// True when the four rows around a row of cells hold one constant label,
// i.e. no cell of the row is mixed.
template <int W>
__device__ __forceinline__ bool quad_uniform(const RowBits<W>& r00, const RowBits<W>& r01, const RowBits<W>& r10,
                                             const RowBits<W>& r11, int dim) {
    const RowBits<W> z = zero_row<W>(), f = full_row<W>(dim);
    const bool allz = row_equal(r00, z) && row_equal(r01, z) && row_equal(r10, z) && row_equal(r11, z);
    const bool allf = row_equal(r00, f) && row_equal(r01, f) && row_equal(r10, f) && row_equal(r11, f);
    return allz || allf;
}

// 8-corner mask of cell (x, y, z) from the four rows around it; corner bit
// 4dx + 2dy + dz (DMCChunk.cpp:200-438).
template <int W>
__device__ __forceinline__ uint32_t cell_mask(const RowBits<W>& r00, const RowBits<W>& r01, const RowBits<W>& r10,
                                              const RowBits<W>& r11, int z) {
    return bit_at<W>(r00, z) | (bit_at<W>(r00, z + 1) << 1) | (bit_at<W>(r01, z) << 2) |
           (bit_at<W>(r01, z + 1) << 3) | (bit_at<W>(r10, z) << 4) | (bit_at<W>(r10, z + 1) << 5) |
           (bit_at<W>(r11, z) << 6) | (bit_at<W>(r11, z + 1) << 7);
}

// ------------------------------------------------------------------------
// plane staging: three consecutive label planes of a chunk, realigned
// to u64 words in LDS (coalesced funnel-shifted loads); rows are then read
// from LDS.  Word pw of each plane is a zero pad for the funnel reads.
// ------------------------------------------------------------------------
// NP planes staged together: plane q is bitmap plane xp[q] of src[q] (all
// zeros outside 0 .. dim-1) into dst[q], a zero word after each plane's
// last; every load is issued before the first LDS write.
template <int NP>
__device__ __forceinline__ void stage_plane_group(uint64_t* const (&dst)[NP], const uint64_t* const (&src)[NP],
                                                  int64_t nwords, int dim, const int (&xp)[NP]) {
    const int64_t pbits = (int64_t)dim * dim;
    const int pw = (int)((pbits + 63) >> 6);
    for (int k = threadIdx.x; k <= pw; k += blockDim.x) {
        uint64_t lo[NP], hi[NP];
        int sh[NP];
        bool ok[NP], has_hi[NP];
#pragma unroll
        for (int q = 0; q < NP; q++) {
            ok[q] = xp[q] >= 0 && xp[q] < dim && k < pw;
            const int64_t bit = ok[q] ? (int64_t)xp[q] * pbits + 64ll * k : 0;
            const int64_t w = bit >> 6;
            sh[q] = (int)(bit & 63);
            has_hi[q] = w + 1 < nwords;
            lo[q] = src[q][w];
            hi[q] = src[q][has_hi[q] ? w + 1 : w];
        }
        const int64_t rem = pbits - 64ll * k;  // bits of the plane from word k on
#pragma unroll
        for (int q = 0; q < NP; q++) {
            uint64_t v = lo[q];
            if (sh[q]) v = (v >> sh[q]) | ((has_hi[q] ? hi[q] : 0ull) << (64 - sh[q]));
            if (rem < 64) v &= rem > 0 ? (1ull << rem) - 1ull : 0ull;
            dst[q][k] = ok[q] ? v : 0ull;
        }
    }
}

__device__ __forceinline__ void stage_planes(uint64_t* p0, uint64_t* p1, uint64_t* p2, const uint64_t* __restrict__ L,
                                             int64_t nwords, int dim, int x0) {
    uint64_t* const dst[3] = {p0, p1, p2};
    const uint64_t* const src[3] = {L, L, L};
    const int xp[3] = {x0, x0 + 1, x0 + 2};
    stage_plane_group<3>(dst, src, nwords, dim, xp);
}

// A wave-uniform row moved to scalar registers: the bit operations on it then
// run on the scalar unit and it holds no vector registers.
template <int W>
__device__ __forceinline__ RowBits<W> uniform_row(const RowBits<W>& a) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a.w[k]);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a.w[k] >> 32));
        r.w[k] = ((uint64_t)hi << 32) | lo;
    }
    return r;
}

// Row y of a staged plane.
// Both words of every funnel read load unconditionally (the staged plane's
// zero pad word keeps w + 1 in bounds) and an invalid row reads row 0, so
// the reads carry no branches and issue together.
template <int W>
__device__ __forceinline__ RowBits<W> lds_row(const uint64_t* P, int dim, int y, bool valid) {
    RowBits<W> r;
    const int base = (valid ? y : 0) * dim;
#pragma unroll
    for (int k = 0; k < W; k++) {
        const int bit = base + 64 * k;
        const int w = bit >> 6;
        const int sh = bit & 63;
        const uint64_t lo = P[w], hi = P[w + 1];
        uint64_t v = (lo >> sh) | ((hi << 1) << (63 - sh));  // the second term is 0 for sh == 0
        const int nb = dim - 64 * k;
        if (nb < 64) v &= nb > 0 ? (1ull << nb) - 1ull : 0ull;
        r.w[k] = valid ? v : 0ull;
    }
    return r;
}

template <int W>
__device__ __forceinline__ RowBits<W> rshl1(const RowBits<W>& a) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) r.w[k] = (a.w[k] << 1) | (k > 0 ? (a.w[k - 1] >> 63) : 0ull);
    return r;
}

// bit z of the result = bit z + 1 of a (0 past the row)
template <int W>
__device__ __forceinline__ RowBits<W> rshr1(const RowBits<W>& a) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) r.w[k] = (a.w[k] >> 1) | (k + 1 < W ? (a.w[k + 1] << 63) : 0ull);
    return r;
}

// Lane-wise views of a wave-uniform 64-bit word whose bit i belongs to lane
// i (word k of a row, for samples z = 64 k + lane): the lane's bit, and the
// number of set bits below it (v_mbcnt).
__device__ __forceinline__ uint32_t lane_bit(uint64_t w, int lane) { return (uint32_t)(w >> lane) & 1u; }

// The same for a word held in scalar registers (the same value in every
// lane), as this lane's bit of the word used as the wave's lane mask: one
// v_cndmask_b32 with the word as its 64-bit condition, no 64-bit shift.  The
// caller's lane must be the executing lane.
__device__ __forceinline__ uint32_t lane_bit_u(uint64_t w) {
    uint32_t r;
    asm("v_cndmask_b32_e64 %0, 0, 1, %1" : "=v"(r) : "s"(w));
    return r;
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t w) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(w >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)w, 0u));
}

template <int W>
__device__ __forceinline__ RowBits<W> ror(const RowBits<W>& a, const RowBits<W>& b) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) r.w[k] = a.w[k] | b.w[k];
    return r;
}

// Samples z in [1, dim-2]: the z-range where both cells beside an x- or
// y-edge lie in the chunk.
template <int W>
__device__ __forceinline__ RowBits<W> inner_z(int dim) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) {
        uint64_t v = ~0ull;
        const int lo = 64 * k;
        if (lo == 0) v &= ~1ull;
        const int hi = dim - 2 - lo;  // last valid bit in this word
        if (hi < 0) v = 0ull;
        else if (hi < 63) v &= (2ull << hi) - 1ull;
        r.w[k] = v;
    }
    return r;
}

template <int W>
__device__ __forceinline__ RowBits<W> rand_(const RowBits<W>& a, const RowBits<W>& b) {
    RowBits<W> r;
#pragma unroll
    for (int k = 0; k < W; k++) r.w[k] = a.w[k] & b.w[k];
    return r;
}

// Interior sign-changing edges owned by the samples of row (x, y): the four
// cells around each lie in the chunk.
template <int W>
__device__ __forceinline__ void interior_edges(const RowBits<W>& r00, const RowBits<W>& r01, const RowBits<W>& r10,
                                               int dim, int x, int y, RowBits<W>& ex, RowBits<W>& ey,
                                               RowBits<W>& ez) {
    const RowBits<W> iz = inner_z<W>(dim);
    const bool xin = x >= 1 && x <= dim - 2, yin = y >= 1 && y <= dim - 2;
    ex = (yin && x + 1 < dim) ? rand_<W>(rxor<W>(r00, r10, true), iz) : zero_row<W>();
    ey = (xin && y + 1 < dim) ? rand_<W>(rxor<W>(r00, r01, true), iz) : zero_row<W>();
    ez = (xin && yin) ? zedges<W>(r00, dim) : zero_row<W>();
}


}  // namespace bmf
