// HIP kernels of the chunk-extraction path (gfx950, wave64).
//
// Per batch:  plan -> sample -> count -> scan -> emit -> sort_adj ->
//             K x (dual, primal) -> finalize
//
// Data layout in HBM (per sampled chunk, offsets from the plan kernel):
//   densities   FP32 [x][y][z]                    (DMCChunk density_block)
//   labels      flat bitmap over the sample index (x*dim + y)*dim + z,
//               u64 word i/64, bit i%64, set when the density is < 0;
//               one word per wave of 64 consecutive samples (one ballot)
//   row_v/row_t per label row (x, y): vertex / triangle counts, then
//               exclusive bases after the scan
//   output      positions float3, boundary u8, chunk-local triangle
//               indices, valence u32, adjacency BMF_MAX_VALENCE x u32
// Cell masks and the reference's u32 label words are never materialised on
// the hot path: the row kernels rebuild them from the bitmap with shifts.
#include <hip/hip_runtime.h>

#include "../../include/bmf_mc_tables.h"
#include "bmf_internal.h"
#include "samplers.h"

namespace bmf {

__constant__ uint64_t c_tri_table[256] = BMF_TRI_TABLE_INIT;

// Local edge e of a cell -> (dx, dy, dz, axis) of the lattice edge it is
// (DMCChunk.cpp:543-565).
__constant__ int8_t c_edge_dx[12] = {0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 1, 1};
__constant__ int8_t c_edge_dy[12] = {0, 0, 1, 1, 0, 0, 0, 0, 0, 1, 0, 1};
__constant__ int8_t c_edge_dz[12] = {0, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0, 0};
__constant__ int8_t c_edge_ax[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};

// ------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
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

// q = n / d for n < 2^24 (float estimate + exact correction).
__device__ __forceinline__ uint32_t udiv24(uint32_t n, uint32_t d, float rcp) {
    int32_t q = (int32_t)((float)n * rcp);
    int32_t r = (int32_t)n - q * (int32_t)d;
    while (r < 0) { q--; r += (int32_t)d; }
    while (r >= (int32_t)d) { q++; r -= (int32_t)d; }
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
// plan: geometry, culling, arena offsets, counters (one block)
// ------------------------------------------------------------------------
constexpr int kPlanThreads = 1024;

__global__ void __launch_bounds__(kPlanThreads) k_plan(Arenas A, int n, int allow_cull, int reset_bumps) {
    __shared__ int64_t s_scan[4][kPlanThreads];
    const int tid = threadIdx.x;
    const SamplerDev* __restrict__ S = A.sampler;
    const int per = (n + kPlanThreads - 1) / kPlanThreads;
    const int j0 = tid * per, j1 = min(n, j0 + per);
    int64_t sum[4] = {0, 0, 0, 0};  // samples, label words, rows, tiles

    for (int j = j0; j < j1; j++) {
        const bmf_chunk_desc d = A.descs[j];
        ChunkMeta m;
        // DMCChunk::label_grid geometry (DMCChunk.cpp:94-101)
        float t = d.overlap * 2.0f;
        t = 1.0f + t;
        t = d.size * t;
        m.delta = t / (float)(d.dim - 1);
        const float so = d.size * d.overlap;
        m.opos[0] = d.pos[0] - so;
        m.opos[1] = d.pos[1] - so;
        m.opos[2] = d.pos[2] - so;
        m.dim = d.dim;
        m.iters = d.iters;
        m.bp = d.boundary_processing;
        m.state = kStateSampled;
        if (allow_cull && S->kind == BMF_SAMPLER_TERRAIN) {
            // |fbm| <= 1 exactly, so |amp*fbm| <= |amp|; sample y is monotone in j.
            const float a = fabsf(S->amp);
            const float ylo = m.opos[1];
            const float yhi = (float)(d.dim - 1) * m.delta + m.opos[1];
            if (ylo - S->y0 > a) m.state = kStateAir;
            else if (yhi - S->y0 < -a) m.state = kStateSolid;
        }
        const int64_t ns = (int64_t)d.dim * d.dim * d.dim;
        const bool sampled = m.state == kStateSampled;
        m.lab_words = sampled ? (ns + 63) / 64 : 0;
        m.any_air = m.state == kStateAir;
        m.any_solid = m.state == kStateSolid;
        m.contains_mesh = 0;
        m.status = BMF_OK;
        m.V = 0;
        m.T = 0;
        m.v_off = 0;
        m.i_off = 0;
        m.dens_off = sum[0];
        m.lab_off = sum[1];
        m.row_off = sum[2];
        sum[0] += sampled ? ns : 0;
        sum[1] += m.lab_words;
        sum[2] += sampled ? (int64_t)d.dim * d.dim : 0;
        sum[3] += sampled ? (ns + kSampleTile - 1) / kSampleTile : 0;
        A.meta[j] = m;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) s_scan[q][tid] = sum[q];
    __syncthreads();
    // Hillis-Steele inclusive scan across threads
    for (int off = 1; off < kPlanThreads; off <<= 1) {
        int64_t v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = tid >= off ? s_scan[q][tid - off] : 0;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) s_scan[q][tid] += v[q];
        __syncthreads();
    }
    int64_t base[4];
#pragma unroll
    for (int q = 0; q < 4; q++) base[q] = s_scan[q][tid] - sum[q];
    const int64_t tot_samples = s_scan[0][kPlanThreads - 1];
    const int64_t tot_words = s_scan[1][kPlanThreads - 1];
    const int64_t tot_rows = s_scan[2][kPlanThreads - 1];
    const int64_t tot_tiles = s_scan[3][kPlanThreads - 1];
    const bool overflow = tot_samples > A.cap_samples || tot_words > A.cap_samples / 64 + A.cap_chunks ||
                          tot_rows > A.cap_rows;
    int64_t tp = base[3], rp = base[2];
    for (int j = j0; j < j1; j++) {
        ChunkMeta& m = A.meta[j];
        m.dens_off += base[0];
        m.lab_off += base[1];
        m.row_off += base[2];
        A.tile_prefix[j] = overflow ? 0 : tp;
        A.row_prefix[j] = overflow ? 0 : rp;
        if (m.state == kStateSampled) {
            const int64_t ns = (int64_t)m.dim * m.dim * m.dim;
            tp += (ns + kSampleTile - 1) / kSampleTile;
            rp += (int64_t)m.dim * m.dim;
            if (overflow) {
                m.state = kStateSkipped;
                m.status = BMF_ECAPACITY;
            }
        } else {
            // culled chunks: DMCChunk.cpp:159-162 on uniform labels
            m.contains_mesh = (m.state == kStateAir) && (m.dim % 32 != 0);
        }
    }
    if (tid == 0) {
        A.tile_prefix[n] = overflow ? 0 : tot_tiles;
        A.row_prefix[n] = overflow ? 0 : tot_rows;
        Counters* c = A.ctr;
        c->sample_head = 0;
        if (reset_bumps) {
            c->v_bump = 0;
            c->i_bump = 0;
            c->need_v = 0;
            c->need_i = 0;
            c->overflow = 0;
            c->error = 0;
        }
        if (overflow) c->overflow = 1;
        c->total_tiles = overflow ? 0 : tot_tiles;
        c->total_rows = overflow ? 0 : tot_rows;
        c->total_samples = tot_samples;
    }
}

// ------------------------------------------------------------------------
// sample: persistent; work-queue tiles of kSampleTile consecutive samples
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kSampleThreads) k_sample(Arenas A, int n) {
    __shared__ int64_t s_tile;
    const SamplerDev* __restrict__ S = A.sampler;
    const int64_t total = A.ctr->total_tiles;
    const int lane = lane_id();
    for (;;) {
        if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(&A.ctr->sample_head, 1ull);
        __syncthreads();
        const int64_t tile = s_tile;
        __syncthreads();
        if (tile >= total) break;
        const int j = find_chunk(A.tile_prefix, n, tile);
        const ChunkMeta& M = A.meta[j];
        const uint32_t dim = (uint32_t)M.dim, dim2 = dim * dim, ns = dim2 * dim;
        const float rd = 1.0f / (float)dim, rd2 = 1.0f / (float)dim2;
        const float delta = M.delta, ox = M.opos[0], oy = M.opos[1], oz = M.opos[2];
        float* __restrict__ dens = A.dens + M.dens_off;
        uint64_t* __restrict__ lab = A.labels + M.lab_off;
        const uint32_t first = (uint32_t)(tile - A.tile_prefix[j]) * kSampleTile;
        bool air = false, solid = false;
#pragma unroll 1
        for (int r = 0; r < kSampleRounds; r++) {
            const uint32_t i = first + r * kSampleThreads + threadIdx.x;
            const bool valid = i < ns;
            float d = 0.0f;
            if (valid) {
                const uint32_t x = udiv24(i, dim2, rd2);
                const uint32_t rem = i - x * dim2;
                const uint32_t y = udiv24(rem, dim, rd);
                const uint32_t z = rem - y * dim;
                // NoiseSampler.cpp:21-27: (float)i * scale + p, two roundings
                const float px = (float)x * delta + ox;
                const float py = (float)y * delta + oy;
                const float pz = (float)z * delta + oz;
                d = sample_density(*S, px, py, pz);
                dens[i] = d;
            }
            const uint64_t b = __ballot(valid && d < 0.0f);
            const uint64_t sb = __ballot(valid && !(d < 0.0f));
            if (lane == 0 && (i - (uint32_t)lane) < ns) lab[i >> 6] = b;
            air = air || b != 0ull;
            solid = solid || sb != 0ull;
        }
        if (lane == 0) {
            if (air) atomicOr(&A.meta[j].any_air, 1u);
            if (solid) atomicOr(&A.meta[j].any_solid, 1u);
        }
    }
}

// ------------------------------------------------------------------------
// count: one wave per label row -> vertices and triangles of the row
// ------------------------------------------------------------------------
template <int W>
__device__ __forceinline__ void count_row(const Arenas& A, const ChunkMeta& M, int64_t r, const uint8_t* s_ntri,
                                          uint32_t& nv, uint32_t& nt) {
    const int dim = M.dim;
    const int lane = lane_id();
    const int x = (int)(r / dim), y = (int)(r - (int64_t)x * dim);
    const uint64_t* L = A.labels + M.lab_off;
    const RowBits<W> r00 = load_row<W>(L, M.lab_words, dim, x, y);
    const RowBits<W> r10 = load_row<W>(L, M.lab_words, dim, x + 1, y);
    const RowBits<W> r01 = load_row<W>(L, M.lab_words, dim, x, y + 1);
    const RowBits<W> r11 = load_row<W>(L, M.lab_words, dim, x + 1, y + 1);
    nv = total_pop<W>(rxor<W>(r00, r10, x + 1 < dim)) + total_pop<W>(rxor<W>(r00, r01, y + 1 < dim)) +
         total_pop<W>(zedges<W>(r00, dim));
    nt = 0;
    if (x < dim - 1 && y < dim - 1 && !quad_uniform<W>(r00, r01, r10, r11, dim)) {
        for (int z0 = 0; z0 < dim - 1; z0 += 64) {
            const int z = z0 + lane;
            const uint32_t c = z < dim - 1 ? s_ntri[cell_mask<W>(r00, r01, r10, r11, z)] : 0u;
            nt += (uint32_t)__popcll(__ballot(c & 1u)) + 2u * (uint32_t)__popcll(__ballot(c & 2u)) +
                  4u * (uint32_t)__popcll(__ballot(c & 4u));
        }
    }
}

__global__ void __launch_bounds__(kRowThreads) k_count(Arenas A, int n) {
    __shared__ uint8_t s_ntri[256];
    for (int t = threadIdx.x; t < 256; t += blockDim.x) s_ntri[t] = (uint8_t)((c_tri_table[t] >> 60) / 3);
    __syncthreads();
    const int wave = uniform_int(threadIdx.x >> 6);
    const int lane = lane_id();
    const int64_t nw = (int64_t)gridDim.x * (kRowThreads / 64);
    const int64_t total = A.ctr->total_rows;
    for (int64_t g = (int64_t)blockIdx.x * (kRowThreads / 64) + wave; g < total; g += nw) {
        const int j = find_chunk(A.row_prefix, n, g);
        const ChunkMeta& M = A.meta[j];
        const int64_t r = g - A.row_prefix[j];
        uint32_t nv = 0, nt = 0;
        if (M.any_air && M.any_solid) {
            switch ((M.dim + 63) >> 6) {
                case 1: count_row<1>(A, M, r, s_ntri, nv, nt); break;
                case 2: count_row<2>(A, M, r, s_ntri, nv, nt); break;
                case 3: count_row<3>(A, M, r, s_ntri, nv, nt); break;
                default: count_row<4>(A, M, r, s_ntri, nv, nt); break;
            }
        }
        if (lane == 0) {
            A.row_v[M.row_off + r] = nv;
            A.row_t[M.row_off + r] = nt;
        }
    }
}

// ------------------------------------------------------------------------
// scan: one block per chunk -> row bases, totals, output reservation
// ------------------------------------------------------------------------
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kScanThreads) k_scan(Arenas A, int n) {
    __shared__ uint32_t s_v[kScanThreads], s_t[kScanThreads];
    __shared__ int64_t s_off[2];
    __shared__ int s_ok;
    const int tid = threadIdx.x;
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        ChunkMeta& M = A.meta[j];
        if (M.state != kStateSampled) continue;
        const int dim = M.dim;
        const int64_t R = (int64_t)dim * dim;
        uint32_t* rv = A.row_v + M.row_off;
        uint32_t* rt = A.row_t + M.row_off;
        const int64_t per = (R + kScanThreads - 1) / kScanThreads;
        const int64_t r0 = tid * per, r1 = min(R, r0 + per);
        uint32_t sv = 0, st = 0;
        for (int64_t r = r0; r < r1; r++) {
            sv += rv[r];
            st += rt[r];
        }
        s_v[tid] = sv;
        s_t[tid] = st;
        __syncthreads();
        for (int off = 1; off < kScanThreads; off <<= 1) {
            uint32_t av = tid >= off ? s_v[tid - off] : 0u, at = tid >= off ? s_t[tid - off] : 0u;
            __syncthreads();
            s_v[tid] += av;
            s_t[tid] += at;
            __syncthreads();
        }
        uint32_t bv = s_v[tid] - sv, bt = s_t[tid] - st;
        for (int64_t r = r0; r < r1; r++) {
            const uint32_t a = rv[r], b = rt[r];
            rv[r] = bv;
            rt[r] = bt;
            bv += a;
            bt += b;
        }
        if (tid == 0) {
            const uint32_t V = s_v[kScanThreads - 1], T = s_t[kScanThreads - 1];
            M.V = (int32_t)V;
            M.T = (int32_t)T;
            M.contains_mesh = M.any_air && (M.any_solid || (dim % 32 != 0));
            int ok = 1;
            if (V > 0) {
                const int64_t vo = (int64_t)atomicAdd(&A.ctr->v_bump, (unsigned long long)V);
                const int64_t io = (int64_t)atomicAdd(&A.ctr->i_bump, (unsigned long long)(3ull * T));
                atomicAdd(&A.ctr->need_v, (unsigned long long)V);
                atomicAdd(&A.ctr->need_i, (unsigned long long)(3ull * T));
                if (vo + V > A.cap_vertices || io + 3ll * T > A.cap_indices) {
                    ok = 0;
                    M.status = BMF_ECAPACITY;
                    A.ctr->overflow = 1;
                }
                M.v_off = vo;
                M.i_off = io;
            }
            s_ok = ok;
            s_off[0] = M.v_off;
        }
        __syncthreads();
        // valence counters start at zero (filled by atomics in k_emit)
        if (s_ok && M.iters > 0) {
            const int64_t V = M.V, vo = s_off[0];
            for (int64_t v = tid; v < V; v += kScanThreads) A.valence[vo + v] = 0u;
        }
        __syncthreads();
    }
}

// tile prefixes for the smoothing kernels (one block)
__global__ void __launch_bounds__(kPlanThreads) k_plan_smooth(Arenas A, int n) {
    __shared__ int64_t s_scan[2][kPlanThreads];
    const int tid = threadIdx.x;
    const int per = (n + kPlanThreads - 1) / kPlanThreads;
    const int j0 = tid * per, j1 = min(n, j0 + per);
    int64_t sum[2] = {0, 0};
    for (int j = j0; j < j1; j++) {
        const ChunkMeta& M = A.meta[j];
        const bool smooth = M.iters > 0 && M.V > 0 && M.status == BMF_OK && M.state == kStateSampled;
        sum[0] += smooth ? (M.V + kSmoothThreads - 1) / kSmoothThreads : 0;
        sum[1] += smooth ? (M.T + kSmoothThreads - 1) / kSmoothThreads : 0;
    }
    s_scan[0][tid] = sum[0];
    s_scan[1][tid] = sum[1];
    __syncthreads();
    for (int off = 1; off < kPlanThreads; off <<= 1) {
        int64_t a = tid >= off ? s_scan[0][tid - off] : 0, b = tid >= off ? s_scan[1][tid - off] : 0;
        __syncthreads();
        s_scan[0][tid] += a;
        s_scan[1][tid] += b;
        __syncthreads();
    }
    int64_t bv = s_scan[0][tid] - sum[0], bt = s_scan[1][tid] - sum[1];
    for (int j = j0; j < j1; j++) {
        const ChunkMeta& M = A.meta[j];
        const bool smooth = M.iters > 0 && M.V > 0 && M.status == BMF_OK && M.state == kStateSampled;
        A.vtile_prefix[j] = bv;
        A.ttile_prefix[j] = bt;
        bv += smooth ? (M.V + kSmoothThreads - 1) / kSmoothThreads : 0;
        bt += smooth ? (M.T + kSmoothThreads - 1) / kSmoothThreads : 0;
    }
    if (tid == 0) {
        A.vtile_prefix[n] = s_scan[0][kPlanThreads - 1];
        A.ttile_prefix[n] = s_scan[1][kPlanThreads - 1];
        A.ctr->total_vtiles = s_scan[0][kPlanThreads - 1];
        A.ctr->total_ttiles = s_scan[1][kPlanThreads - 1];
    }
}

// ------------------------------------------------------------------------
// emit: one wave per label row; lanes = z.  Vertices of the row's samples
// (ids by popcount prefix over the row's edge masks), triangles of the
// row's interior cells (ids by a bit-sliced ballot prefix of the per-cell
// triangle counts), adjacency by atomic append (sorted afterwards).
// ------------------------------------------------------------------------
template <int W>
__device__ __forceinline__ void emit_row(const Arenas& A, const ChunkMeta& M, int64_t r, const uint8_t* s_ntri,
                                         uint32_t (*s_vid)[64]) {
    const int dim = M.dim;
    const int lane = lane_id();
    const uint64_t lt = lanemask_lt();
    const int x = (int)(r / dim), y = (int)(r - (int64_t)x * dim);
    const uint64_t* L = A.labels + M.lab_off;
    const int64_t nwd = M.lab_words;
    // label rows (x+a, y+b)
    const RowBits<W> R00 = load_row<W>(L, nwd, dim, x, y), R01 = load_row<W>(L, nwd, dim, x, y + 1);
    const RowBits<W> R10 = load_row<W>(L, nwd, dim, x + 1, y), R11 = load_row<W>(L, nwd, dim, x + 1, y + 1);
    const bool tris = x < dim - 1 && y < dim - 1 && !quad_uniform<W>(R00, R01, R10, R11, dim);
    // edge masks of row (x, y)
    RowBits<W> EX00 = rxor<W>(R00, R10, x + 1 < dim), EY00 = rxor<W>(R00, R01, y + 1 < dim);
    RowBits<W> EZ00 = zedges<W>(R00, dim);
    const uint32_t nvr = total_pop<W>(EX00) + total_pop<W>(EY00) + total_pop<W>(EZ00);
    if (nvr == 0 && !tris) return;
    // edge masks of rows (x, y+1), (x+1, y), (x+1, y+1)
    RowBits<W> EX01 = zero_row<W>(), EY01 = zero_row<W>(), EZ01 = zero_row<W>();
    RowBits<W> EX10 = zero_row<W>(), EY10 = zero_row<W>(), EZ10 = zero_row<W>();
    RowBits<W> EX11 = zero_row<W>(), EY11 = zero_row<W>(), EZ11 = zero_row<W>();
    uint32_t vb00 = A.row_v[M.row_off + r], vb01 = 0, vb10 = 0, vb11 = 0;
    if (tris) {
        const RowBits<W> R02 = load_row<W>(L, nwd, dim, x, y + 2), R12 = load_row<W>(L, nwd, dim, x + 1, y + 2);
        const RowBits<W> R20 = load_row<W>(L, nwd, dim, x + 2, y), R21 = load_row<W>(L, nwd, dim, x + 2, y + 1);
        EX01 = rxor<W>(R01, R11, x + 1 < dim);
        EY01 = rxor<W>(R01, R02, y + 2 < dim);
        EZ01 = zedges<W>(R01, dim);
        EX10 = rxor<W>(R10, R20, x + 2 < dim);
        EY10 = rxor<W>(R10, R11, y + 1 < dim);
        EZ10 = zedges<W>(R10, dim);
        EX11 = rxor<W>(R11, R21, x + 2 < dim);
        EY11 = rxor<W>(R11, R12, y + 2 < dim);
        EZ11 = zedges<W>(R11, dim);
        // x, y < dim-1 here, so all four rows exist
        vb01 = A.row_v[M.row_off + r + 1];
        vb10 = A.row_v[M.row_off + r + dim];
        vb11 = A.row_v[M.row_off + r + dim + 1];
    }
    const int64_t vo = M.v_off, io = M.i_off;
    float* __restrict__ P = A.pos + vo * 3;
    uint8_t* __restrict__ B = A.boundary + vo;
    const float* __restrict__ D = A.dens + M.dens_off;
    const bool adjacency = M.iters > 0;
    uint32_t tcarry = A.row_t[M.row_off + r];
    for (int z0 = 0; z0 < dim; z0 += 64) {
        const int z = z0 + lane;
        // ---- vertices owned by sample (x, y, z): x, y, z edges in order ----
        if (z < dim) {
            const uint32_t fx = bit_at<W>(EX00, z), fy = bit_at<W>(EY00, z), fz = bit_at<W>(EZ00, z);
            if (fx | fy | fz) {
                uint32_t id = vb00 + prefix_pop<W>(EX00, z) + prefix_pop<W>(EY00, z) + prefix_pop<W>(EZ00, z);
                const float s0 = D[((int64_t)x * dim + y) * dim + z];
#pragma unroll
                for (int ax = 0; ax < 3; ax++) {
                    const uint32_t f = ax == 0 ? fx : (ax == 1 ? fy : fz);
                    if (!f) continue;
                    if (id >= (uint32_t)M.V) {  // counts and ids disagree: report, never write past the chunk
                        A.ctr->error = 2;
                        break;
                    }
                    const int x1 = x + (ax == 0), y1 = y + (ax == 1), z1 = z + (ax == 2);
                    const float s1 = D[((int64_t)x1 * dim + y1) * dim + z1];
                    // DMCChunk::calculate_isovertex / _get_intersection (DMCChunk.cpp:657-674)
                    const float mu = (0.0f - s0) / (s1 - s0);
                    const float dx = (float)x1 - (float)x, dy = (float)y1 - (float)y, dz = (float)z1 - (float)z;
                    P[3 * (int64_t)id + 0] = dx * mu + (float)x;
                    P[3 * (int64_t)id + 1] = dy * mu + (float)y;
                    P[3 * (int64_t)id + 2] = dz * mu + (float)z;
                    B[id] = (uint8_t)(x == 0 || y == 0 || z == 0 || x == dim - 1 || y == dim - 1 || z == dim - 1 ||
                                      x1 == dim - 1 || y1 == dim - 1 || z1 == dim - 1);
                    id++;
                }
            }
        }
        // ---- triangles of interior cell (x, y, z) ----
        if (tris) {
            const uint32_t mask = z < dim - 1 ? cell_mask<W>(R00, R01, R10, R11, z) : 0u;
            const uint32_t nt = s_ntri[mask];
            const uint64_t b0 = __ballot(nt & 1u), b1 = __ballot(nt & 2u), b2 = __ballot(nt & 4u);
            const uint32_t pre = (uint32_t)__popcll(b0 & lt) + 2u * (uint32_t)__popcll(b1 & lt) +
                                 4u * (uint32_t)__popcll(b2 & lt);
            if (nt) {
                // vertex ids of the 12 local edges (DMCChunk.cpp:543-565):
                // x-edges e = 2*dy + dz on rows (x, y+dy); y-edges e = 4 + 2*dx + dz
                // on rows (x+dx, y); z-edges e = 8 + 2*dx + dy on rows (x+dx, y+dy)
#define BMF_VIDS(EX, EY, EZ, VB, DX, DY)                                                          \
    {                                                                                            \
        const uint32_t p0 = VB + prefix_pop<W>(EX, z) + prefix_pop<W>(EY, z) + prefix_pop<W>(EZ, z); \
        const uint32_t ex0 = bit_at<W>(EX, z), ey0 = bit_at<W>(EY, z), ez0 = bit_at<W>(EZ, z);   \
        const uint32_t p1 = p0 + ex0 + ey0 + ez0;                                                \
        if (DX == 0) {                                                                           \
            s_vid[2 * DY + 0][lane] = p0;                                                        \
            s_vid[2 * DY + 1][lane] = p1;                                                        \
        }                                                                                        \
        if (DY == 0) {                                                                           \
            s_vid[4 + 2 * DX + 0][lane] = p0 + ex0;                                              \
            s_vid[4 + 2 * DX + 1][lane] = p1 + bit_at<W>(EX, z + 1);                             \
        }                                                                                        \
        s_vid[8 + 2 * DX + DY][lane] = p0 + ex0 + ey0;                                           \
    }
                BMF_VIDS(EX00, EY00, EZ00, vb00, 0, 0)
                BMF_VIDS(EX01, EY01, EZ01, vb01, 0, 1)
                BMF_VIDS(EX10, EY10, EZ10, vb10, 1, 0)
                BMF_VIDS(EX11, EY11, EZ11, vb11, 1, 1)
#undef BMF_VIDS
                const uint32_t tb = tcarry + pre;
                const uint64_t row = c_tri_table[mask];
                const int cnt = (int)(row >> 60);
                if (tb + (uint32_t)(cnt / 3) > (uint32_t)M.T) {
                    A.ctr->error = 3;
                    continue;
                }
                uint32_t* __restrict__ I = A.idx + io + 3ll * tb;
                for (int s = 0; s < cnt; s++) {
                    const int e = (int)((row >> (4 * s)) & 15ull);
                    const uint32_t v = s_vid[e][lane];
                    if (v >= (uint32_t)M.V) {
                        A.ctr->error = 4;
                        break;
                    }
                    I[s] = v;
                    if (adjacency) {
                        const uint32_t k = atomicAdd(&A.valence[vo + v], 1u);
                        if (k < BMF_MAX_VALENCE) A.adj[(vo + v) * BMF_MAX_VALENCE + k] = tb + (uint32_t)(s / 3);
                        else A.ctr->error = 1;
                    }
                }
            }
            tcarry += (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
        }
    }
}

__global__ void __launch_bounds__(kRowThreads) k_emit(Arenas A, int n) {
    __shared__ uint8_t s_ntri[256];
    __shared__ uint32_t s_vid[kRowThreads / 64][12][64];
    for (int t = threadIdx.x; t < 256; t += blockDim.x) s_ntri[t] = (uint8_t)((c_tri_table[t] >> 60) / 3);
    __syncthreads();
    const int wave = uniform_int(threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * (kRowThreads / 64);
    const int64_t total = A.ctr->total_rows;
    for (int64_t g = (int64_t)blockIdx.x * (kRowThreads / 64) + wave; g < total; g += nw) {
        const int j = find_chunk(A.row_prefix, n, g);
        const ChunkMeta& M = A.meta[j];
        if (M.status != BMF_OK || M.V == 0) continue;
        const int64_t r = g - A.row_prefix[j];
        switch ((M.dim + 63) >> 6) {
            case 1: emit_row<1>(A, M, r, s_ntri, s_vid[wave]); break;
            case 2: emit_row<2>(A, M, r, s_ntri, s_vid[wave]); break;
            case 3: emit_row<3>(A, M, r, s_ntri, s_vid[wave]); break;
            default: emit_row<4>(A, M, r, s_ntri, s_vid[wave]); break;
        }
    }
}

// ------------------------------------------------------------------------
// sort each vertex's adjacency list into ascending triangle id
// (MeshProcessor::init_primitives appends in triangle order, :114-127)
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kSmoothThreads) k_sort_adj(Arenas A, int n) {
    const int64_t total = A.ctr->total_vtiles;
    for (int64_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
        const int j = find_chunk(A.vtile_prefix, n, tile);
        const ChunkMeta& M = A.meta[j];
        const int64_t v = (tile - A.vtile_prefix[j]) * kSmoothThreads + threadIdx.x;
        if (v >= M.V) continue;
        const int64_t gv = M.v_off + v;
        const uint32_t cnt = min(A.valence[gv], (uint32_t)BMF_MAX_VALENCE);
        uint32_t* __restrict__ lst = A.adj + gv * BMF_MAX_VALENCE;
        uint32_t a[BMF_MAX_VALENCE];
#pragma unroll
        for (int k = 0; k < BMF_MAX_VALENCE; k++) a[k] = (uint32_t)k < cnt ? lst[k] : 0xFFFFFFFFu;
        // odd-even transposition sort with compile-time indices
#pragma unroll
        for (int rnd = 0; rnd < BMF_MAX_VALENCE; rnd++)
#pragma unroll
            for (int k = rnd & 1; k + 1 < BMF_MAX_VALENCE; k += 2) {
                const uint32_t lo = min(a[k], a[k + 1]), hi = max(a[k], a[k + 1]);
                a[k] = lo;
                a[k + 1] = hi;
            }
#pragma unroll
        for (int k = 0; k < BMF_MAX_VALENCE; k++)
            if ((uint32_t)k < cnt) lst[k] = a[k];
    }
}

// ------------------------------------------------------------------------
// smoothing: dual step per triangle, primal step per vertex
// (MeshProcessor.cpp:130-306), iteration m of chunks with iters > m
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kSmoothThreads) k_dual(Arenas A, int n, int m) {
    const int64_t total = A.ctr->total_ttiles;
    for (int64_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
        const int j = find_chunk(A.ttile_prefix, n, tile);
        const ChunkMeta& M = A.meta[j];
        if (M.iters <= m) continue;
        const int64_t t = (tile - A.ttile_prefix[j]) * kSmoothThreads + threadIdx.x;
        if (t >= M.T) continue;
        const uint32_t* __restrict__ I = A.idx + M.i_off + 3 * t;
        const float* __restrict__ P = A.pos + M.v_off * 3;
        const uint32_t i0 = I[0], i1 = I[1], i2 = I[2];
        if (i0 >= (uint32_t)M.V || i1 >= (uint32_t)M.V || i2 >= (uint32_t)M.V) {
            A.ctr->error = 5;
            continue;
        }
        float* __restrict__ C = A.cen + M.i_off + 3 * t;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float acc = 0.0f;
            acc = acc + P[3 * (int64_t)i0 + c];
            acc = acc + P[3 * (int64_t)i1 + c];
            acc = acc + P[3 * (int64_t)i2 + c];
            C[c] = acc / 3.0f;
        }
    }
}

__global__ void __launch_bounds__(kSmoothThreads) k_primal(Arenas A, int n, int m) {
    const int64_t total = A.ctr->total_vtiles;
    for (int64_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
        const int j = find_chunk(A.vtile_prefix, n, tile);
        const ChunkMeta& M = A.meta[j];
        if (M.iters <= m) continue;
        const int64_t v = (tile - A.vtile_prefix[j]) * kSmoothThreads + threadIdx.x;
        if (v >= M.V) continue;
        const int64_t gv = M.v_off + v;
        const uint32_t cnt = A.valence[gv];
        if (cnt == 0 || (!M.bp && A.boundary[gv])) continue;
        const uint32_t* __restrict__ lst = A.adj + gv * BMF_MAX_VALENCE;
        const float* __restrict__ C = A.cen + M.i_off;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        if (cnt > BMF_MAX_VALENCE) {
            A.ctr->error = 6;
            continue;
        }
        for (uint32_t k = 0; k < cnt; k++) {
            const int64_t t = lst[k];
            if (t >= M.T) {
                A.ctr->error = 7;
                break;
            }
            sx = sx + C[3 * t + 0];
            sy = sy + C[3 * t + 1];
            sz = sz + C[3 * t + 2];
        }
        const float fc = (float)cnt;
        float* __restrict__ P = A.pos + gv * 3;
        P[0] = sx / fc;
        P[1] = sy / fc;
        P[2] = sz / fc;
    }
}

// ------------------------------------------------------------------------
// finalize: per-chunk results for the host
// ------------------------------------------------------------------------
__global__ void k_finalize(Arenas A, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const ChunkMeta& M = A.meta[j];
    bmf_chunk_result r;
    const bool ok = M.status == BMF_OK;
    r.vertex_offset = ok ? M.v_off : 0;
    r.index_offset = ok ? M.i_off : 0;
    r.vertex_count = ok ? M.V : 0;
    r.index_count = ok ? 3 * M.T : 0;
    r.contains_mesh = M.contains_mesh;
    r.status = M.status;
    r.sampled = M.state == kStateSampled;
    r.delta = M.delta;
    r.overlap_pos[0] = M.opos[0];
    r.overlap_pos[1] = M.opos[1];
    r.overlap_pos[2] = M.opos[2];
    A.results[j] = r;
}

// ------------------------------------------------------------------------
// stage inspection in the reference's layouts (chunk 0 of the last plan)
// ------------------------------------------------------------------------
// DMCChunk::label_grid words [x][y][ceil(dim/32)], bit z%32 (DMCChunk.cpp:143)
__global__ void k_ref_labels(Arenas A, uint32_t* words) {
    const ChunkMeta& M = A.meta[0];
    const int dim = M.dim, zc = (dim + 31) / 32;
    const int64_t nwords = (int64_t)dim * dim * zc;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = w / zc;
        const int zb = (int)(w - row * zc);
        const int x = (int)(row / dim), y = (int)(row - (int64_t)x * dim);
        uint32_t v = 0;
        if (M.state == kStateSampled) {
            const RowBits<kMaxRowWords> rb = load_row<kMaxRowWords>(A.labels + M.lab_off, M.lab_words, dim, x, y);
            const uint64_t q = word_at<kMaxRowWords>(rb, zb >> 1);
            v = (uint32_t)(q >> (32 * (zb & 1)));
        } else if (M.state == kStateAir) {
            const int nb = min(32, dim - 32 * zb);
            v = nb >= 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u);
        }
        words[w] = v;
    }
}

// DMCChunk::label_edges masks [x][y][ceil(dim/8)], byte z%8 (DMCChunk.cpp:200-438)
__global__ void k_ref_masks(Arenas A, uint64_t* masks) {
    const ChunkMeta& M = A.meta[0];
    const int dim = M.dim, z8 = (dim + 7) / 8;
    const int64_t nwords = (int64_t)dim * dim * z8;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = w / z8;
        const int zw = (int)(w - row * z8);
        const int x = (int)(row / dim), y = (int)(row - (int64_t)x * dim);
        constexpr int W = kMaxRowWords;
        RowBits<W> r00 = zero_row<W>(), r01 = zero_row<W>(), r10 = zero_row<W>(), r11 = zero_row<W>();
        if (M.state == kStateSampled) {
            const uint64_t* L = A.labels + M.lab_off;
            r00 = load_row<W>(L, M.lab_words, dim, x, y);
            r01 = load_row<W>(L, M.lab_words, dim, x, y + 1);
            r10 = load_row<W>(L, M.lab_words, dim, x + 1, y);
            r11 = load_row<W>(L, M.lab_words, dim, x + 1, y + 1);
        } else if (M.state == kStateAir) {
            RowBits<W> full = zero_row<W>();
            for (int k = 0; k < W; k++) {
                const int nb = dim - 64 * k;
                full.w[k] = nb >= 64 ? ~0ull : (nb > 0 ? (1ull << nb) - 1ull : 0ull);
            }
            r00 = full;
            if (y + 1 < dim) r01 = full;
            if (x + 1 < dim) r10 = full;
            if (x + 1 < dim && y + 1 < dim) r11 = full;
        }
        uint64_t v = 0;
        for (int b = 0; b < 8; b++) {
            const int z = 8 * zw + b;
            if (z >= dim) break;
            v |= (uint64_t)cell_mask<W>(r00, r01, r10, r11, z) << (8 * b);
        }
        masks[w] = v;
    }
}

// ------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------
hipError_t launch_plan(const Arenas& A, int n, int allow_cull, int reset_bumps, hipStream_t s) {
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, s, A, n, allow_cull, reset_bumps);
    return hipGetLastError();
}
hipError_t launch_sample(const Arenas& A, int n, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_sample, dim3(blocks), dim3(kSampleThreads), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_count(const Arenas& A, int n, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_count, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_scan(const Arenas& A, int n, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_scan, dim3(blocks), dim3(kScanThreads), 0, s, A, n);
    hipLaunchKernelGGL(k_plan_smooth, dim3(1), dim3(kPlanThreads), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_emit(const Arenas& A, int n, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_emit, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_sort_adj(const Arenas& A, int n, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_adj, dim3(blocks), dim3(kSmoothThreads), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_dual(const Arenas& A, int n, int m, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_dual, dim3(blocks), dim3(kSmoothThreads), 0, s, A, n, m);
    return hipGetLastError();
}
hipError_t launch_primal(const Arenas& A, int n, int m, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_primal, dim3(blocks), dim3(kSmoothThreads), 0, s, A, n, m);
    return hipGetLastError();
}
hipError_t launch_finalize(const Arenas& A, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3((n + 255) / 256), dim3(256), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_ref_labels(const Arenas& A, uint32_t* words, hipStream_t s) {
    hipLaunchKernelGGL(k_ref_labels, dim3(512), dim3(256), 0, s, A, words);
    return hipGetLastError();
}
hipError_t launch_ref_masks(const Arenas& A, uint64_t* masks, hipStream_t s) {
    hipLaunchKernelGGL(k_ref_masks, dim3(512), dim3(256), 0, s, A, masks);
    return hipGetLastError();
}

}  // namespace bmf
