// HIP kernels of the chunk-extraction path (gfx950, wave64).
//
// Per batch:  plan -> sample -> count -> scan -> emit -> smooth
// (the per-chunk results are written by plan for culled or skipped chunks
// and by scan for sampled ones, so no launch follows the smoothing)
//
// Data layout in HBM (per sampled chunk, offsets from the plan kernel):
//   densities   FP32 [x][y][z]                    (DMCChunk density_block)
//   labels      flat bitmap over the sample index (x*dim + y)*dim + z,
//               u64 word i/64, bit i%64, set when the density is < 0;
//               one word per wave of 64 consecutive samples (one ballot)
//   row_v/row_t per label row (x, y): vertex / triangle counts, then
//               exclusive bases after the scan
//   cell_tb/    per surface cell (same index as the density): first triangle
//   cell_mask   id and corner mask, written only for chunks that smooth
//   output      positions float3, boundary u8, vertex edge code u32,
//               chunk-local triangle indices
// Cell masks and the reference's u32 label words are never materialised on
// the hot path: the plane kernels rebuild them from the bitmap with shifts.
#include <hip/hip_runtime.h>

#include "bmf_device_tables.h"
#include "bmf_internal.h"
#include "bmf_rows.h"
#include "samplers.h"

namespace bmf {

// ------------------------------------------------------------------------
// plan: geometry, culling, arena offsets, counters (one block; three passes
// over many blocks for batches beyond one block)
// ------------------------------------------------------------------------
constexpr int kPlanThreads = 1024;
// Up to this many descriptors the plans run in one block (a few per thread,
// one launch each); larger batches take the multi-block passes, whose extra
// launches cost more than they save below it.
constexpr int kPlanOneBlock = 2 * kPlanThreads;

// The host-visible result of chunk j from its final record: written by
// k_plan for the chunks it culls or skips and by k_scan for the sampled ones
// (every field is final there: later kernels report errors through the
// counters, not the chunk records).
__device__ __forceinline__ void write_result(const Arenas& A, int j, const ChunkMeta& M) {
    bmf_chunk_result r;
    const bool ok = M.status == BMF_OK;
    r.vertex_offset = ok ? M.v_off : 0;
    r.index_offset = ok ? M.i_off : 0;
    r.vertex_count = ok ? M.V : 0;
    r.index_count = ok ? M.ipp * M.T : 0;
    r.contains_mesh = M.contains_mesh;
    r.status = M.status;
    r.sampled = M.state == kStateSampled;
    r.delta = M.delta;
    r.overlap_pos[0] = M.opos[0];
    r.overlap_pos[1] = M.opos[1];
    r.overlap_pos[2] = M.opos[2];
    A.results[j] = r;
}

// a[i] of a register array by unrolled selects (a runtime index into a
// register array would place it in scratch memory)
template <int N>
__device__ __forceinline__ int64_t pick(const int64_t (&a)[N], int i) {
    int64_t r = 0;
#pragma unroll
    for (int k = 0; k < N; k++) r = k == i ? a[k] : r;
    return r;
}

// Plan record of one descriptor: DMCChunk::label_grid geometry
// (DMCChunk.cpp:94-101), analytic culling, sizes; `add` gets its
// contributions to the four arena sums (samples, label words, rows, sampling
// tiles) and to the x-planes of its row-word class (add[4 + class]).
__device__ __forceinline__ ChunkMeta plan_record(const Arenas& A, const SamplerDev* __restrict__ S,
                                                 const bmf_chunk_desc& d, int allow_cull, int64_t (&add)[8]) {
    ChunkMeta m;
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
    } else if (allow_cull && S->kind == BMF_SAMPLER_NOISE && S->cull) {
        // d = (-dy) - n h with dy = (y g) ym monotone in y and n in
        // [nlo, nhi]: every FP32 step is monotone, so the extremes of d
        // over the chunk come from the end values
        const float ylo = m.opos[1];
        const float yhi = (float)(d.dim - 1) * m.delta + m.opos[1];
        const float dy0 = (ylo * S->g_scale) * S->y_scale, dy1 = (yhi * S->g_scale) * S->y_scale;
        const float p0 = S->nlo * S->height, p1 = S->nhi * S->height;
        const float dmax = (-fminf(dy0, dy1)) - fminf(p0, p1);
        const float dmin = (-fmaxf(dy0, dy1)) - fmaxf(p0, p1);
        if (dmax < 0.0f) m.state = kStateAir;
        else if (dmin >= 0.0f) m.state = kStateSolid;
    }
    m.csg_ns = m.csg_nb = 0;  // k_csg_cull fills them
    const int64_t ns = (int64_t)d.dim * d.dim * d.dim;
    const bool sampled = m.state == kStateSampled;
    m.lab_words = sampled ? (ns + 63) / 64 : 0;
    m.any_air = m.state == kStateAir;
    m.any_solid = m.state == kStateSolid;
    m.contains_mesh = 0;
    m.status = BMF_OK;
    m.V = 0;
    m.T = 0;
    m.ipp = (A.mesh_mode == BMF_MESH_QUADS && !A.quad_tris) ? 4 : 3;
    m.v_off = 0;
    m.i_off = 0;
    m.dens_off = 0;
    m.lab_off = 0;
    m.row_off = 0;
    m.plane_slot = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) add[q] = 0;
    add[0] = sampled ? ns : 0;
    add[1] = m.lab_words;
    add[2] = sampled ? (int64_t)d.dim * d.dim : 0;
    add[3] = sampled ? (ns + A.sample_tile - 1) / A.sample_tile : 0;
    const int cls = ((d.dim + 63) >> 6) - 1;
#pragma unroll
    for (int q = 0; q < 4; q++) add[4 + q] = sampled && q == cls ? d.dim : 0;
    return m;
}

__device__ __forceinline__ bool plan_overflow(const Arenas& A, const int64_t (&tot)[8]) {
    return tot[0] > A.cap_samples || tot[1] > A.cap_samples / 64 + A.cap_chunks || tot[2] > A.cap_rows;
}

// A record's final state once its offsets are known (base: where its block's
// arena ranges start; cbase: its row-word class's first plane slot).
__device__ __forceinline__ void plan_finish(const Arenas& A, int j, ChunkMeta& m, const int64_t (&base)[8],
                                            const int64_t (&cbase)[4], bool overflow) {
    m.dens_off += base[0];
    m.lab_off += base[1];
    m.row_off += base[2];
    if (m.state == kStateSampled) {
        const int q = ((m.dim + 63) >> 6) - 1;
        int64_t cb = 0, bb = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            cb = k == q ? cbase[k] : cb;
            bb = k == q ? base[4 + k] : bb;
        }
        m.plane_slot += cb + bb;
        if (overflow) {
            m.state = kStateSkipped;
            m.status = BMF_ECAPACITY;
        }
    } else {
        // culled chunks: DMCChunk.cpp:159-162 on uniform labels
        m.contains_mesh = (m.state == kStateAir) && (m.dim % 32 != 0);
    }
    A.meta[j] = m;
    if (m.state != kStateSampled) write_result(A, j, m);  // final: k_scan skips it
}

// The batch counters from the totals: sampling tiles, overflow, and per
// row-word class its planes in plane_map and the row segments per plane
// (few sampled planes, e.g. single large chunks: split each plane's rows so
// that the class still has kSplitTarget items).
__device__ __forceinline__ void plan_counters(const Arenas& A, int n, const int64_t (&tot)[8], bool overflow,
                                              int reset_bumps) {
    A.tile_prefix[n] = overflow ? 0 : tot[3];
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
    c->total_tiles = overflow ? 0 : tot[3];
    int64_t cb = 0;
    for (int q = 0; q < 4; q++) {
        const int64_t np = overflow ? 0 : tot[4 + q];
        int split = A.row_split > 0 ? A.row_split : 1;
        if (A.row_split <= 0)
            while (split < kMaxPlaneSplit && np * split < kSplitTarget) split <<= 1;
        c->plane_split[q] = split;
        c->class_base[q] = cb;
        c->class_planes[q] = np;
        cb += np;
    }
    c->n_smooth = 0;
    c->total_samples = tot[0];
}

// Batches of up to kPlanOneBlock descriptors: one block, each thread a
// segment of consecutive descriptors, one launch.
__global__ void __launch_bounds__(kPlanThreads) k_plan(Arenas A, int n, int allow_cull, int reset_bumps) {
    __shared__ int64_t s_w[8][kPlanThreads / 64 + 1];
    const int per = (n + kPlanThreads - 1) / kPlanThreads;
    const int j0 = threadIdx.x * per, j1 = min(n, j0 + per);
    int64_t sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = j0; j < j1; j++) {
        int64_t add[8];
        ChunkMeta m = plan_record(A, A.sampler, A.descs[j], allow_cull, add);
        // offsets within the thread's segment; tile_prefix likewise
        m.dens_off = sum[0];
        m.lab_off = sum[1];
        m.row_off = sum[2];
        A.tile_prefix[j] = sum[3];
        if (m.state == kStateSampled) m.plane_slot = pick<8>(sum, 4 + ((m.dim + 63) >> 6) - 1);
        A.meta[j] = m;
#pragma unroll
        for (int q = 0; q < 8; q++) sum[q] += add[q];
    }
    int64_t excl[8], tot[8];
    block_scan<8, kPlanThreads>(sum, excl, tot, s_w);
    const bool overflow = plan_overflow(A, tot);
    // plane_map slots: the classes one after another, chunk order inside a class
    int64_t cbase[4], cb = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        cbase[q] = cb;
        cb += tot[4 + q];
    }
    for (int j = j0; j < j1; j++) {
        A.tile_prefix[j] = overflow ? 0 : A.tile_prefix[j] + excl[3];
        ChunkMeta m = A.meta[j];
        plan_finish(A, j, m, excl, cbase, overflow);
    }
    if (threadIdx.x == 0) plan_counters(A, n, tot, overflow, reset_bumps);
}

// Larger batches, pass 1 of 3 (one descriptor per thread over many blocks):
// records with offsets relative to their block, and each block's totals.
__global__ void __launch_bounds__(kPlanThreads) k_plan_part(Arenas A, int n, int allow_cull) {
    __shared__ int64_t s_w[8][kPlanThreads / 64 + 1];
    const int j = blockIdx.x * kPlanThreads + threadIdx.x;
    int64_t add[8];
    ChunkMeta m{};
    if (j < n) m = plan_record(A, A.sampler, A.descs[j], allow_cull, add);
    else
        for (int q = 0; q < 8; q++) add[q] = 0;
    int64_t excl[8], tot[8];
    block_scan<8, kPlanThreads>(add, excl, tot, s_w);
    if (j < n) {
        A.tile_prefix[j] = excl[3];
        m.dens_off = excl[0];
        m.lab_off = excl[1];
        m.row_off = excl[2];
        if (m.state == kStateSampled) m.plane_slot = pick<8>(excl, 4 + ((m.dim + 63) >> 6) - 1);
        A.meta[j] = m;
    }
    if (threadIdx.x < 8) A.plan_part[8 * blockIdx.x + threadIdx.x] = tot[threadIdx.x];
}

// Pass 2 (one block, one thread per pass-1 block, at most kPlanThreads):
// block bases, batch totals, overflow flag, counters.  Bases go to
// plan_part[8 (nb + b) ...], the overflow flag after them.
__global__ void __launch_bounds__(kPlanThreads) k_plan_total(Arenas A, int n, int nb, int reset_bumps) {
    __shared__ int64_t s_w[8][kPlanThreads / 64 + 1];
    const int b = threadIdx.x;
    int64_t v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = b < nb ? A.plan_part[8 * b + q] : 0;
    int64_t excl[8], tot[8];
    block_scan<8, kPlanThreads>(v, excl, tot, s_w);
    const bool overflow = plan_overflow(A, tot);
    if (b < nb)
#pragma unroll
        for (int q = 0; q < 8; q++) A.plan_part[8 * (nb + b) + q] = excl[q];
    if (b == 0) {
        A.plan_part[16 * nb] = overflow;
        int64_t cb = 0;
        for (int q = 0; q < 4; q++) {
            A.plan_part[16 * nb + 1 + q] = cb;
            cb += tot[4 + q];
        }
        plan_counters(A, n, tot, overflow, reset_bumps);
    }
}

// Pass 3: final offsets, sampling-tile prefixes, plane slots, the results
// of culled chunks.
__global__ void __launch_bounds__(kPlanThreads) k_plan_fix(Arenas A, int n, int nb) {
    const int j = blockIdx.x * kPlanThreads + threadIdx.x;
    if (j >= n) return;
    int64_t base[8], cbase[4];
#pragma unroll
    for (int q = 0; q < 8; q++) base[q] = A.plan_part[8 * (nb + blockIdx.x) + q];
    const bool overflow = A.plan_part[16 * nb] != 0;
#pragma unroll
    for (int q = 0; q < 4; q++) cbase[q] = A.plan_part[16 * nb + 1 + q];
    A.tile_prefix[j] = overflow ? 0 : A.tile_prefix[j] + base[3];
    ChunkMeta m = A.meta[j];
    plan_finish(A, j, m, base, cbase, overflow);
}

// ------------------------------------------------------------------------
// CSG culling: per chunk, the primitives whose value can exceed the floor
// inside its sample box (spheres within r + |floor| of the box, boxes whose
// box grown by |floor| meets it; 0.5 of slack covers FP32 rounding), in
// their original order (sphere ids, then box ids).  Block per chunk; order
// kept by a ballot prefix per 256 primitives.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_csg_cull(Arenas A, int n) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const SamplerDev* __restrict__ S = A.sampler;
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        ChunkMeta& M = A.meta[j];
        const float lo[3] = {M.opos[0], M.opos[1], M.opos[2]};
        const float ext = (float)(M.dim - 1) * M.delta;
        const float reach = -kCsgFloor + 0.5f;
        uint16_t* L = A.csg_list + (int64_t)j * A.csg_stride;
        if (threadIdx.x == 0) s_base = 0;
        __syncthreads();
        for (int pass = 0; pass < 2; pass++) {
            const int np = pass == 0 ? S->n_spheres : S->n_boxes;
            for (int i0 = 0; i0 < np; i0 += 256) {
                const int i = i0 + (int)threadIdx.x;
                bool keep = false;
                if (i < np) {
                    if (pass == 0) {
                        const float* sp = S->spheres + 4 * i;
                        float dd = 0.0f;
                        for (int a = 0; a < 3; a++) {
                            const float t = fmaxf(fmaxf(lo[a] - sp[a], sp[a] - (lo[a] + ext)), 0.0f);
                            dd += t * t;
                        }
                        const float r = sp[3] + reach;
                        keep = dd < r * r;
                    } else {
                        const float* b = S->boxes + 6 * i;
                        keep = true;
                        for (int a = 0; a < 3; a++)
                            keep = keep && (b[a] - b[3 + a] - reach <= lo[a] + ext) && (b[a] + b[3 + a] + reach >= lo[a]);
                    }
                }
                const uint64_t bal = __ballot(keep);
                if (lane == 0) s_wave[wave] = (uint32_t)__popcll(bal);
                __syncthreads();
                uint32_t off = s_base;
                for (int w = 0; w < wave; w++) off += s_wave[w];
                if (keep) L[off + (uint32_t)__popcll(bal & lt)] = (uint16_t)i;
                __syncthreads();
                if (threadIdx.x == 0) s_base += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                if (pass == 0) M.csg_ns = (int32_t)s_base;
                else M.csg_nb = (int32_t)s_base - M.csg_ns;
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------------
// sample: persistent; workgroups pull tiles of Arenas::sample_tile consecutive
// samples of one chunk from an atomic work queue.
//
// Terrain chunks use exact adaptive octave evaluation.  After k octaves the
// FP32 partial sum S_k bounds the final sum: each remaining step is
// fma(a_o, n_o, S) with |n_o| <= 1 and rounding is monotone, so
// S_O in [L, H] with L = (..(S_k - |a_k|) - ..), H = (..(S_k + |a_k|) + ..)
// evaluated in FP32; every value involved is at most A = sum |a_o| in
// magnitude, so L and H differ from the exact S_k -/+ sum_{o>=k} |a_o| by at
// most (O - k) half-ulps of A, and reach[k] = that sum + (O - k + 2) ulp(A)
// (rounded up, host side) gives FP32 bounds S_k -/+ reach[k] outside [L, H].
// The density is monotone in S, so d(S_k - reach), d(S_k + reach) bound the
// final density and decide the label once they agree in sign.  Undecided
// samples are compacted into an LDS queue (wave ballot + popcount) so each
// octave runs on dense waves.  Samples that run all octaves get their exact
// density written and their bit set in the `exact` bitmap; the others only
// get their label (k_count queues the ones that are endpoints of a
// sign-changing edge and k_refine finishes their densities).  Without
// adaptation every
// sample is evaluated fully (bit-identical values either way).
// ------------------------------------------------------------------------
// The terrain terms of the sampler and a chunk's sample geometry as values:
// read once per tile, so the loops below never reload them after their
// own stores.
struct TerrainTerms {
    float amp, y0, inv_norm;
};
struct ChunkGeom {
    float ox, oy, oz, delta;
};

__device__ __forceinline__ float density_from_sum(const TerrainTerms& S, float sum, float t2) {
    float fb = sum * S.inv_norm;
    fb = fminf(fmaxf(fb, -1.0f), 1.0f);
    return S.amp * fb - t2;
}

// 0: undecided, 1: certainly air (d < 0), 2: certainly solid.  The final
// FP32 sum lies in [sum - reach[k], sum + reach[k]] (see SamplerDev::reach:
// the interval the remaining fma steps can reach, widened by their rounding
// and that of the two operations here, so both ends are safe bounds).
__device__ __forceinline__ int decide_label(const TerrainTerms& S, float reach, float sum, float t2) {
    const float lo = sum - reach, hi = sum + reach;
    const float d0 = density_from_sum(S, lo, t2), d1 = density_from_sum(S, hi, t2);
    if (fmaxf(d0, d1) < 0.0f) return 1;
    if (!(fminf(d0, d1) < 0.0f)) return 2;
    return 0;
}

__device__ __forceinline__ void sample_index(uint32_t i, uint32_t dim, uint32_t dim2, float rd, float rd2, uint32_t& x,
                                             uint32_t& y, uint32_t& z) {
    x = udiv24(i, dim2, rd2);
    const uint32_t rem = i - x * dim2;
    y = udiv24(rem, dim, rd);
    z = rem - y * dim;
}

// NoiseSampler.cpp:21-27: (float)i * scale + p, two roundings
__device__ __forceinline__ void sample_position(const ChunkGeom& G, uint32_t x, uint32_t y, uint32_t z, float& px,
                                                float& py, float& pz) {
    px = (float)x * G.delta + G.ox;
    py = (float)y * G.delta + G.oy;
    pz = (float)z * G.delta + G.oz;
}

__device__ __forceinline__ void sample_coords(const ChunkGeom& G, uint32_t i, uint32_t dim, uint32_t dim2, float rd,
                                              float rd2, float& px, float& py, float& pz) {
    uint32_t x, y, z;
    sample_index(i, dim, dim2, rd, rd2, x, y, z);
    sample_position(G, x, y, z, px, py, pz);
}

// Sampling modes, fixed per launch by the host: adaptive octaves for the
// terrain sampler; the CSG union with the chunk's culled primitives staged in
// LDS once per tile; a 2D noise field with the noise of the tile's (x, z)
// columns evaluated once into LDS and broadcast along y; every other case
// (external densities, implicit shapes, 3D noise fields, the terrain in full)
// evaluated directly.
constexpr int kSampleGeneric = 0, kSampleAdaptive = 1, kSampleCsg = 2, kSampleSlab = 3;
// Adaptive mode: undecided samples of a wave (of 64) from which phase 0
// evaluates octave 0 in place rather than queueing them: queueing costs a
// push, a pop and the sample coordinates again, in-place evaluation the
// idle lanes of one octave.
constexpr int kInPlaceOctave = 52;
// Octave queue entries: sample index within the tile (kTile <= 4096) in the
// low 12 bits, the sample's next octave (< BMF_MAX_OCTAVES = 16) above.
constexpr int kQueueOctaveShift = 12;
constexpr uint32_t kQueueSampleMask = (1u << kQueueOctaveShift) - 1u;
static_assert(16 * kSampleThreads <= (1 << kQueueOctaveShift), "tile index fits the queue entry");
static_assert(BMF_MAX_OCTAVES <= 16, "octave index fits the queue entry");

template <int ROUNDS, int MODE>
__global__ void __launch_bounds__(kSampleThreads) k_sample(Arenas A, int n) {
    constexpr int kTile = ROUNDS * kSampleThreads;
    constexpr bool adapt = MODE == kSampleAdaptive;
    constexpr int kQ = adapt ? kTile : 1;                       // octave queues (adaptive mode only)
    constexpr int kC = MODE == kSampleCsg ? 2 * kCsgStage : 1;  // staged CSG primitives
    // 2D noise columns of a tile: its samples span at most kTile / dim^2 + 2
    // x-planes of dim columns, i.e. at most kTile / dim + 2 dim <= kTile / 2 + 2 BMF_MAX_DIM
    constexpr int kSlab = MODE == kSampleSlab ? kTile / 2 + 2 * BMF_MAX_DIM : 1;
    __shared__ float s_slab[kSlab];
    __shared__ int64_t s_tile;
    __shared__ uint32_t s_qn[2];
    __shared__ uint16_t s_qi[2][kQ];
    __shared__ float s_qs[2][kQ];
    __shared__ float4 s_csg[kC];
    __shared__ unsigned long long s_lab[kTile / 64];
    __shared__ unsigned long long s_exact[kTile / 64];
    __shared__ uint32_t s_flags;
    const SamplerDev* __restrict__ S = A.sampler;
    const TerrainTerms TT{S->amp, S->y0, S->inv_norm};
    const float reach0 = S->reach[0];
    const int O = S->octaves;
    const uint32_t seed0 = S->seed;  // octave k uses seed0 + k
    // adaptive mode: per-octave frequency, amplitude and reach (visible to
    // every thread after the first barrier of the tile loop)
    __shared__ float s_oct[3][BMF_MAX_OCTAVES + 1];
    if constexpr (adapt) {
        for (int t = threadIdx.x; t <= O; t += blockDim.x) {
            s_oct[0][t] = t < O ? S->fo[t] : 0.0f;
            s_oct[1][t] = t < O ? S->ao[t] : 0.0f;
            s_oct[2][t] = S->reach[t];
        }
    }
    const int64_t total = A.ctr->total_tiles;
    const int lane = lane_id();
    const uint64_t lt = lanemask_lt();
    unsigned long long evals = 0;  // adaptive mode: octaves this wave evaluated (wave-uniform)
    __shared__ unsigned long long s_evals;
    if (threadIdx.x == 0) s_evals = 0;  // ordered before its use by the tile loop's barriers
    // (chunk, x) of every x-plane item of the plane kernels that follow
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        const ChunkMeta& M = A.meta[j];
        if (M.state != kStateSampled) continue;
        const int64_t ps = M.plane_slot;
        for (int x = threadIdx.x; x < M.dim; x += blockDim.x) A.plane_map[ps + x] = make_int2(j, x);
    }
    for (;;) {
        if (threadIdx.x == 0) {
            s_tile = (int64_t)atomicAdd(&A.ctr->sample_head, 1ull);
            s_qn[0] = 0;
            s_qn[1] = 0;
            s_flags = 0;
        }
        __syncthreads();
        const int64_t tile = s_tile;
        if (tile >= total) break;
        const int j = find_chunk(A.tile_prefix, n, tile);
        const ChunkMeta& M = A.meta[j];
        const ChunkGeom G{M.opos[0], M.opos[1], M.opos[2], M.delta};
        const uint32_t dim = (uint32_t)M.dim, dim2 = dim * dim, ns = dim2 * dim;
        const float rd = 1.0f / (float)dim, rd2 = 1.0f / (float)dim2;
        float* __restrict__ dens = A.dens + M.dens_off;
        const uint32_t first = (uint32_t)(tile - A.tile_prefix[j]) * kTile;
        bool staged = false;
        if constexpr (MODE == kSampleCsg) {
            staged = M.csg_ns + M.csg_nb <= kCsgStage;  // uniform across the block
            if (staged) {
                stage_csg(*S, A.csg_list + (int64_t)j * A.csg_stride, M.csg_ns, M.csg_nb, s_csg);
                __syncthreads();
            }
        }
        uint32_t slab_x0 = 0;
        if constexpr (MODE == kSampleSlab) {
            const uint32_t last = min(first + (uint32_t)kTile, ns) - 1u;
            slab_x0 = first / dim2;
            const uint32_t ncol = (last / dim2 - slab_x0 + 1u) * dim;
            if (ncol > (uint32_t)kSlab) {
                if (threadIdx.x == 0) A.ctr->error = 12;  // unreachable by the bound above
            } else {
                for (uint32_t q = threadIdx.x; q < ncol; q += blockDim.x) {
                    const uint32_t qx = q / dim, qz = q - qx * dim;
                    float px, py, pz;
                    sample_position(G, slab_x0 + qx, 0u, qz, px, py, pz);
                    s_slab[q] = noise_value(*S, px, 0.0f, pz);
                }
            }
            __syncthreads();
        }
        // ---- phase 0: every sample of the tile ----
        // A thread's sample advances by kSampleThreads per round: its (x, y, z)
        // advances by (0, q, r) with q, r = kSampleThreads / dim, % dim and at
        // most one carry per axis while q + 1 < dim; smaller chunks divide.
        const uint32_t step_y = kSampleThreads / dim, step_z = kSampleThreads - step_y * dim;
        const bool stepping = step_y + 1 < dim;
        uint32_t cx, cy, cz;
        sample_index(first + threadIdx.x, dim, dim2, rd, rd2, cx, cy, cz);
#pragma unroll 1
        for (int r = 0; r < ROUNDS; r++) {
            const uint32_t li = r * kSampleThreads + threadIdx.x;  // waves cover 64 consecutive samples
            const uint32_t i = first + li;
            const bool valid = i < ns;
            if (r > 0) {
                if (stepping) {
                    cz += step_z;
                    cy += step_y;
                    if (cz >= dim) {
                        cz -= dim;
                        cy++;
                    }
                    if (cy >= dim) {
                        cy -= dim;
                        cx++;
                    }
                } else {
                    sample_index(i, dim, dim2, rd, rd2, cx, cy, cz);
                }
            }
            bool air = false, done = true, fin = false;
            int kq = 0;         // adaptive: the next octave of the wave's undecided samples
            float sum0 = 0.0f;  // and their partial sum
            if constexpr (adapt) {
                float px = 0.0f, py = 0.0f, pz = 0.0f, t2 = 0.0f;
                int dec = 2;
                if (valid) {
                    sample_position(G, cx, cy, cz, px, py, pz);
                    t2 = py - TT.y0;
                    dec = decide_label(TT, reach0, 0.0f, t2);
                }
                // while nearly all of the wave's samples are undecided it
                // evaluates their octaves here, as the rounds would (same
                // operations); the survivors of a sparse wave are queued with
                // their next octave, so that the rounds run on dense waves
                for (;;) {
                    const uint32_t live = (uint32_t)__popcll(__ballot(valid && dec == 0));
                    if (kq >= O || live < (uint32_t)kInPlaceOctave) break;
                    evals += live;
                    if (valid && dec == 0) {
                        const float f = s_oct[0][kq];
                        sum0 = __builtin_fmaf(s_oct[1][kq], snoise3(px * f, py * f, pz * f, seed0 + (uint32_t)kq), sum0);
                        if (kq == O - 1) {
                            const float d = density_from_sum(TT, sum0, t2);
                            dens[i] = d;
                            dec = d < 0.0f ? 1 : 2;
                            fin = true;
                        } else {
                            dec = decide_label(TT, s_oct[2][kq + 1], sum0, t2);
                        }
                    }
                    kq++;
                }
                done = dec != 0;
                air = dec == 1;
            } else if (valid) {
                float px, py, pz;
                sample_position(G, cx, cy, cz, px, py, pz);
                float d;
                if constexpr (MODE == kSampleSlab) {
                    const uint32_t q = (cx - slab_x0) * dim + cz;
                    d = noise_density_from(*S, q < (uint32_t)kSlab ? s_slab[q] : 0.0f, py);
                } else if (staged) d = csg_density_lds(s_csg, M.csg_ns, M.csg_nb, px, py, pz);
                else if (S->kind == BMF_SAMPLER_EXTERNAL) d = S->ext[A.ext_base + M.dens_off + i];
                else if (S->kind == BMF_SAMPLER_CSG && A.csg_list)
                    d = csg_density_list(*S, A.csg_list + (int64_t)j * A.csg_stride, M.csg_ns, M.csg_nb, px, py, pz);
                else d = sample_density(*S, px, py, pz);
                dens[i] = d;
                air = d < 0.0f;
            }
            const uint64_t bl = __ballot(valid && done && air);
            const uint64_t push = __ballot(valid && !done);
            const uint64_t vb = __ballot(valid);
            const uint64_t fb = __ballot(fin);
            if (lane == 0) {
                s_lab[li >> 6] = bl;
                s_exact[li >> 6] = adapt ? fb : vb;
            }
            if constexpr (adapt) {
                if (push) {
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&s_qn[0], (uint32_t)__popcll(push));
                    base = __shfl(base, 0);
                    if (valid && !done) {
                        const uint32_t q = base + (uint32_t)__popcll(push & lt);
                        s_qi[0][q] = (uint16_t)(li | ((uint32_t)kq << kQueueOctaveShift));
                        s_qs[0][q] = sum0;
                    }
                }
            }
        }
        __syncthreads();
        // ---- rounds on the compacted queue: each entry (sample, next octave,
        // partial sum) advances one octave per round, so O rounds finish all ----
        if constexpr (adapt) {
            for (int rnd = 0; rnd < O; rnd++) {
                const int cur = rnd & 1, nxt = cur ^ 1;
                const uint32_t nq = s_qn[cur];
                if (nq == 0) break;
                for (uint32_t e0 = 0; e0 < nq; e0 += kSampleThreads) {
                    const uint32_t e = e0 + threadIdx.x;
                    const bool active = e < nq;
                    uint32_t li = 0, k = 0;
                    float sum = 0.0f;
                    int dec = 2;
                    if (active) {
                        const uint32_t qv = s_qi[cur][e];
                        li = qv & kQueueSampleMask;
                        k = qv >> kQueueOctaveShift;
                        sum = s_qs[cur][e];
                        float px, py, pz;
                        sample_coords(G, first + li, dim, dim2, rd, rd2, px, py, pz);
                        const float t2 = py - TT.y0;
                        const float f = s_oct[0][k];
                        const float nz = snoise3(px * f, py * f, pz * f, seed0 + k);
                        sum = __builtin_fmaf(s_oct[1][k], nz, sum);
                        if ((int)k == O - 1) {
                            const float d = density_from_sum(TT, sum, t2);
                            dens[first + li] = d;
                            dec = d < 0.0f ? 1 : 2;
                            atomicOr(&s_exact[li >> 6], 1ull << (li & 63));
                        } else {
                            dec = decide_label(TT, s_oct[2][k + 1], sum, t2);
                        }
                        if (dec == 1) atomicOr(&s_lab[li >> 6], 1ull << (li & 63));
                    }
                    evals += (unsigned long long)__popcll(__ballot(active));
                    const uint64_t push = __ballot(active && dec == 0);
                    if (push) {
                        uint32_t base = 0;
                        if (lane == 0) base = atomicAdd(&s_qn[nxt], (uint32_t)__popcll(push));
                        base = __shfl(base, 0);
                        if (active && dec == 0) {
                            const uint32_t q = base + (uint32_t)__popcll(push & lt);
                            s_qi[nxt][q] = (uint16_t)(li | ((k + 1) << kQueueOctaveShift));
                            s_qs[nxt][q] = sum;
                        }
                    }
                }
                __syncthreads();
                if (threadIdx.x == 0) s_qn[cur] = 0;
                __syncthreads();
            }
        }
        // ---- label and exactness words of the tile ----
        if (threadIdx.x < kTile / 64) {
            const uint32_t w0 = first + 64u * threadIdx.x;
            if (w0 < ns) {
                const uint32_t nb = ns - w0;
                const uint64_t vm = nb >= 64 ? ~0ull : ((1ull << nb) - 1ull);
                const uint64_t lb = s_lab[threadIdx.x];
                A.labels[M.lab_off + (w0 >> 6)] = lb;
                A.exact[M.lab_off + (w0 >> 6)] = s_exact[threadIdx.x];
                uint32_t fl = 0;
                if (lb & vm) fl |= 1u;
                if (~lb & vm) fl |= 2u;
                if (fl) atomicOr(&s_flags, fl);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (s_flags & 1u) atomicOr(&A.meta[j].any_air, 1u);
            if (s_flags & 2u) atomicOr(&A.meta[j].any_solid, 1u);
        }
        __syncthreads();
    }
    // the workgroup's count: one LDS add per wave, one plain store per workgroup
    if (adapt && lane == 0 && evals) atomicAdd(&s_evals, evals);
    __syncthreads();
    if (threadIdx.x == 0 && (int)blockIdx.x < A.stat_cap) A.stat_blocks[blockIdx.x] = s_evals;
}

// ------------------------------------------------------------------------
// count: one block per (chunk, x-plane, row segment), one wave per label row ->
// vertices and triangles of each row
// ------------------------------------------------------------------------
// Staged planes: s_lab[0..2] = labels of x-1, x, x+1; s_lab[3] = exactness
// of x.  Counts the row's vertices and triangles (in the quad mode its
// cells' polygons and the quads of its interior sign-changing edges, with
// s_ntri holding polygons per mask), and queues for k_refine the row's
// samples that are endpoints of sign-changing edges but were labelled before
// their last octave (k_sample's adaptive evaluation).
template <int W>
__device__ __forceinline__ void count_row(const Arenas& A, const ChunkMeta& M, const uint64_t* const* s_lab,
                                          const uint8_t* s_ntri, const uint8_t (*s_cls)[BMF_MAX_DIM], int x, int y,
                                          uint32_t* __restrict__ q, uint32_t* s_qn, uint32_t& nv, uint32_t& nt) {
    const int dim = M.dim;
    const int lane = lane_id();
    const bool x1 = x + 1 < dim, y1 = y + 1 < dim;
    nv = 0;
    nt = 0;
    // a row inside one uniform region (the row, and its neighbours at x-1,
    // x+1, y-1, y+1, x+1/y+1, all one constant label) has nothing; decided
    // from the staged rows' classes (0: every label 0, 1: every label 1,
    // 2: mixed) before any row is read
    const uint32_t c00 = s_cls[1][y];
    if (c00 != 2u && (!y1 || s_cls[1][y + 1] == c00) && (!x1 || s_cls[2][y] == c00) &&
        (!(x1 && y1) || s_cls[2][y + 1] == c00) && (x < 1 || s_cls[0][y] == c00) && (y < 1 || s_cls[1][y - 1] == c00))
        return;
    // wave-uniform rows (the same words in every lane); per lane only single bits
    const RowBits<W> r00 = lds_row<W>(s_lab[1], dim, y, true);
    const RowBits<W> r01 = lds_row<W>(s_lab[1], dim, y + 1, y1);
    const RowBits<W> r10 = lds_row<W>(s_lab[2], dim, y, x1);
    const RowBits<W> r11 = lds_row<W>(s_lab[2], dim, y + 1, x1 && y1);
    const RowBits<W> rxm = lds_row<W>(s_lab[0], dim, y, x >= 1);
    const RowBits<W> rym = lds_row<W>(s_lab[1], dim, y - 1, y >= 1);
    const RowBits<W> ex = rxor<W>(r00, r10, x1), ey = rxor<W>(r00, r01, y1), ez = zedges<W>(r00, dim);
    const bool quads = A.mesh_mode == BMF_MESH_QUADS;
    if (quads) {
        RowBits<W> ix, iy, iz;
        interior_edges<W>(r00, r01, r10, dim, x, y, ix, iy, iz);
        nv = 0;
        nt = (total_pop<W>(ix) + total_pop<W>(iy) + total_pop<W>(iz)) * (A.quad_tris ? 2u : 1u);
    } else {
        nv = total_pop<W>(ex) + total_pop<W>(ey) + total_pop<W>(ez);
    }
    uint32_t ncell = 0;  // triangles (polygons in the quad mode) of the row's cells
    if (x < dim - 1 && y < dim - 1 && !quad_uniform<W>(r00, r01, r10, r11, dim)) {
        const RowBits<W> s00 = rshr1<W>(r00), s01 = rshr1<W>(r01), s10 = rshr1<W>(r10), s11 = rshr1<W>(r11);
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (64 * k >= dim - 1) break;
            const int z = 64 * k + lane;
            // corner bit 4 dx + 2 dy + dz
            const uint32_t c =
                z < dim - 1 ? s_ntri[lane_bit(r00.w[k], lane) | (lane_bit(s00.w[k], lane) << 1) |
                                     (lane_bit(r01.w[k], lane) << 2) | (lane_bit(s01.w[k], lane) << 3) |
                                     (lane_bit(r10.w[k], lane) << 4) | (lane_bit(s10.w[k], lane) << 5) |
                                     (lane_bit(r11.w[k], lane) << 6) | (lane_bit(s11.w[k], lane) << 7)]
                            : 0u;
            ncell += (uint32_t)__popcll(__ballot(c & 1u)) + 2u * (uint32_t)__popcll(__ballot(c & 2u)) +
                     4u * (uint32_t)__popcll(__ballot(c & 4u));
        }
    }
    if (quads) nv = ncell;
    else nt = ncell;
    // endpoints of the sign-changing edges incident to the row's samples
    RowBits<W> ep = ror<W>(ror<W>(ex, ey), ror<W>(ez, rshl1<W>(ez)));
    ep = ror<W>(ep, ror<W>(rxor<W>(r00, rxm, x >= 1), rxor<W>(r00, rym, y >= 1)));
    if (total_pop<W>(ep) == 0) return;
    const RowBits<W> done = lds_row<W>(s_lab[3], dim, y, true);
    RowBits<W> need;
#pragma unroll
    for (int k = 0; k < W; k++) need.w[k] = ep.w[k] & ~done.w[k];
    if (total_pop<W>(need) == 0) return;
    // append them to this plane's region of the refine queue (an LDS counter
    // per plane, no global atomics); k_refine evaluates them on dense waves.
    // Bits past the row's last sample are clear, so word k of `need` is the
    // set of lanes that append.
    const uint32_t rowbase = (uint32_t)(y * dim);
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint64_t b = need.w[k];
        if (!b) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(s_qn, (uint32_t)__popcll(b));
        base = __shfl(base, 0);
        if (lane_bit(b, lane)) q[base + mbcnt64(b)] = rowbase + (uint32_t)(64 * k + lane);
    }
}

// ------------------------------------------------------------------------
// refine: full density of the queued samples (endpoints of sign-changing
// edges labelled early by k_sample); one block per (x-plane, row segment) queue.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kSampleThreads) k_refine(Arenas A, int n) {
    const SamplerDev* __restrict__ S = A.sampler;
    unsigned long long evals = 0;  // terrain octaves this workgroup evaluated (uniform)
    for (int cw = 0; cw < 4; cw++) {  // the row-word classes in turn
        const int split = A.ctr->plane_split[cw];
        const int64_t base = A.ctr->class_base[cw];
        const int64_t total = A.ctr->class_planes[cw] * split;
        for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
            const int64_t item = it / split;
            const int seg = (int)(it - item * split);
            const uint32_t cnt = A.plane_refine_n[(base + item) * kMaxPlaneSplit + seg];
            if (cnt == 0) continue;
            const int2 pm = A.plane_map[base + item];
            const ChunkMeta M = read_chunk(A.meta, pm.x);
            const uint32_t dim = (uint32_t)M.dim;
            const uint32_t x = (uint32_t)uniform_int(pm.y);
            const uint32_t y0 = (uint32_t)((int64_t)seg * dim / split);
            const uint32_t* __restrict__ q = A.refine + M.dens_off + (int64_t)x * dim * dim + (int64_t)y0 * dim;
            float* __restrict__ D = A.dens + M.dens_off + (int64_t)x * dim * dim;
            const float px = (float)x * M.delta + M.opos[0];
            if (S->kind == BMF_SAMPLER_TERRAIN) evals += (unsigned long long)cnt * (unsigned long long)S->octaves;
            for (uint32_t e = threadIdx.x; e < cnt; e += blockDim.x) {
                const uint32_t r = q[e];  // y * dim + z within the plane
                const uint32_t y = r / dim, z = r - y * dim;
                const float py = (float)y * M.delta + M.opos[1];
                const float pz = (float)z * M.delta + M.opos[2];
                D[r] = sample_density(*S, px, py, pz);
            }
        }
    }
    if (threadIdx.x == 0 && (int)blockIdx.x < A.stat_cap) A.stat_blocks[A.stat_cap + blockIdx.x] = evals;
}

// Instantiated per row-word class W = ceil(dim/64); blocks of other classes
// skip their items (the host launches only the classes present).
template <int W>
__global__ void __launch_bounds__(kRowThreads) k_count(Arenas A, int n) {
    __shared__ uint8_t s_ntri[256];
    __shared__ uint64_t s_planes[4][plane_words(W)];
    const uint64_t* const s_lab[4] = {s_planes[0], s_planes[1], s_planes[2], s_planes[3]};
    __shared__ uint8_t s_cls[3][BMF_MAX_DIM];  // rows of the staged label planes: 0 / 1 every label 0 / 1, 2 mixed
    __shared__ uint32_t s_qn;
    const bool quads = A.mesh_mode == BMF_MESH_QUADS;
    for (int t = threadIdx.x; t < 256; t += blockDim.x)
        s_ntri[t] = (uint8_t)(quads ? dmc_polygons(c_dmc_table[t]) : (uint32_t)((c_tri_table[t] >> 60) / 3));
    const int wave = uniform_int(threadIdx.x >> 6);
    const int lane = lane_id();
    const int split = A.ctr->plane_split[W - 1];
    const int64_t base = A.ctr->class_base[W - 1];
    const int64_t total = A.ctr->class_planes[W - 1] * split;
    for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int64_t item = it / split;
        const int seg = (int)(it - item * split);
        const int2 pm = A.plane_map[base + item];
        const ChunkMeta M = read_chunk(A.meta, pm.x);
        const int dim = M.dim;
        const int x = uniform_int(pm.y);
        const int64_t qslot = (base + item) * kMaxPlaneSplit + seg;  // this segment's refine-queue count
        const int y0 = (int)((int64_t)seg * dim / split), y1 = (int)((int64_t)(seg + 1) * dim / split);
        uint32_t* rv = A.row_v + M.row_off + (int64_t)x * dim;
        uint32_t* rt = A.row_t + M.row_off + (int64_t)x * dim;
        if (!(M.any_air && M.any_solid)) {  // no surface in this chunk
            for (int y = y0 + threadIdx.x; y < y1; y += blockDim.x) rv[y] = rt[y] = 0u;
            if (threadIdx.x == 0) A.plane_refine_n[qslot] = 0;
            continue;
        }
        __syncthreads();  // previous item is done with the shared arrays
        if (threadIdx.x == 0) s_qn = 0;
        {
            // label planes x-1, x, x+1 and the exact-sample plane x
            uint64_t* const dst[4] = {s_planes[0], s_planes[1], s_planes[2], s_planes[3]};
            const uint64_t* const src[4] = {A.labels + M.lab_off, A.labels + M.lab_off, A.labels + M.lab_off,
                                            A.exact + M.lab_off};
            const int xp[4] = {x - 1, x, x + 1, x};
            stage_plane_group<4>(dst, src, M.lab_words, dim, xp);
        }
        __syncthreads();
        // class of every staged label row, one thread per row
        for (int t = threadIdx.x; t < 3 * dim; t += blockDim.x) {
            const int p = t / dim, y = t - p * dim;
            const uint64_t* P = s_planes[p];
            bool any1 = false, any0 = false;
#pragma unroll
            for (int k = 0; k < W; k++) {
                const int bit = y * dim + 64 * k, w = bit >> 6, sh = bit & 63;
                uint64_t v = (P[w] >> sh) | ((P[w + 1] << 1) << (63 - sh));  // the pad word keeps w + 1 in bounds
                const int nb = dim - 64 * k;
                const uint64_t m = nb >= 64 ? ~0ull : ((1ull << nb) - 1ull);
                v &= m;
                any1 = any1 || v != 0ull;
                any0 = any0 || v != m;
            }
            s_cls[p][y] = (uint8_t)(any1 ? (any0 ? 2 : 1) : 0);
        }
        __syncthreads();
        // this segment's region of the plane's refine queue: its rows' samples
        uint32_t* q = A.refine + M.dens_off + (int64_t)x * dim * dim + (int64_t)y0 * dim;
        for (int y = y0 + wave; y < y1; y += kRowThreads / 64) {
            uint32_t nv = 0, nt = 0;
            count_row<W>(A, M, s_lab, s_ntri, s_cls, x, y, q, &s_qn, nv, nt);
            if (lane == 0) {
                rv[y] = nv;
                rt[y] = nt;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) A.plane_refine_n[qslot] = s_qn;
    }
}

// ------------------------------------------------------------------------
// scan: one block per chunk -> row bases, totals, output reservation
// ------------------------------------------------------------------------
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kScanThreads) k_scan(Arenas A, int n) {
    constexpr int NW = kScanThreads / 64;
    __shared__ uint32_t s_wv[NW], s_wt[NW];  // wave totals
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        ChunkMeta& M = A.meta[j];
        if (M.state != kStateSampled) continue;
        const int dim = M.dim;
        const int64_t R = (int64_t)dim * dim;
        uint32_t* rv = A.row_v + M.row_off;
        uint32_t* rt = A.row_t + M.row_off;
        // exclusive scan of the rows: each wave scans a contiguous run of
        // 64-row tiles (coalesced, wave shuffles, a running carry; four tiles'
        // loads issued together), then one exchange of the wave totals and a
        // second pass that adds each wave's base
        const int64_t ntile = (R + 63) / 64;
        const int64_t tpw = (ntile + NW - 1) / NW;
        const int64_t t0 = (int64_t)wave * tpw, t1 = t0 + tpw < ntile ? t0 + tpw : ntile;
        uint32_t wv = 0, wt = 0;  // running totals of this wave's run
        for (int64_t tg = t0; tg < t1; tg += 4) {
            uint32_t av[4], at[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int64_t r = (tg + k) * 64 + lane;
                const bool in = tg + k < t1 && r < R;
                av[k] = in ? rv[r] : 0u;
                at[k] = in ? rt[r] : 0u;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint32_t sv = av[k], sw = at[k];
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t uv = __shfl_up(sv, off), ut = __shfl_up(sw, off);
                    if (lane >= off) {
                        sv += uv;
                        sw += ut;
                    }
                }
                const int64_t r = (tg + k) * 64 + lane;
                if (tg + k < t1 && r < R) {
                    rv[r] = wv + sv - av[k];
                    rt[r] = wt + sw - at[k];
                }
                wv += (uint32_t)__shfl(sv, 63);
                wt += (uint32_t)__shfl(sw, 63);
            }
        }
        if (lane == 0) {
            s_wv[wave] = wv;
            s_wt[wave] = wt;
        }
        __syncthreads();
        uint32_t bv = 0, bt = 0, cv = 0, ct = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const uint32_t xv = s_wv[w], xt = s_wt[w];
            if (w < wave) {
                bv += xv;
                bt += xt;
            }
            cv += xv;
            ct += xt;
        }
        if (bv | bt)
            for (int64_t t = t0; t < t1; t++) {
                const int64_t r = t * 64 + lane;
                if (r < R) {
                    rv[r] += bv;
                    rt[r] += bt;
                }
            }
        if (tid == 0) {
            const uint32_t V = cv, T = ct;
            M.V = (int32_t)V;
            M.T = (int32_t)T;
            M.contains_mesh = M.any_air && (M.any_solid || (dim % 32 != 0));
            int ok = 1;
            if (V > 0) {
                const int64_t vo = (int64_t)atomicAdd(&A.ctr->v_bump, (unsigned long long)V);
                const uint64_t ni = (uint64_t)M.ipp * T;
                const int64_t io = (int64_t)atomicAdd(&A.ctr->i_bump, (unsigned long long)ni);
                atomicAdd(&A.ctr->need_v, (unsigned long long)V);
                atomicAdd(&A.ctr->need_i, (unsigned long long)ni);
                if (vo + V > A.cap_vertices || io + (int64_t)ni > A.cap_indices) {
                    ok = 0;
                    M.status = BMF_ECAPACITY;
                    A.ctr->overflow = 1;
                }
                M.v_off = vo;
                M.i_off = io;
            }
            (void)ok;
            write_result(A, j, M);
        }
        __syncthreads();
    }
}

// Ordered list of the chunks that run the smoothing (K > 0, a mesh, fitted)
// and the prefix of their vertex tiles, for k_smooth (one block).
__device__ __forceinline__ bool smooths(const ChunkMeta& M, int normals) {
    return (M.iters > 0 || normals) && M.V > 0 && M.T > 0 && M.status == BMF_OK && M.state == kStateSampled;
}

// The list is grouped by XCD: the k-th smoothing chunk goes to XCD k % 8,
// so that all tiles of a chunk run on one XCD and its positions, indices and
// centroids stay in that XCD's L2 (workgroups are dispatched to the 8 XCDs
// round-robin by block index; a different dispatch order only costs
// locality, not correctness).
__global__ void __launch_bounds__(kPlanThreads) k_plan_smooth(Arenas A, int n) {
    __shared__ int64_t s_w[2][kPlanThreads / 64 + 1];
    const int tid = threadIdx.x;
    const int per = (n + kPlanThreads - 1) / kPlanThreads;
    const int j0 = tid * per, j1 = min(n, j0 + per);
    // ---- rank of each smoothing chunk (ordered compaction) ----
    int64_t cnt = 0;
    for (int j = j0; j < j1; j++) cnt += smooths(A.meta[j], A.normals);
    int64_t c1[1] = {cnt}, e1[1], t1[1];
    block_scan<1, kPlanThreads>(c1, e1, t1, s_w);
    const int ns = (int)t1[0];
    int k = (int)e1[0];
    for (int j = j0; j < j1; j++) {
        if (!smooths(A.meta[j], A.normals)) continue;
        // grouped position: groups g = k % 8 in order, rank k / 8 inside
        const int g = k & (kXcds - 1), q = k / kXcds;
        int pos = q;
        for (int h = 0; h < g; h++) pos += (ns - h + kXcds - 1) / kXcds;
        A.smooth_list[pos] = j;
        k++;
    }
    __syncthreads();  // smooth_list written by this block is visible to it
    // ---- tile prefixes in grouped order ----
    const int per2 = (ns + kPlanThreads - 1) / kPlanThreads;
    const int p0 = tid * per2, p1 = min(ns, p0 + per2);
    int64_t tv = 0, tt = 0;
    for (int p = p0; p < p1; p++) {
        const ChunkMeta& M = A.meta[A.smooth_list[p]];
        tv += (M.V + kSmoothThreads - 1) / kSmoothThreads;
        tt += (M.T + kSmoothThreads - 1) / kSmoothThreads;
    }
    int64_t c2[2] = {tv, tt}, e2[2], t2[2];
    block_scan<2, kPlanThreads>(c2, e2, t2, s_w);
    int64_t bv = e2[0], bt = e2[1];
    for (int p = p0; p < p1; p++) {
        const ChunkMeta& M = A.meta[A.smooth_list[p]];
        A.smooth_prefix[p] = bv;
        A.smooth_tprefix[p] = bt;
        bv += (M.V + kSmoothThreads - 1) / kSmoothThreads;
        bt += (M.T + kSmoothThreads - 1) / kSmoothThreads;
    }

    __syncthreads();
    if (tid == kPlanThreads - 1) {
        A.ctr->n_smooth = ns;
        A.smooth_prefix[ns] = t2[0];
        A.smooth_tprefix[ns] = t2[1];
        A.ctr->total_vtiles = t2[0];
        A.ctr->total_ttiles = t2[1];
    }
    __syncthreads();
    // ---- tile range of each XCD group ----
    if (tid <= kXcds) {
        int pos = 0;
        for (int h = 0; h < tid; h++) pos += (ns - h + kXcds - 1) / kXcds;
        A.smooth_xcd_v[tid] = pos < ns ? A.smooth_prefix[pos] : A.smooth_prefix[ns];
        A.smooth_xcd_t[tid] = pos < ns ? A.smooth_tprefix[pos] : A.smooth_tprefix[ns];
    }
}

// Larger batches (more than kPlanOneBlock chunks): the same list and
// prefixes in five passes over many blocks.  The grouped order puts rank k
// (k-th smoothing chunk in chunk order) in group g = k % kXcds at place
// k / kXcds, so the prefix at a position is the totals of the groups before
// its own plus the sum over the earlier ranks of its residue class: block
// scans carry one running sum per residue class.
__device__ __forceinline__ int grouped_pos(int k, int ns) {
    const int g = k & (kXcds - 1);
    int pos = k / kXcds;
    for (int h = 0; h < g; h++) pos += (ns - h + kXcds - 1) / kXcds;
    return pos;
}

// Pass 1: smoothing chunks per block.
__global__ void __launch_bounds__(kPlanThreads) k_smooth_count(Arenas A, int n) {
    __shared__ int64_t s_w[1][kPlanThreads / 64 + 1];
    const int j = blockIdx.x * kPlanThreads + threadIdx.x;
    int64_t f[1] = {j < n && smooths(A.meta[j], A.normals)}, e[1], t[1];
    block_scan<1, kPlanThreads>(f, e, t, s_w);
    if (threadIdx.x == 0) A.plan_part[blockIdx.x] = t[0];
}

// Pass 2 (one block): rank base of each block, and the count.
__global__ void __launch_bounds__(kPlanThreads) k_smooth_ranks(Arenas A, int nb) {
    __shared__ int64_t s_w[1][kPlanThreads / 64 + 1];
    const int b = threadIdx.x;
    int64_t v[1] = {b < nb ? A.plan_part[b] : 0}, e[1], t[1];
    block_scan<1, kPlanThreads>(v, e, t, s_w);
    if (b < nb) A.plan_part[nb + b] = e[0];
    if (b == 0) {
        A.plan_part[2 * nb] = t[0];
        A.ctr->n_smooth = (int)t[0];
    }
}

// Pass 3: each smoothing chunk's rank and place in the list, and the block's
// sums of vertex / triangle tiles per residue class (exclusive sums of the
// chunk kept in plan_tmp).
__global__ void __launch_bounds__(kPlanThreads) k_smooth_place(Arenas A, int n, int nb) {
    __shared__ int64_t s_r[1][kPlanThreads / 64 + 1];
    __shared__ int64_t s_w[2 * kXcds][kPlanThreads / 64 + 1];
    const int j = blockIdx.x * kPlanThreads + threadIdx.x;
    const bool on = j < n && smooths(A.meta[j], A.normals);
    int64_t r1[1] = {on}, re[1], rt[1];
    block_scan<1, kPlanThreads>(r1, re, rt, s_r);
    const int k = (int)(A.plan_part[nb + blockIdx.x] + re[0]);
    const int g = k & (kXcds - 1);
    int64_t f[2 * kXcds], e[2 * kXcds], t[2 * kXcds];
    int64_t vt = 0, tt = 0;
    if (on) {
        const ChunkMeta& M = A.meta[j];
        vt = (M.V + kSmoothThreads - 1) / kSmoothThreads;
        tt = (M.T + kSmoothThreads - 1) / kSmoothThreads;
        A.smooth_list[grouped_pos(k, (int)A.plan_part[2 * nb])] = j;
    }
#pragma unroll
    for (int q = 0; q < kXcds; q++) {
        f[q] = q == g ? vt : 0;
        f[kXcds + q] = q == g ? tt : 0;
    }
    block_scan<2 * kXcds, kPlanThreads>(f, e, t, s_w);
    if (on) {
        A.plan_tmp[2 * (int64_t)j] = pick<2 * kXcds>(e, g);
        A.plan_tmp[2 * (int64_t)j + 1] = pick<2 * kXcds>(e, kXcds + g);
    }
    if (threadIdx.x < 2 * kXcds)
        A.plan_part[2 * nb + 1 + 2 * kXcds * blockIdx.x + threadIdx.x] = pick<2 * kXcds>(t, (int)threadIdx.x);
}

// Pass 4 (one block): per block and class the sums of the blocks before it,
// the group totals, their bases (= each XCD group's first tile) and totals.
__global__ void __launch_bounds__(kPlanThreads) k_smooth_bases(Arenas A, int nb) {
    __shared__ int64_t s_w[2 * kXcds][kPlanThreads / 64 + 1];
    const int b = threadIdx.x;
    int64_t* part = A.plan_part + 2 * nb + 1;
    int64_t* base = part + 2 * kXcds * nb;
    int64_t v[2 * kXcds], e[2 * kXcds], t[2 * kXcds];
#pragma unroll
    for (int q = 0; q < 2 * kXcds; q++) v[q] = b < nb ? part[2 * kXcds * b + q] : 0;
    block_scan<2 * kXcds, kPlanThreads>(v, e, t, s_w);
    if (b < nb)
#pragma unroll
        for (int q = 0; q < 2 * kXcds; q++) base[2 * kXcds * b + q] = e[q];
    if (b == 0) {
        int64_t gv = 0, gt = 0;
        int64_t* gbase = base + 2 * kXcds * nb;
#pragma unroll
        for (int g = 0; g < kXcds; g++) {
            gbase[g] = gv;
            gbase[kXcds + g] = gt;
            A.smooth_xcd_v[g] = gv;
            A.smooth_xcd_t[g] = gt;
            gv += t[g];
            gt += t[kXcds + g];
        }
        A.smooth_xcd_v[kXcds] = gv;
        A.smooth_xcd_t[kXcds] = gt;
        const int ns = (int)A.plan_part[2 * nb];
        A.smooth_prefix[ns] = gv;
        A.smooth_tprefix[ns] = gt;
        A.ctr->total_vtiles = gv;
        A.ctr->total_ttiles = gt;
    }
}

// Pass 5: the tile prefixes at each list position.
__global__ void __launch_bounds__(kPlanThreads) k_smooth_prefix(Arenas A, int n, int nb) {
    __shared__ int64_t s_r[1][kPlanThreads / 64 + 1];
    const int j = blockIdx.x * kPlanThreads + threadIdx.x;
    const bool on = j < n && smooths(A.meta[j], A.normals);
    int64_t r1[1] = {on}, re[1], rt[1];
    block_scan<1, kPlanThreads>(r1, re, rt, s_r);
    if (!on) return;
    const int k = (int)(A.plan_part[nb + blockIdx.x] + re[0]);
    const int ns = (int)A.plan_part[2 * nb];
    const int g = k & (kXcds - 1);
    const int64_t* base = A.plan_part + 2 * nb + 1 + 2 * kXcds * nb;
    const int64_t* gbase = base + 2 * kXcds * nb;
    const int pos = grouped_pos(k, ns);
    A.smooth_prefix[pos] = gbase[g] + base[2 * kXcds * blockIdx.x + g] + A.plan_tmp[2 * (int64_t)j];
    A.smooth_tprefix[pos] = gbase[kXcds + g] + base[2 * kXcds * blockIdx.x + kXcds + g] + A.plan_tmp[2 * (int64_t)j + 1];
}

// ------------------------------------------------------------------------
// emit: one block per (chunk, x-plane, row segment), lanes = z, in slabs of
// at most kEmitRows label rows:
//  pass 1, one wave per row of planes x and x+1 (rows y0 .. y0 + kEmitRows):
//    the first vertex id of every sample (row base + popcount prefix over
//    the row's x, y and z edge masks) with its x- and y-edge bits, into an
//    LDS table; rows of plane x owned by the slab also write their vertices.
//  pass 2, one wave per cell row: the triangles of the row's interior cells
//    (ids by a bit-sliced ballot prefix of the per-cell triangle counts),
//    their corner vertex ids looked up in the table, and for chunks that
//    smooth the per-cell triangle base and mask.
// ------------------------------------------------------------------------
constexpr int kEmitRows = 16;
static_assert((kEmitRows + 1) * 3 * BMF_MAX_DIM < (1 << 14), "a slab's vertex ids fit the 14-bit offsets of the id table");

// Local edge e of a cell (the 12-edge numbering of DMCChunk.cpp:543-565):
// x-edges e = 2 dy + dz, y-edges e = 4 + 2 dx + dz, z-edges e = 8 + 2 dx + dy,
// the edge of axis `ax` starting at corner (dx, dy, dz); 5 bits per edge,
// dx | dy << 1 | dz << 2 | ax << 3.
__host__ __device__ constexpr uint64_t edge_corner_code(int e) {
    return e < 4 ? (uint64_t)(0 | ((e >> 1) << 1) | ((e & 1) << 2) | (0 << 3))
                 : e < 8 ? (uint64_t)(((e - 4) >> 1) | (0 << 1) | ((e & 1) << 2) | (1 << 3))
                         : (uint64_t)(((e - 8) >> 1) | ((e & 1) << 1) | (0 << 2) | (2 << 3));
}
__host__ __device__ constexpr uint64_t edge_corner_table() {
    uint64_t t = 0;
    for (int e = 0; e < 12; e++) t |= edge_corner_code(e) << (5 * e);
    return t;
}
constexpr uint64_t kEdgeCorners = edge_corner_table();

// The same per edge, as emit's pass 2 uses it: the edge's id-table entry
// relative to the cell's own (plane dx, row + dy, column + dz of the
// [2][kEmitRows + 1][64 W] table) in bits 16-31, the entry's x- and y-edge
// bits that precede the edge's vertex among its sample's (bits 14 and 15:
// none for an x-edge, the x-edge for a y-edge, both for a z-edge), the
// plane in bit 0.
template <int W>
__host__ __device__ constexpr uint32_t edge_decode(int e) {
    const uint32_t g = (uint32_t)(kEdgeCorners >> (5 * e)) & 31u;
    const uint32_t dx = g & 1u, dy = (g >> 1) & 1u, dz = (g >> 2) & 1u, ax = g >> 3;
    const uint32_t off = dx * (uint32_t)((kEmitRows + 1) * 64 * W) + dy * (uint32_t)(64 * W) + dz;
    return (off << 16) | (ax >= 1 ? 0x4000u : 0u) | (ax == 2 ? 0x8000u : 0u) | dx;
}

// Pass 1 for label row (x + p, y): vertex ids into s_id, and for p == 0 with
// `own` the vertices themselves (DMCChunk::calculate_isovertex, DMCChunk.cpp:
// 657-674; vertices of a sample in x, y, z edge order).  The row masks are
// wave-uniform (scalar registers); per lane only the bit and the count of
// set bits below it are taken.
template <int W>
__device__ __forceinline__ void emit_ids_row(const Arenas& A, const ChunkMeta& M,
                                             const uint64_t (*s_lab)[plane_words(W)], uint32_t base,
                                             uint16_t* s_idrow, uint32_t slab_base, int x, int p, int y, bool own) {
    const int dim = M.dim;
    const int lane = lane_id();
    const int xp = x + p;
    const bool vx = xp + 1 < dim, vy = y + 1 < dim;
    const RowBits<W> R0 = uniform_row<W>(lds_row<W>(s_lab[p], dim, y, true));
    const RowBits<W> EX = rxor<W>(R0, uniform_row<W>(lds_row<W>(s_lab[p + 1], dim, y, vx)), vx);
    const RowBits<W> EY = rxor<W>(R0, uniform_row<W>(lds_row<W>(s_lab[p], dim, y + 1, vy)), vy);
    const RowBits<W> EZ = zedges<W>(R0, dim);
    const int64_t vo = M.v_off;
    const uint32_t Vc = (uint32_t)M.V;
    float* __restrict__ P = A.pos + vo * 3;
    const bool cells = M.iters > 0 || A.normals;  // the chunk runs the smoothing kernels
    // with K > 0 the last primal step writes the output position of every
    // vertex it moves (every vertex has a ring; boundary vertices move only
    // with boundary processing), so only the others are written here
    const bool moves = M.iters > 0 && A.smoothing_follows;
    const bool moves_boundary = moves && M.bp;
    float4* __restrict__ P4 = cells ? A.pos4 + vo : nullptr;  // smoothing state
    float4* __restrict__ N4 = cells && A.normals ? A.nrm4 + vo : nullptr;
    uint8_t* __restrict__ B = A.boundary + vo;
    uint32_t* __restrict__ E = A.vedge + vo;
    const float* __restrict__ D = A.dens + M.dens_off;
    uint32_t kb = base;  // id of the first vertex of word k's samples
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint64_t ex = EX.w[k], ey = EY.w[k], ez = EZ.w[k];
        const int z = 64 * k + lane;
        if (z < dim) {
            const uint32_t fx = lane_bit_u(ex), fy = lane_bit_u(ey), fz = lane_bit_u(ez);
            uint32_t id = kb + mbcnt64(ex) + mbcnt64(ey) + mbcnt64(ez);
            s_idrow[z] = (uint16_t)((id - slab_base) | (fx << 14) | (fy << 15));
            if (own && (fx | fy | fz)) {
                // the sample and the far ends of its edges load together (an
                // edge that is not set reads the sample again)
                const int64_t b = ((int64_t)x * dim + y) * dim + z;
                const float s0 = D[b];
                const float e1[3] = {D[fx ? b + (int64_t)dim * dim : b], D[fy ? b + dim : b], D[fz ? b + 1 : b]};
                const bool fits = id + fx + fy + fz <= Vc;
                if (!fits) A.ctr->error = 2;  // counts and ids disagree: report, never write past the chunk
#pragma unroll
                for (int ax = 0; ax < 3; ax++) {
                    const uint32_t f = ax == 0 ? fx : (ax == 1 ? fy : fz);
                    if (!f || !fits) continue;
                    const int xe = x + (ax == 0), ye = y + (ax == 1), ze = z + (ax == 2);
                    const float s1 = e1[ax];
                    const float mu = (0.0f - s0) / (s1 - s0);
                    const float dx = (float)xe - (float)x, dy = (float)ye - (float)y, dz = (float)ze - (float)z;
                    const float qx = dx * mu + (float)x, qy = dy * mu + (float)y, qz = dz * mu + (float)z;
                    const bool bnd = x == 0 || y == 0 || z == 0 || x == dim - 1 || y == dim - 1 || z == dim - 1 ||
                                     xe == dim - 1 || ye == dim - 1 || ze == dim - 1;
                    if (!(bnd ? moves_boundary : moves)) {
                        P[3 * (int64_t)id + 0] = qx;
                        P[3 * (int64_t)id + 1] = qy;
                        P[3 * (int64_t)id + 2] = qz;
                    }
                    if (P4) P4[id] = make_float4(qx, qy, qz, 0.0f);
                    if (N4) N4[id] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    B[id] = (uint8_t)bnd;
                    E[id] = (uint32_t)x | ((uint32_t)y << 8) | ((uint32_t)z << 16) | ((uint32_t)ax << 24);
                    id++;
                }
            }
        }
        kb += (uint32_t)(__popcll(ex) + __popcll(ey) + __popcll(ez));
    }
}

// Pass 2 for cell row (x, y); s_id rows start at label row ys.
template <int W>
__device__ __forceinline__ void emit_tris_row(const Arenas& A, const ChunkMeta& M,
                                              const uint64_t (*s_lab)[plane_words(W)], const uint8_t* s_ntri,
                                              const uint64_t* s_tri, const uint32_t* s_edge, uint32_t tcarry,
                                              const uint16_t (*s_id)[kEmitRows + 1][64 * W], uint32_t slab0,
                                              uint32_t slab1, int x, int y, int ys) {
    const int dim = M.dim;
    const int lane = lane_id();
    // corner rows of the cell row, and the same shifted by one sample (z + 1)
    const RowBits<W> R00 = uniform_row<W>(lds_row<W>(s_lab[0], dim, y, true));
    const RowBits<W> R01 = uniform_row<W>(lds_row<W>(s_lab[0], dim, y + 1, true));
    const RowBits<W> R10 = uniform_row<W>(lds_row<W>(s_lab[1], dim, y, true));
    const RowBits<W> R11 = uniform_row<W>(lds_row<W>(s_lab[1], dim, y + 1, true));
    const RowBits<W> S00 = rshr1<W>(R00), S01 = rshr1<W>(R01), S10 = rshr1<W>(R10), S11 = rshr1<W>(R11);
    const int64_t io = M.i_off;
    const uint32_t Vc = (uint32_t)M.V, Tc = (uint32_t)M.T;
    const bool cells = M.iters > 0 || A.normals;
#pragma unroll
    for (int k = 0; k < W; k++) {
        if (64 * k >= dim - 1) break;
        const int z = 64 * k + lane;
        // corner bit 4 dx + 2 dy + dz (DMCChunk.cpp:200-438)
        const uint32_t mask =
            z < dim - 1 ? lane_bit_u(R00.w[k]) | (lane_bit_u(S00.w[k]) << 1) | (lane_bit_u(R01.w[k]) << 2) |
                              (lane_bit_u(S01.w[k]) << 3) | (lane_bit_u(R10.w[k]) << 4) |
                              (lane_bit_u(S10.w[k]) << 5) | (lane_bit_u(R11.w[k]) << 6) |
                              (lane_bit_u(S11.w[k]) << 7)
                        : 0u;
        const uint32_t nt = s_ntri[mask];
        const uint64_t b0 = __ballot(nt & 1u), b1 = __ballot(nt & 2u), b2 = __ballot(nt & 4u);
        if (nt) {
            const uint32_t tb = tcarry + mbcnt64(b0) + 2u * mbcnt64(b1) + 4u * mbcnt64(b2);
            const uint64_t row = s_tri[mask];
            if (tb + nt > Tc) {
                A.ctr->error = 3;
            } else {
                // per triangle: the three corner ids read together, written
                // together; an id past the chunk's vertices (counts and ids
                // disagree) is reported, its slot is inside the chunk's range
                uint32_t* __restrict__ I = A.idx + io + 3ll * tb;
                bool bad = false;
                // the cell's own entry of the id table; a corner's entry is
                // its edge's offset from it (edge_decode)
                const uint16_t* __restrict__ cid = &s_id[0][0][0] + (y - ys) * (64 * W) + z;
                const uint32_t sdiff = slab1 - slab0;
                // at most 5 triangles per cell: unrolled, so each slot's table
                // bits are a constant field of the row
#pragma unroll
                for (int q = 0; q < 5; q++) {
                    if (q >= (int)nt) break;
                    uint32_t v[3];
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const int sl = 3 * q + c;
                        const uint32_t d = s_edge[(uint32_t)(row >> (4 * sl)) & 15u];
                        const uint32_t e = cid[d >> 16];
                        v[c] = slab0 + (d & 1u) * sdiff + (e & 0x3FFFu) + (uint32_t)__popc(e & d & 0xC000u);
                        bad = bad || v[c] >= Vc;
                    }
                    I[3 * q] = v[0];
                    I[3 * q + 1] = v[1];
                    I[3 * q + 2] = v[2];
                }
                if (bad) A.ctr->error = 4;
                if (cells) {
                    const int64_t ci = M.dens_off + ((int64_t)x * dim + y) * dim + z;
                    A.cell_tb[ci] = tb;
                    A.cell_mask[ci] = (uint8_t)mask;
                }
            }
        }
        tcarry += (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
    }
}

template <int W>
__global__ void __launch_bounds__(kRowThreads) k_emit(Arenas A, int n) {
    constexpr int NW = kRowThreads / 64;
    __shared__ uint8_t s_ntri[256];
    __shared__ uint64_t s_planes[3][plane_words(W)];  // labels of planes x, x+1, x+2
    __shared__ uint32_t s_base[3][BMF_MAX_DIM + 1];  // row vertex bases of planes x, x+1; triangle bases of plane x
    // first vertex id of each sample of the slab's rows, less the slab's first
    // row base of its plane (a slab's rows hold at most (kEmitRows + 1) *
    // 3 * 256 < 2^14 vertices), | x-edge bit 14 | y-edge bit 15
    __shared__ uint16_t s_id[2][kEmitRows + 1][64 * W];
    __shared__ uint64_t s_tri[256];                       // the triangle table
    __shared__ uint32_t s_edge[16];                       // edge_decode of edges 0..11
    if (threadIdx.x < 16) s_edge[threadIdx.x] = threadIdx.x < 12 ? edge_decode<W>(threadIdx.x) : 0u;
    for (int t = threadIdx.x; t < 256; t += blockDim.x) {
        s_tri[t] = c_tri_table[t];
        s_ntri[t] = (uint8_t)((c_tri_table[t] >> 60) / 3);
    }
    const int wave = uniform_int(threadIdx.x >> 6);
    const int split = A.ctr->plane_split[W - 1];
    const int64_t base = A.ctr->class_base[W - 1];
    const int64_t total = A.ctr->class_planes[W - 1] * split;
    for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int64_t item = it / split;
        const int seg = (int)(it - item * split);
        const int2 pm = A.plane_map[base + item];
        const ChunkMeta M = read_chunk(A.meta, pm.x);
        const int dim = M.dim;
        if (M.status != BMF_OK || M.V == 0) continue;  // uniform across the block
        const int x = uniform_int(pm.y);
        const int y0 = (int)((int64_t)seg * dim / split), y1 = (int)((int64_t)(seg + 1) * dim / split);
        __syncthreads();  // previous item is done with the shared arrays
        stage_planes(s_planes[0], s_planes[1], s_planes[2], A.labels + M.lab_off, M.lab_words, dim, x);
        const uint32_t* rv = A.row_v + M.row_off + (int64_t)x * dim;
        const uint32_t* rt = A.row_t + M.row_off + (int64_t)x * dim;
        // row bases of planes x and x+1 and the triangle bases of plane x; the
        // entry after a plane's last row is the next row's base (or the chunk
        // total), so base[y+1] - base[y] is the row's count
        const int64_t R = (int64_t)dim * dim, r0 = (int64_t)x * dim;
        for (int y = threadIdx.x; y <= dim; y += blockDim.x) {
            const int64_t ra = r0 + y, rb = r0 + dim + y;
            const uint32_t va = rv[ra < R ? y : 0], vb = rv[rb < R ? dim + y : 0], ta = rt[ra < R ? y : 0];
            s_base[0][y] = ra < R ? va : (uint32_t)M.V;
            s_base[1][y] = rb < R ? vb : (uint32_t)M.V;
            s_base[2][y] = ra < R ? ta : (uint32_t)M.T;
        }
        const bool cellsx = x < dim - 1;  // plane x has cells (and plane x+1 exists)
        for (int ys = y0; ys < y1; ys += kEmitRows) {
            const int ye = min(y1, ys + kEmitRows);         // cell / owned rows [ys, ye)
            const int yl = min(dim - 1, ye);                // id rows [ys, yl]
            __syncthreads();  // staging done; the previous slab is done with s_id
            const uint32_t slab0 = s_base[0][ys], slab1 = s_base[1][ys];  // the id table's offsets
            // pass 1: rows of plane x (and x+1 when it has cells); rows that own no
            // vertex are never referenced by a triangle and skip the label work
            const int nrow = yl - ys + 1;
            for (int t = wave; t < (cellsx ? 2 : 1) * nrow; t += NW) {
                const int p = t >= nrow ? 1 : 0, y = ys + (t - p * nrow);
                if (s_base[p][y + 1] == s_base[p][y]) continue;
                emit_ids_row<W>(A, M, s_planes, s_base[p][y], s_id[p][y - ys], p ? slab1 : slab0, x, p, y, p == 0 && y < ye);
            }
            if (!cellsx) continue;
            __syncthreads();
            // pass 2: cell rows with triangles
            for (int y = ys + wave; y < ye && y < dim - 1; y += NW) {
                if (s_base[2][y + 1] == s_base[2][y]) continue;
                emit_tris_row<W>(A, M, s_planes, s_ntri, s_tri, s_edge, s_base[2][y], s_id, slab0, slab1, x, y, ys);
            }
        }
    }
}

// Tile -> (chunk, tile within the chunk) for the smoothing kernels, one
// block per listed chunk, written once per frame before the first step.
__global__ void __launch_bounds__(kSmoothThreads) k_smooth_map(Arenas A) {
    const int ns = A.ctr->n_smooth;
    for (int p = blockIdx.x; p < ns; p += gridDim.x) {
        const int j = A.smooth_list[p];
        const ChunkMeta& M = A.meta[j];
        const int vt = (M.V + kSmoothThreads - 1) / kSmoothThreads, tt = (M.T + kSmoothThreads - 1) / kSmoothThreads;
        const int64_t bv = A.smooth_prefix[p], bt = A.smooth_tprefix[p];
        for (int r = threadIdx.x; r < vt; r += blockDim.x) A.smooth_vmap[bv + r] = make_int2(j, r);
        for (int r = threadIdx.x; r < tt; r += blockDim.x) A.smooth_tmap[bt + r] = make_int2(j, r);
    }
}

// ------------------------------------------------------------------------
// smoothing, iteration m of MeshProcessor<3> (MeshProcessor.cpp:130-306 as
// driven by ChunkGenerator.cpp:112-124), over the chunks with K > m:
//   k_dual   per triangle: centroid ((0 + p0) + p1) + p2, then / 3
//   k_primal per vertex: sum of the adjacent centroids from +0 in ascending
//            triangle id, / count; boundary vertices stay unless
//            boundary_processing.  The adjacent triangles come from the ring
//            of cells around the vertex's lattice edge (c_ring) and each
//            cell's table slots, which is ascending id order.
// Positions live in a float4 buffer during the iterations (updated in place
// by k_primal: k_dual has consumed them); the last iteration also writes the
// float3 output.
// ------------------------------------------------------------------------
// The chunk fields a smoothing tile reads, loaded together and kept in
// scalar registers: a tile, so its chunk, is the whole block's.
struct SmoothChunk {
    int iters, dim, bp, V, T;
    int64_t v_off, i_off, dens_off;
};
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
    const uint64_t u = (uint64_t)v;
    return (int64_t)(((uint64_t)(uint32_t)uniform_int((int)(u >> 32)) << 32) | (uint32_t)uniform_int((int)(uint32_t)u));
}
__device__ __forceinline__ SmoothChunk smooth_chunk(const ChunkMeta& M) {
    const int iters = M.iters, dim = M.dim, bp = M.bp, V = M.V, T = M.T;
    const int64_t v_off = M.v_off, i_off = M.i_off, dens_off = M.dens_off;
    return SmoothChunk{uniform_int(iters), uniform_int(dim), uniform_int(bp), uniform_int(V), uniform_int(T),
                       uniform_i64(v_off), uniform_i64(i_off), uniform_i64(dens_off)};
}

__global__ void __launch_bounds__(kSmoothThreads) k_dual(Arenas A, int m) {
    const int xcd = blockIdx.x % kXcds;
    const int64_t end = A.smooth_xcd_t[xcd + 1];
    for (int64_t tile = A.smooth_xcd_t[xcd] + blockIdx.x / kXcds; tile < end; tile += gridDim.x / kXcds) {
        const int2 tm = A.smooth_tmap[tile];  // (chunk, tile within the chunk)
        const int chunk = uniform_int(tm.x), ty = uniform_int(tm.y);
        const SmoothChunk M = smooth_chunk(A.meta[chunk]);
        const int64_t T = M.T, i_off = M.i_off, v_off = M.v_off;
        const uint32_t V = (uint32_t)M.V;
        if (M.iters <= m) continue;
        const int64_t t = (int64_t)ty * kSmoothThreads + threadIdx.x;
        if (t >= T) continue;
        const uint32_t* __restrict__ I = A.idx + i_off + 3 * t;
        const float4* __restrict__ P = A.pos4 + v_off;
        const uint32_t i0 = I[0], i1 = I[1], i2 = I[2];
        if (i0 >= V || i1 >= V || i2 >= V) {
            A.ctr->error = 5;
            continue;
        }
        const float4 a = P[i0], b = P[i1], c = P[i2];
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        cx = cx + a.x;
        cy = cy + a.y;
        cz = cz + a.z;
        cx = cx + b.x;
        cy = cy + b.y;
        cz = cz + b.z;
        cx = cx + c.x;
        cy = cy + c.y;
        cz = cz + c.z;
        A.cen4[i_off / 3 + t] = make_float4(cx / 3.0f, cy / 3.0f, cz / 3.0f, 0.0f);
    }
}

// A vertex's ring, packed: word c for ring cell c (the cells around the
// vertex's lattice edge in c_ring order) = the cell's first triangle | the
// bits of its triangles that use the edge << 27.  A chunk holds at most
// 5 * 255^3 < 2^27 triangles.  The first smoothing step packs it (four
// independent cell reads), later steps load it as one 16-byte word.
constexpr uint32_t kRingFirst = (1u << 27) - 1u;
constexpr int kRingRound = 8;  // centroids loaded per round (a ring has about six triangles)

// Ring cell c around an edge along `ax` is the edge origin offset by
// (c >> 1) - 1 and (c & 1) - 1 along the other two axes in x, y, z order,
// with local edge 4 ax + 3 - c: the rows of c_ring.
__device__ __forceinline__ uint4 pack_ring(const uint32_t* __restrict__ CT, const uint8_t* __restrict__ CM,
                                           uint32_t code, int dim, uint32_t T, bool& bad) {
    const int ex = (int)(code & 255u), ey = (int)((code >> 8) & 255u), ez = (int)((code >> 16) & 255u);
    const int ax = (int)(code >> 24);
    int64_t ci[4];
    bool in[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int u = (c >> 1) - 1, w = (c & 1) - 1;
        const int cx = ex + (ax == 0 ? 0 : u), cy = ey + (ax == 0 ? u : (ax == 1 ? 0 : w)), cz = ez + (ax == 2 ? 0 : w);
        in[c] = cx >= 0 && cy >= 0 && cz >= 0 && cx <= dim - 2 && cy <= dim - 2 && cz <= dim - 2;
        ci[c] = in[c] ? ((int64_t)cx * dim + cy) * dim + cz : 0;
    }
    uint32_t tb[4], mk[4], r[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        tb[c] = CT[ci[c]];
        mk[c] = CM[ci[c]];
    }
#pragma unroll
    for (int c = 0; c < 4; c++) mk[c] = c_edge_tris.m[mk[c]][4 * ax + 3 - c];  // any mask byte is a valid row
#pragma unroll
    for (int c = 0; c < 4; c++) {
        uint32_t tm = in[c] ? mk[c] : 0u;
        if (tm && tb[c] + (uint32_t)(31 - __clz(tm)) >= T) {  // past the chunk's triangles
            bad = true;
            tm = 0u;
        }
        r[c] = (tm ? tb[c] & kRingFirst : 0u) | (tm << 27);
    }
    return make_uint4(r[0], r[1], r[2], r[3]);
}

// Sums of the ring's centroids (and face normals) from +0 in ascending
// triangle id, the order of MeshProcessor<3>'s adjacency lists; returns the
// ring size.  kRingRound centroids are loaded per round, a lane's unused slots
// reading its round's first triangle and adding nothing.
template <bool NRM>
__device__ __forceinline__ uint32_t ring_sums(uint4 r, const float4* __restrict__ C, const float4* __restrict__ F,
                                              float (&s)[6]) {
    uint32_t bits = (r.x >> 27) | ((r.y >> 27) << 5) | ((r.z >> 27) << 10) | ((r.w >> 27) << 15);
    const uint32_t cnt = (uint32_t)__popc(bits);
    // triangle of bit b = 5 c + slot: first(c) + slot = (first(c) - 5 c) + b, modulo 2^32
    const uint32_t f0 = r.x & kRingFirst, f1 = (r.y & kRingFirst) - 5u, f2 = (r.z & kRingFirst) - 10u,
                   f3 = (r.w & kRingFirst) - 15u;
    while (bits) {
        uint32_t t[kRingRound];
        bool ok[kRingRound];
#pragma unroll
        for (int k = 0; k < kRingRound; k++) {
            ok[k] = bits != 0u;
            const uint32_t b = ok[k] ? (uint32_t)(__ffs(bits) - 1) : 0u;
            const uint32_t f = b < 5u ? f0 : (b < 10u ? f1 : (b < 15u ? f2 : f3));
            t[k] = ok[k] ? f + b : (k == 0 ? 0u : t[0]);  // slot 0 is always used
            bits &= bits - 1u;
        }
        float4 q[kRingRound], n[kRingRound];
#pragma unroll
        for (int k = 0; k < kRingRound; k++) {
            q[k] = C[t[k]];
            if (NRM) n[k] = F[t[k]];
        }
#pragma unroll
        for (int k = 0; k < kRingRound; k++) {
            s[0] = ok[k] ? s[0] + q[k].x : s[0];
            s[1] = ok[k] ? s[1] + q[k].y : s[1];
            s[2] = ok[k] ? s[2] + q[k].z : s[2];
            if (NRM) {
                s[3] = ok[k] ? s[3] + n[k].x : s[3];
                s[4] = ok[k] ? s[4] + n[k].y : s[4];
                s[5] = ok[k] ? s[5] + n[k].z : s[5];
            }
        }
    }
    return cnt;
}

__global__ void __launch_bounds__(kSmoothThreads) k_primal(Arenas A, int m) {
    const int xcd = blockIdx.x % kXcds;
    const int64_t end = A.smooth_xcd_v[xcd + 1];
    for (int64_t tile = A.smooth_xcd_v[xcd] + blockIdx.x / kXcds; tile < end; tile += gridDim.x / kXcds) {
        const int2 tm = A.smooth_vmap[tile];  // (chunk, tile within the chunk)
        const int chunk = uniform_int(tm.x), ty = uniform_int(tm.y);
        const SmoothChunk M = smooth_chunk(A.meta[chunk]);
        const int K = M.iters, dim = M.dim, bp = M.bp;
        const int64_t V = M.V, v_off = M.v_off, dens_off = M.dens_off, t_off = M.i_off / 3;
        const uint32_t T = (uint32_t)M.T;
        if (K <= m) continue;
        const int64_t v = (int64_t)ty * kSmoothThreads + threadIdx.x;
        if (v >= V) continue;
        const int64_t g = v_off + v;
        // the boundary flag, the lattice edge and the packed ring load together
        const uint8_t bnd = A.boundary[g];
        const uint32_t code = m == 0 ? A.vedge[g] : 0u;
        uint4 r = m == 0 ? make_uint4(0u, 0u, 0u, 0u) : A.vring[g];
        if (!bp && bnd) continue;  // stays at its position in both buffers
        if (m == 0) {
            bool bad = false;
            r = pack_ring(A.cell_tb + dens_off, A.cell_mask + dens_off, code, dim, T, bad);
            if (bad) A.ctr->error = 7;
        }
        float s[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const uint32_t cnt = ring_sums<false>(r, A.cen4 + t_off, nullptr, s);
        if (m == 0 && K > 1) A.vring[g] = r;  // stored after the loads: a store here is waited for with them
        if (cnt == 0) {  // MeshProcessor.cpp:247: no adjacency, unchanged (the emit kernel left its output to here)
            if (m == K - 1) {
                const float4 q = A.pos4[g];
                float* __restrict__ P = A.pos + g * 3;
                P[0] = q.x;
                P[1] = q.y;
                P[2] = q.z;
            }
            continue;
        }
        const float fc = (float)cnt;
        const float px = s[0] / fc, py = s[1] / fc, pz = s[2] / fc;
        A.pos4[g] = make_float4(px, py, pz, 0.0f);
        if (m == K - 1) {
            float* __restrict__ P = A.pos + g * 3;
            P[0] = px;
            P[1] = py;
            P[2] = pz;
        }
    }
}


// With vertex normals (the smooth_normals branch of MeshProcessor<3>,
// MeshProcessor.cpp:159-216, 271-287, for N = 3): every chunk of the list
// runs max(K, 1) steps; step m's face normal is
// -cross(normalize(p0 - p1), normalize(p0 - p2)) while m == 0, m < max_norms
// (min(K/2 - 3, 10)) or m < 3, else the sum of its corners' normals; the
// primal step averages them over the ring for every vertex with a ring,
// normalises when m == 3 or (m == 0 and K <= 3) and keeps the result when
// its y is not 0; positions move as in k_dual/k_primal (K > 0 only); the
// last step writes the unit normal.
__global__ void __launch_bounds__(kSmoothThreads) k_dual_nrm(Arenas A, int m) {
    const int xcd = blockIdx.x % kXcds;
    const int64_t end = A.smooth_xcd_t[xcd + 1];
    for (int64_t tile = A.smooth_xcd_t[xcd] + blockIdx.x / kXcds; tile < end; tile += gridDim.x / kXcds) {
        const int2 tm = A.smooth_tmap[tile];  // (chunk, tile within the chunk)
        const int chunk = uniform_int(tm.x), ty = uniform_int(tm.y);
        const SmoothChunk M = smooth_chunk(A.meta[chunk]);
        const int K = M.iters;
        if (max(K, 1) <= m) continue;
        const int64_t t = (int64_t)ty * kSmoothThreads + threadIdx.x;
        if (t >= M.T) continue;
        const uint32_t* __restrict__ I = A.idx + M.i_off + 3 * t;
        const float4* __restrict__ P = A.pos4 + M.v_off;
        const uint32_t i0 = I[0], i1 = I[1], i2 = I[2];
        if (i0 >= (uint32_t)M.V || i1 >= (uint32_t)M.V || i2 >= (uint32_t)M.V) {
            A.ctr->error = 5;
            continue;
        }
        const float4 a = P[i0], b = P[i1], c = P[i2];
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        cx = cx + a.x;
        cy = cy + a.y;
        cz = cz + a.z;
        cx = cx + b.x;
        cy = cy + b.y;
        cz = cz + b.z;
        cx = cx + c.x;
        cy = cy + c.y;
        cz = cz + c.z;
        const int64_t tt = M.i_off / 3 + t;
        A.cen4[tt] = make_float4(cx / 3.0f, cy / 3.0f, cz / 3.0f, 0.0f);
        const int max_norms = min(K / 2 - 3, 10);
        float4 fn;
        if (m == 0 || m < max_norms || m < 3) {
            float ux = a.x - b.x, uy = a.y - b.y, uz = a.z - b.z;
            float wx = a.x - c.x, wy = a.y - c.y, wz = a.z - c.z;
            const float lu = sqrtf((ux * ux + uy * uy) + uz * uz), lw = sqrtf((wx * wx + wy * wy) + wz * wz);
            ux = ux / lu;
            uy = uy / lu;
            uz = uz / lu;
            wx = wx / lw;
            wy = wy / lw;
            wz = wz / lw;
            fn = make_float4(-(uy * wz - wy * uz), -(uz * wx - wz * ux), -(ux * wy - wx * uy), 0.0f);
        } else {
            const float4* __restrict__ N = A.nrm4 + M.v_off;
            const float4 na = N[i0], nb = N[i1], nc = N[i2];
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            sx = sx + na.x;
            sy = sy + na.y;
            sz = sz + na.z;
            sx = sx + nb.x;
            sy = sy + nb.y;
            sz = sz + nb.z;
            sx = sx + nc.x;
            sy = sy + nc.y;
            sz = sz + nc.z;
            fn = make_float4(sx, sy, sz, 0.0f);
        }
        A.fn4[tt] = fn;
    }
}

__global__ void __launch_bounds__(kSmoothThreads) k_primal_nrm(Arenas A, int m) {
    const int xcd = blockIdx.x % kXcds;
    const int64_t end = A.smooth_xcd_v[xcd + 1];
    for (int64_t tile = A.smooth_xcd_v[xcd] + blockIdx.x / kXcds; tile < end; tile += gridDim.x / kXcds) {
        const int2 tm = A.smooth_vmap[tile];  // (chunk, tile within the chunk)
        const int chunk = uniform_int(tm.x), ty = uniform_int(tm.y);
        const SmoothChunk M = smooth_chunk(A.meta[chunk]);
        const int K = M.iters;
        const int steps = max(K, 1);
        if (steps <= m) continue;
        const int64_t v = (int64_t)ty * kSmoothThreads + threadIdx.x;
        if (v >= M.V) continue;
        const int64_t g = M.v_off + v, dens_off = M.dens_off, t_off = M.i_off / 3;
        const int dim = M.dim, bp = M.bp;
        const uint32_t T = (uint32_t)M.T;
        uint4 r;
        if (m == 0) {
            bool bad = false;
            r = pack_ring(A.cell_tb + dens_off, A.cell_mask + dens_off, A.vedge[g], dim, T, bad);
            if (bad) A.ctr->error = 7;
        } else {
            r = A.vring[g];
        }
        float s[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const uint32_t cnt = ring_sums<true>(r, A.cen4 + t_off, A.fn4 + t_off, s);
        if (m == 0 && steps > 1) A.vring[g] = r;
        const float sx = s[0], sy = s[1], sz = s[2], nx = s[3], ny = s[4], nz = s[5];
        float4 n = A.nrm4[g];
        if (cnt > 0) {
            const float fc = (float)cnt;
            float ax2 = nx / fc, ay2 = ny / fc, az2 = nz / fc;
            const bool unit = m != steps - 1 && (m == 3 || (m == 0 && K <= 3));
            if (unit) {
                const float l = sqrtf((ax2 * ax2 + ay2 * ay2) + az2 * az2);
                ax2 = ax2 / l;
                ay2 = ay2 / l;
                az2 = az2 / l;
            }
            if (ay2 != 0.0f) {
                n = make_float4(ax2, ay2, az2, 0.0f);
                A.nrm4[g] = n;
            }
            if (K > 0 && (bp || !A.boundary[g])) {
                const float px = sx / fc, py = sy / fc, pz = sz / fc;
                A.pos4[g] = make_float4(px, py, pz, 0.0f);
                if (m == K - 1) {
                    float* __restrict__ P = A.pos + g * 3;
                    P[0] = px;
                    P[1] = py;
                    P[2] = pz;
                }
            }
        } else if (K > 0 && m == K - 1 && (bp || !A.boundary[g])) {
            // no ring: unchanged, and the emit kernel left its output to here
            const float4 q = A.pos4[g];
            float* __restrict__ P = A.pos + g * 3;
            P[0] = q.x;
            P[1] = q.y;
            P[2] = q.z;
        }
        if (m == steps - 1) {
            const float l = sqrtf((n.x * n.x + n.y * n.y) + n.z * n.z);
            float* __restrict__ O = A.nrm + g * 3;
            O[0] = l > 0.0f ? n.x / l : 0.0f;
            O[1] = l > 0.0f ? n.y / l : 0.0f;
            O[2] = l > 0.0f ? n.z / l : 0.0f;
        }
    }
}

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
    if (n <= kPlanOneBlock) {
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, s, A, n, allow_cull, reset_bumps);
    } else {
        const int nb = (n + kPlanThreads - 1) / kPlanThreads;
        if (nb > kPlanThreads) return hipErrorInvalidValue;  // beyond BMF's 2^20 chunks
        hipLaunchKernelGGL(k_plan_part, dim3(nb), dim3(kPlanThreads), 0, s, A, n, allow_cull);
        hipLaunchKernelGGL(k_plan_total, dim3(1), dim3(kPlanThreads), 0, s, A, n, nb, reset_bumps);
        hipLaunchKernelGGL(k_plan_fix, dim3(nb), dim3(kPlanThreads), 0, s, A, n, nb);
    }
    if (A.csg_list) hipLaunchKernelGGL(k_csg_cull, dim3((unsigned)(n < 2048 ? n : 2048)), dim3(256), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_sample(const Arenas& A, int n, int blocks, int adaptive, int kind, int slab, hipStream_t s) {
    const int mode = adaptive && kind == BMF_SAMPLER_TERRAIN ? kSampleAdaptive
                     : kind == BMF_SAMPLER_CSG && A.csg_list  ? kSampleCsg
                     : kind == BMF_SAMPLER_NOISE && slab      ? kSampleSlab
                                                              : kSampleGeneric;
    const dim3 g(blocks), b(kSampleThreads);
#define BMF_SAMPLE(R)                                                                             \
    switch (mode) {                                                                               \
        case kSampleAdaptive: hipLaunchKernelGGL((k_sample<R, kSampleAdaptive>), g, b, 0, s, A, n); break; \
        case kSampleCsg: hipLaunchKernelGGL((k_sample<R, kSampleCsg>), g, b, 0, s, A, n); break;           \
        case kSampleSlab: hipLaunchKernelGGL((k_sample<R, kSampleSlab>), g, b, 0, s, A, n); break;         \
        default: hipLaunchKernelGGL((k_sample<R, kSampleGeneric>), g, b, 0, s, A, n); break;               \
    }
    switch (A.sample_tile) {
        case 1024: BMF_SAMPLE(4) break;
        case 2048: BMF_SAMPLE(8) break;
        case 4096: BMF_SAMPLE(16) break;
        default: return hipErrorInvalidValue;
    }
#undef BMF_SAMPLE
    return hipGetLastError();
}
hipError_t launch_count(const Arenas& A, int n, int blocks, int wclasses, hipStream_t s) {
    if (wclasses & 1) hipLaunchKernelGGL(k_count<1>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 2) hipLaunchKernelGGL(k_count<2>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 4) hipLaunchKernelGGL(k_count<3>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 8) hipLaunchKernelGGL(k_count<4>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_scan(const Arenas& A, int n, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_scan, dim3(blocks), dim3(kScanThreads), 0, s, A, n);
    if (n <= kPlanOneBlock) {
        hipLaunchKernelGGL(k_plan_smooth, dim3(1), dim3(kPlanThreads), 0, s, A, n);
    } else {
        const int nb = (n + kPlanThreads - 1) / kPlanThreads;
        hipLaunchKernelGGL(k_smooth_count, dim3(nb), dim3(kPlanThreads), 0, s, A, n);
        hipLaunchKernelGGL(k_smooth_ranks, dim3(1), dim3(kPlanThreads), 0, s, A, nb);
        hipLaunchKernelGGL(k_smooth_place, dim3(nb), dim3(kPlanThreads), 0, s, A, n, nb);
        hipLaunchKernelGGL(k_smooth_bases, dim3(1), dim3(kPlanThreads), 0, s, A, nb);
        hipLaunchKernelGGL(k_smooth_prefix, dim3(nb), dim3(kPlanThreads), 0, s, A, n, nb);
    }
    return hipGetLastError();
}
hipError_t launch_emit(const Arenas& A, int n, int blocks, int wclasses, hipStream_t s) {
    if (wclasses & 1) hipLaunchKernelGGL(k_emit<1>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 2) hipLaunchKernelGGL(k_emit<2>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 4) hipLaunchKernelGGL(k_emit<3>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 8) hipLaunchKernelGGL(k_emit<4>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_refine(const Arenas& A, int n, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_refine, dim3(blocks), dim3(kSampleThreads), 0, s, A, n);
    return hipGetLastError();
}
hipError_t launch_smooth_map(const Arenas& A, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_smooth_map, dim3(std::min(blocks, 1024)), dim3(kSmoothThreads), 0, s, A);
    return hipGetLastError();
}
hipError_t launch_smooth(const Arenas& A, int m, int blocks, hipStream_t s) {
    blocks = blocks < kXcds ? kXcds : blocks - blocks % kXcds;  // whole XCD groups
    if (m == 0) hipLaunchKernelGGL(k_smooth_map, dim3(std::min(blocks, 1024)), dim3(kSmoothThreads), 0, s, A);
    if (A.normals) {
        hipLaunchKernelGGL(k_dual_nrm, dim3(blocks), dim3(kSmoothThreads), 0, s, A, m);
        hipLaunchKernelGGL(k_primal_nrm, dim3(blocks), dim3(kSmoothThreads), 0, s, A, m);
    } else {
        hipLaunchKernelGGL(k_dual, dim3(blocks), dim3(kSmoothThreads), 0, s, A, m);
        hipLaunchKernelGGL(k_primal, dim3(blocks), dim3(kSmoothThreads), 0, s, A, m);
    }
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
