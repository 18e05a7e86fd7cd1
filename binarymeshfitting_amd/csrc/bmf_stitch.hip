// Seams between the chunks of a world octree (C ABI bmf_stitch in
// include/bmf.h): the marching-cubes triangles of the dual cells that
// straddle chunk boundaries, the output of WorldStitcher::stitch_all
// (WorldStitcher.cpp:26-49) over the chunks' shell octrees
// (DMCChunk::generate_octree, DMCChunk.cpp:700-778).
//
// The reference walks the world octree recursively (stitch_cell, the face
// and edge recursions, stitch_indexes: WorldStitcher.cpp:184-572); here the
// same dual cells are enumerated point by point, one thread per candidate
// point, in a fixed order (the chunks of the leaf list in order, the lattice
// points of each chunk's closed cube [0, dim]^3 on the cube's boundary in
// x-major order):
//   - integer geometry in units of the finest unit node: a node at level l
//     with morton code 1 d1 .. dl spans [sum_k d_k 2^(L-k) dim, + 2^(L-l) dim)
//     per axis (L the deepest level);
//   - a point is stitched when its 8 octant neighbourhoods lie in the world
//     but not all in one chunk, by the finest of those chunks (the lowest
//     node index among equals);
//   - corner c (4*dx + 2*dy + dz) is the unit node of octant c's chunk that
//     contains the point's octant c, placed where the shell octree places it
//     (pos + MCD * size/2 + MCD * size/4 + ..., FP32 in that order), valued
//     by the sampler there (every octave);
//   - marching cubes as stitch_indexes: crossing (v1 - v0) * mu + v0,
//     mu = (0 - s0) / (s1 - s0), triangles from the tri_table.
// Three launches: count (triangles per candidate, and per 256-candidate
// tile), a one-workgroup scan of the tile totals, emit (recomputes the
// cells of candidates with triangles and writes them at their offsets).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/bmf.h"
#include "bmf_context.h"
#include "bmf_device_tables.h"
#include "bmf_internal.h"
#include "samplers.h"

namespace bmf {
namespace {

constexpr int kStitchThreads = 256;  // candidates per tile
constexpr int kScanThreads = 1024;

#define BMF_HIP(call)                          \
    do {                                       \
        hipError_t e_ = (call);                \
        if (e_ != hipSuccess) return BMF_EHIP; \
    } while (0)

// the child of a node in direction (dx, dy, dz), index dx | dy << 1 | dz << 2
// (children are numbered by the corner tables MCDX/Y/Z, Tables.hpp:7-9)
__constant__ int c_child_of[8] = {0, 1, 4, 5, 3, 2, 7, 6};
// corners of the 12 local edges in the tables' numbering
__constant__ int c_edge[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                                  {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};

struct World {
    const bmf_world_node* nodes;
    int n_nodes;
    int root;
    int L;     // deepest level
    int dim;   // unit nodes per chunk edge (a power of two)
};

__device__ __forceinline__ void node_box(const World& W, int i, int64_t o[3], int64_t& ext) {
    const int l = W.nodes[i].level;
    const uint64_t m = W.nodes[i].morton;
    o[0] = o[1] = o[2] = 0;
    for (int k = 1; k <= l; k++) {
        const int d = (int)((m >> (3 * (l - k))) & 7u);
        const int64_t w = ((int64_t)1 << (W.L - k)) * W.dim;
        o[0] += (d & 1) * w;
        o[1] += ((d >> 1) & 1) * w;
        o[2] += ((d >> 2) & 1) * w;
    }
    ext = ((int64_t)1 << (W.L - l)) * W.dim;
}

// The world leaf containing the point p2 / 2 (half units), or -1.
__device__ int locate(const World& W, const int64_t p2[3]) {
    int n = W.root;
    for (int guard = 0; guard <= W.L + 1 && n >= 0 && n < W.n_nodes; guard++) {
        int64_t o[3], e;
        node_box(W, n, o, e);
        for (int a = 0; a < 3; a++)
            if (p2[a] < 2 * o[a] || p2[a] >= 2 * (o[a] + e)) return -1;
        if (W.nodes[n].world_leaf) return n;
        const int dx = p2[0] >= 2 * o[0] + e, dy = p2[1] >= 2 * o[1] + e, dz = p2[2] >= 2 * o[2] + e;
        n = W.nodes[n].children[c_child_of[dx | (dy << 1) | (dz << 2)]];
    }
    return -1;
}

// Lattice index (i, j, k) of candidate t among the boundary points of the
// cube [0, d]^3 in x-major order.
__device__ __forceinline__ void boundary_point(int64_t t, int d, int& i, int& j, int& k) {
    const int64_t face = (int64_t)(d + 1) * (d + 1), ring = 4ll * d;
    if (t < face) {
        i = 0;
        j = (int)(t / (d + 1));
        k = (int)(t % (d + 1));
        return;
    }
    t -= face;
    if (t >= (int64_t)(d - 1) * ring) {
        t -= (int64_t)(d - 1) * ring;
        i = d;
        j = (int)(t / (d + 1));
        k = (int)(t % (d + 1));
        return;
    }
    i = 1 + (int)(t / ring);
    int s = (int)(t % ring);
    if (s < d + 1) {
        j = 0;
        k = s;
    } else if (s < d + 1 + 2 * (d - 1)) {
        const int u = s - (d + 1);
        j = 1 + u / 2;
        k = (u & 1) * d;
    } else {
        j = d;
        k = s - (d + 1) - 2 * (d - 1);
    }
}

// The dual cell of candidate (leaf list entry li, boundary point t): returns
// false when the point is not stitched by this chunk; else its corners.
__device__ bool dual_cell(const World& W, const SamplerDev& S, const int32_t* __restrict__ leaves, int li, int64_t t,
                          float pos[8][3], float smp[8]) {
    const int Lf = leaves[li];
    if (Lf < 0 || Lf >= W.n_nodes) return false;
    int64_t o[3], e;
    node_box(W, Lf, o, e);
    const int64_t u = e / W.dim;
    int i, j, k;
    boundary_point(t, W.dim, i, j, k);
    const int64_t P[3] = {o[0] + i * u, o[1] + j * u, o[2] + k * u};
    int oc[8];
    bool same = true;
    int owner = -1;
    int64_t owner_ext = 0;
    for (int c = 0; c < 8; c++) {
        const int64_t p2[3] = {2 * P[0] + ((c & 4) ? 1 : -1), 2 * P[1] + ((c & 2) ? 1 : -1),
                               2 * P[2] + ((c & 1) ? 1 : -1)};
        oc[c] = locate(W, p2);
        if (oc[c] < 0) return false;
        same = same && oc[c] == oc[0];
        int64_t oo[3], ee;
        node_box(W, oc[c], oo, ee);
        if (owner < 0 || ee < owner_ext || (ee == owner_ext && oc[c] < owner)) {
            owner = oc[c];
            owner_ext = ee;
        }
    }
    if (same || owner != Lf) return false;
    for (int c = 0; c < 8; c++) {
        const bmf_world_node& q = W.nodes[oc[c]];
        int64_t oo[3], ee;
        node_box(W, oc[c], oo, ee);
        const int64_t uc = ee / W.dim;
        const int d[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
        int64_t idx[3];
        for (int a = 0; a < 3; a++) {
            const int64_t rel = P[a] - oo[a];
            idx[a] = d[a] ? rel / uc : (rel + uc - 1) / uc - 1;
        }
        float p[3] = {q.pos[0], q.pos[1], q.pos[2]};
        float size = q.size;
        for (int half = W.dim >> 1; half >= 1; half >>= 1) {
            const float cs = size * 0.5f;
            for (int a = 0; a < 3; a++) p[a] = p[a] + (float)((idx[a] / half) & 1) * cs;
            size = cs;
        }
        pos[c][0] = p[0];
        pos[c][1] = p[1];
        pos[c][2] = p[2];
        smp[c] = sample_density(S, p[0], p[1], p[2]);
    }
    return true;
}

__device__ __forceinline__ int cell_mask(const float smp[8]) {
    int mask = 0;
    for (int c = 0; c < 8; c++) mask |= (smp[c] < 0.0f) << c;
    return mask;
}

__device__ __forceinline__ int tri_count(int mask) {
    return (mask == 0 || mask == 255) ? 0 : (int)(c_tri_table[mask] >> 60) / 3;
}

__global__ void __launch_bounds__(kStitchThreads) k_stitch_count(World W, const SamplerDev* __restrict__ S,
                                                                 const int32_t* __restrict__ leaves, int64_t per_leaf,
                                                                 int64_t n_cand, uint8_t* __restrict__ counts,
                                                                 int64_t* __restrict__ tile_tot) {
    __shared__ int s_w[kStitchThreads / 64];
    const int64_t cand = (int64_t)blockIdx.x * kStitchThreads + threadIdx.x;
    int n = 0;
    if (cand < n_cand) {
        float pos[8][3], smp[8];
        if (dual_cell(W, *S, leaves, (int)(cand / per_leaf), cand % per_leaf, pos, smp))
            n = tri_count(cell_mask(smp));
        counts[cand] = (uint8_t)n;
    }
    int x = n;
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < kStitchThreads / 64; w++) t += s_w[w];
        tile_tot[blockIdx.x] = t;
    }
}

// Exclusive scan of the tile totals in place; out_total[0] = the sum.
__global__ void __launch_bounds__(kScanThreads) k_stitch_scan(int64_t* __restrict__ tile, int64_t n_tiles,
                                                              int64_t* __restrict__ out_total) {
    __shared__ int64_t s[kScanThreads];
    const int t = threadIdx.x;
    const int64_t seg = (n_tiles + kScanThreads - 1) / kScanThreads;
    const int64_t a0 = std::min<int64_t>(n_tiles, t * seg), a1 = std::min<int64_t>(n_tiles, a0 + seg);
    int64_t sum = 0;
    for (int64_t i = a0; i < a1; i++) sum += tile[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int64_t v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int64_t run = s[t] - sum;
    for (int64_t i = a0; i < a1; i++) {
        const int64_t v = tile[i];
        tile[i] = run;
        run += v;
    }
    if (t == kScanThreads - 1) out_total[0] = s[t];
}

__global__ void __launch_bounds__(kStitchThreads) k_stitch_emit(World W, const SamplerDev* __restrict__ S,
                                                                const int32_t* __restrict__ leaves, int64_t per_leaf,
                                                                int64_t n_cand, const uint8_t* __restrict__ counts,
                                                                const int64_t* __restrict__ tile_off,
                                                                const int64_t* __restrict__ total,
                                                                float* __restrict__ out, int64_t cap) {
    __shared__ int s_w[kStitchThreads / 64];
    if (total[0] > cap) return;  // reported by the host: nothing is written
    const int64_t cand = (int64_t)blockIdx.x * kStitchThreads + threadIdx.x;
    const int n = cand < n_cand ? counts[cand] : 0;
    // exclusive prefix of the counts within the tile, in candidate order
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = n;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; w++) before += s_w[w];
    const int64_t base = tile_off[blockIdx.x] + before + (x - n);
    if (n == 0) return;
    float pos[8][3], smp[8];
    if (!dual_cell(W, *S, leaves, (int)(cand / per_leaf), cand % per_leaf, pos, smp)) return;
    const int mask = cell_mask(smp);
    float cross[12][3];
    for (int e = 0; e < 12; e++) {
        const int a = c_edge[e][0], b = c_edge[e][1];
        if (((mask >> a) & 1) == ((mask >> b) & 1)) continue;
        const float mu = (0.0f - smp[a]) / (smp[b] - smp[a]);
        for (int q = 0; q < 3; q++) {
            const float d = (pos[b][q] - pos[a][q]) * mu;
            cross[e][q] = d + pos[a][q];
        }
    }
    const uint64_t row = c_tri_table[mask];
    for (int tr = 0; tr < n; tr++) {
        float* o = out + 9 * (base + tr);
        for (int v = 0; v < 3; v++) {
            const int e = (int)((row >> (4 * (3 * tr + v))) & 15);
            o[3 * v] = cross[e][0];
            o[3 * v + 1] = cross[e][1];
            o[3 * v + 2] = cross[e][2];
        }
    }
}

__global__ void k_stitch_world(const bmf_world_node* __restrict__ nodes, int n_nodes, int* __restrict__ out) {
    // out[0]: the first node without a parent, out[1]: the deepest level
    __shared__ int s_root, s_level;
    if (threadIdx.x == 0) {
        s_root = 0x7FFFFFFF;
        s_level = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_nodes; i += blockDim.x) {
        if (nodes[i].parent < 0) atomicMin(&s_root, i);
        atomicMax(&s_level, nodes[i].level);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = s_root == 0x7FFFFFFF ? -1 : s_root;
        out[1] = s_level;
    }
}

}  // namespace
}  // namespace bmf

using namespace bmf;

extern "C" {

int bmf_stitch(bmf_ctx* c, const bmf_world_node* nodes, int32_t n_nodes, const int32_t* leaves, int32_t n_leaves,
               const bmf_stitch_params* params, float* triangles, int64_t cap_triangles, int64_t* n_triangles) {
    if (!c || !params || n_nodes < 0 || n_leaves < 0 || cap_triangles < 0) return BMF_EINVAL;
    if ((n_nodes > 0 && !nodes) || (n_leaves > 0 && !leaves) || (cap_triangles > 0 && !triangles)) return BMF_EINVAL;
    const int dim = params->chunk_resolution;
    if (dim < 2 || dim > BMF_MAX_DIM || (dim & (dim - 1))) return BMF_EINVAL;
    if (!c->sampler_set) return BMF_ESTATE;
    if (c->h_sampler.kind == BMF_SAMPLER_EXTERNAL) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t per_leaf = 2ll * (dim + 1) * (dim + 1) + 4ll * dim * (dim - 1);
    const int64_t n_cand = per_leaf * n_leaves;
    const int64_t n_tiles = (n_cand + kStitchThreads - 1) / kStitchThreads;
    if (n_tiles > 0x7FFFFFFF) return BMF_EINVAL;
    if (n_cand > c->stitch_cap_cand) {
        if (c->d_stitch_counts) (void)hipFree(c->d_stitch_counts);
        if (c->d_stitch_tiles) (void)hipFree(c->d_stitch_tiles);
        c->d_stitch_counts = nullptr;
        c->d_stitch_tiles = nullptr;
        c->stitch_cap_cand = 0;
        BMF_HIP(hipMalloc((void**)&c->d_stitch_counts, (size_t)std::max<int64_t>(n_cand, 1)));
        BMF_HIP(hipMalloc((void**)&c->d_stitch_tiles, sizeof(int64_t) * (size_t)std::max<int64_t>(n_tiles, 1)));
        c->stitch_cap_cand = n_cand;
    }
    if (!c->d_stitch_misc) BMF_HIP(hipMalloc((void**)&c->d_stitch_misc, sizeof(int64_t) * 2));
    if (!c->h_stitch_total) BMF_HIP(hipHostMalloc((void**)&c->h_stitch_total, sizeof(int64_t), hipHostMallocDefault));
    // the world's root and deepest level (needed on the host for the launch)
    int* dw = reinterpret_cast<int*>(c->d_stitch_misc);
    hipLaunchKernelGGL(k_stitch_world, dim3(1), dim3(1024), 0, s, nodes, n_nodes, dw);
    int hw[2] = {-1, 0};
    BMF_HIP(hipMemcpyAsync(hw, dw, sizeof(hw), hipMemcpyDeviceToHost, s));
    BMF_HIP(hipStreamSynchronize(s));
    if (n_cand == 0 || hw[0] < 0) {
        if (n_triangles) *n_triangles = 0;
        return BMF_OK;
    }
    if ((int64_t)hw[1] + 8 + 1 > 62) return BMF_EINVAL;  // integer coordinates must fit
    const World W{nodes, n_nodes, hw[0], hw[1], dim};
    int64_t* total = c->d_stitch_misc + 1;
    hipLaunchKernelGGL(k_stitch_count, dim3((unsigned)n_tiles), dim3(kStitchThreads), 0, s, W, c->A.sampler, leaves,
                       per_leaf, n_cand, c->d_stitch_counts, c->d_stitch_tiles);
    hipLaunchKernelGGL(k_stitch_scan, dim3(1), dim3(kScanThreads), 0, s, c->d_stitch_tiles, n_tiles, total);
    hipLaunchKernelGGL(k_stitch_emit, dim3((unsigned)n_tiles), dim3(kStitchThreads), 0, s, W, c->A.sampler, leaves,
                       per_leaf, n_cand, c->d_stitch_counts, c->d_stitch_tiles, total, triangles, cap_triangles);
    BMF_HIP(hipGetLastError());
    BMF_HIP(hipMemcpyAsync(c->h_stitch_total, total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    BMF_HIP(hipStreamSynchronize(s));
    if (n_triangles) *n_triangles = *c->h_stitch_total;
    return *c->h_stitch_total > cap_triangles ? BMF_ECAPACITY : BMF_OK;
}

}  // extern "C"
