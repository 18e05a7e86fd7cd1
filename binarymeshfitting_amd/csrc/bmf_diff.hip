// Differentiable extraction of one chunk (C ABI in include/bmf.h):
//   forward   the extraction kernels with every octave evaluated, then the
//             smoothing on the tape's own copy of the mesh, keeping the
//             positions every `checkpoint_every` iterations;
//   loss      mean squared distance to the nearest of a set of targets
//             (exact nearest neighbour: an implicit binary box hierarchy
//             over Morton-sorted targets, traversed nearest child first);
//   backward  the transpose of the K smoothing steps (linear in the
//             positions for the frozen topology) and of vertex placement
//             mu = (0 - s0) / (s1 - s0) (DMCChunk.cpp:657-674), gathered per
//             density sample in a fixed order (deterministic, no atomics).
//
// Tape layout in HBM (one chunk, V vertices, T triangles):
//   dens  FP32 [dim^3]      the densities of the forward
//   vid0  u32 [dim^3]       first vertex id owned by each sample (valid
//                           where the sample owns a sign-changing edge)
//   vedge u32 [V], boundary u8 [V], idx u32 [3T]   copies of the mesh
//   ring  u32x2 [kMaxRing][V] ring of each vertex: its triangles in
//                           ascending id (the adjacency of
//                           MeshProcessor::set_adj, MeshProcessor.cpp:98-128),
//                           each as its two other corners and the vertex's
//                           place among the three; slot-major so that a
//                           wave's loads of one slot are contiguous
//   deg u8 [V], mv u8 [V]   ring size; whether the primal step moves the
//                           vertex (MeshProcessor.cpp:247)
//   bscale f32 [V]          1 / (3 |ring|) for moved vertices, else 1 (the
//                           scale of the backward's gradient records)
//   p4, q4 float4 [V]       working positions (ping-pong); ckpt float4 [slots][V]
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>

#include "../../include/bmf.h"
#include "bmf_context.h"
#include "bmf_device_tables.h"
#include "bmf_internal.h"

using namespace bmf;

namespace {

constexpr int kThreads = 256;
constexpr int kStepThreads = 64;  // smoothing steps: one wave per workgroup spreads a small mesh over more CUs
// A vertex's lattice edge borders 4 cells and one cell's triangulation uses
// one edge in at most 5 table slots (MCTable.h:27-283), so a ring holds at
// most 20 triangles.
constexpr int kMaxRing = 20;
// Vertex ids share a word with a corner place (2 bits) in the ring records.
constexpr uint32_t kVidMask = 0x3FFFFFFFu;
constexpr int64_t kMaxTapeVertices = int64_t(1) << 30;
constexpr int kLeaf = 64;      // targets per leaf box of the nearest-target hierarchy
constexpr int kNNThreads = 256;  // four vertices per workgroup, one wave each
constexpr int kNNFan = 6;        // a wave tests 2^kNNFan = 64 boxes of one subtree level at once
constexpr int kNNFrames = 6;     // frames of the wave's walk: root + ceil(28 / kNNFan) levels for n < 2^31

#define BMF_HIP(call)                          \
    do {                                       \
        hipError_t e_ = (call);                \
        if (e_ != hipSuccess) return BMF_EHIP; \
    } while (0)

template <class T>
hipError_t dalloc(T** p, int64_t count) {
    return hipMalloc((void**)p, (size_t)std::max<int64_t>(count, 1) * sizeof(T));
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kThreads - 1) / kThreads); }
unsigned step_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kStepThreads - 1) / kStepThreads); }

// ------------------------------------------------------------------------
// forward
// ------------------------------------------------------------------------
// Takes the chunk's extraction into the tape in one launch: densities, edge
// codes, boundary flags and triangles copied out of the context's arenas,
// the positions as float4 (working copy and checkpoint 0, and as float3
// when no smoothing follows), and each sample's first vertex id (vertices
// are numbered in sample order, x-, y-, z-edge within a sample).
__global__ void k_tape_take(const float* __restrict__ dens_in, const uint32_t* __restrict__ vedge_in,
                            const uint8_t* __restrict__ bnd_in, const uint32_t* __restrict__ idx_in,
                            const float* __restrict__ pos_in, int64_t n3, int V, int64_t n_idx, int dim,
                            float* __restrict__ dens, uint32_t* __restrict__ vedge, uint8_t* __restrict__ bnd,
                            uint32_t* __restrict__ idx, float4* __restrict__ p4, float4* __restrict__ ck0,
                            float* __restrict__ pos3, uint32_t* __restrict__ vid0) {
    const int64_t total = std::max<int64_t>(n3, std::max<int64_t>(n_idx, V));
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        if (i < n3) dens[i] = dens_in[i];
        if (i < n_idx) idx[i] = idx_in[i];
        if (i < V) {
            const uint32_t code = vedge_in[i];
            vedge[i] = code;
            bnd[i] = bnd_in[i];
            const float x = pos_in[3 * i], y = pos_in[3 * i + 1], z = pos_in[3 * i + 2];
            p4[i] = make_float4(x, y, z, 0.0f);
            ck0[i] = make_float4(x, y, z, 0.0f);
            if (pos3) {
                pos3[3 * i] = x;
                pos3[3 * i + 1] = y;
                pos3[3 * i + 2] = z;
            }
            const uint32_t smp = code & 0xFFFFFFu;
            if (i == 0 || (vedge_in[i - 1] & 0xFFFFFFu) != smp) {
                const int sx = (int)(smp & 255u), sy = (int)((smp >> 8) & 255u), sz = (int)(smp >> 16);
                vid0[((int64_t)sx * dim + sy) * dim + sz] = (uint32_t)i;
            }
        }
    }
}

// Ring of each vertex from the per-cell triangle bases and masks written by
// the emit kernel: the 4 cells around its edge in lexicographic order and,
// in each, the table slots on its local edge in slot order.  Each triangle
// is stored as its other two corners in corner order, the first with the
// vertex's own place (0, 1, 2) in bits 30-31.  Also the vertex's backward
// scale: 1 / (3 |ring|) if the primal step moves it, else 1.
__global__ void __launch_bounds__(kStepThreads) k_tape_adj(
    const uint32_t* __restrict__ vedge, const uint8_t* __restrict__ boundary, const uint32_t* __restrict__ cell_tb,
    const uint8_t* __restrict__ cell_mask, const uint32_t* __restrict__ idx, int dim, int V, int T, int bp,
    int stride, uint2* __restrict__ ring, uint8_t* __restrict__ deg, uint8_t* __restrict__ mv,
    float* __restrict__ bscale, int* __restrict__ err) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t code = vedge[v];
    const int ex = (int)(code & 255u), ey = (int)((code >> 8) & 255u), ez = (int)((code >> 16) & 255u);
    const int ax = (int)(code >> 24);
    int n = 0;
    for (int c = 0; c < 4; c++) {
        const int cx = ex + c_ring[ax][c][0], cy = ey + c_ring[ax][c][1], cz = ez + c_ring[ax][c][2];
        if (cx < 0 || cy < 0 || cz < 0 || cx > dim - 2 || cy > dim - 2 || cz > dim - 2) continue;
        const int64_t ci = ((int64_t)cx * dim + cy) * dim + cz;
        const uint32_t le = (uint32_t)c_ring[ax][c][3];
        const uint32_t tb = cell_tb[ci];
        const uint64_t row = c_tri_table[cell_mask[ci]];
        const int ns = (int)(row >> 60);
        for (int s = 0; s < ns; s++) {
            if (((row >> (4 * s)) & 15ull) != le) continue;
            const uint32_t t = tb + (uint32_t)(s / 3);
            if (t >= (uint32_t)T || n >= kMaxRing) {
                *err = 1;
                continue;
            }
            const uint32_t a = idx[3 * t], b = idx[3 * t + 1], d = idx[3 * t + 2];
            const uint32_t place = a == (uint32_t)v ? 0u : (b == (uint32_t)v ? 1u : 2u);
            if (place == 2u && d != (uint32_t)v) *err = 1;
            const uint32_t o1 = place == 0u ? b : a, o2 = place == 2u ? b : d;
            ring[(int64_t)n * stride + v] = make_uint2(o1 | (place << 30), o2);
            n++;
        }
    }
    const bool moved = (bp || !boundary[v]) && n > 0;
    deg[v] = (uint8_t)n;
    mv[v] = (uint8_t)moved;
    bscale[v] = moved ? 1.0f / (3.0f * (float)n) : 1.0f;
}

// x / 3 rounded to nearest in three FP32 operations: q0 = x RN(1/3), the
// exact remainder r = x - 3 q0 (one fma), then q0 + r RN(1/3) (one fma) --
// the correction step of Markstein's division (IA-64 and Elementary
// Functions, 2000, ch. 8); zeros and infinities take q0.  Equal to the IEEE
// quotient for every float (tests/test_exact_division.py runs the same
// operations on the host over all 2^32 of them).
__device__ __forceinline__ float div3(float x) {
    const float third = 1.0f / 3.0f;
    const float q0 = x * third;
    const float q = __builtin_fmaf(__builtin_fmaf(-q0, 3.0f, x), third, q0);
    // v_cmp_class mask: -inf (bit 2), -0 (5), +0 (6), +inf (9)
    return __builtin_amdgcn_classf(x, (1 << 2) | (1 << 5) | (1 << 6) | (1 << 9)) ? q0 : q;
}

// The new position of one moved vertex: its ring's centroids summed from +0
// in ascending triangle order, divided by the ring size.
__device__ __forceinline__ float4 smooth_vertex(const uint2* __restrict__ ring, int stride, int n,
                                               const float4* __restrict__ in, float4 pv, int v) {
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll 4
    for (int k = 0; k < n; k++) {
        const uint2 e = ring[(int64_t)k * stride + v];
        const uint32_t place = e.x >> 30;
        const float4 a = in[e.x & kVidMask], b = in[e.y];
        const float4 c0 = place == 0u ? pv : a;
        const float4 c1 = place == 0u ? a : (place == 1u ? pv : b);
        const float4 c2 = place == 2u ? pv : b;
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        cx = cx + c0.x;
        cy = cy + c0.y;
        cz = cz + c0.z;
        cx = cx + c1.x;
        cy = cy + c1.y;
        cz = cz + c1.z;
        cx = cx + c2.x;
        cy = cy + c2.y;
        cz = cz + c2.z;
        sx = sx + div3(cx);
        sy = sy + div3(cy);
        sz = sz + div3(cz);
    }
    const float fc = (float)n;
    return make_float4(sx / fc, sy / fc, sz / fc, 0.0f);
}

// One smoothing iteration (MeshProcessor.cpp:130-306) in one launch: the new
// position of a moved vertex is its ring's centroids summed from +0 in
// ascending triangle order and divided by the ring size, each centroid
// ((0 + p0) + p1) + p2, / 3 recomputed from the corners by every vertex of
// the triangle (the same FP32 value each time, so no separate dual pass);
// other vertices keep theirs.  Reads `in`, writes `out` (ping-pong buffers)
// and, when given, a checkpoint copy `ck` and float3 positions `pos3`.
__global__ void __launch_bounds__(kStepThreads) k_tape_step(const uint2* __restrict__ ring, int stride,
                                                            const uint8_t* __restrict__ deg,
                                                            const uint8_t* __restrict__ mv,
                                                            const float4* __restrict__ in, int V,
                                                            float4* __restrict__ out, float4* __restrict__ ck,
                                                            float* __restrict__ pos3) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const float4 pv = in[v];
    float4 r = pv;
    if (mv[v]) r = smooth_vertex(ring, stride, deg[v], in, pv, v);
    out[v] = r;
    if (ck) ck[v] = r;
    if (pos3) {
        pos3[3 * v] = r.x;
        pos3[3 * v + 1] = r.y;
        pos3[3 * v + 2] = r.z;
    }
}

__global__ void k_tape_out(const float4* __restrict__ p4, int V, float* __restrict__ pos) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const float4 q = p4[v];
    pos[3 * v] = q.x;
    pos[3 * v + 1] = q.y;
    pos[3 * v + 2] = q.z;
}

// ------------------------------------------------------------------------
// backward
// ------------------------------------------------------------------------
// The backward's working record of vertex v: .w = [v moved] and .xyz =
// dL/dp_v scaled by bscale_v = 1 / (3 |ring(v)|) for moved vertices, so that
// a row of the transposed step needs nothing of its neighbours but their
// records.  With no smoothing (`scaled` = 0) the record is dL/dp_v itself.
__global__ void k_bwd_init(const float* __restrict__ g3, int V, const uint8_t* __restrict__ mv,
                           const float* __restrict__ bscale, int scaled, float4* __restrict__ g) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const float sc = scaled ? bscale[v] : 1.0f;
    g[v] = make_float4(g3[3 * v] * sc, g3[3 * v + 1] * sc, g3[3 * v + 2] * sc, scaled ? (float)mv[v] : 0.0f);
}

// One transposed smoothing iteration.  The forward step is
// p'_u = [u not moved] p_u + [u moved] / (3 |ring(u)|) * (sum over the
// triangles of ring(u) of their corners), so
//   dL/dp_v = [v not moved] dL/dp'_v
//             + sum over the triangles t of ring(v), over the corners u of t,
//               of [u moved] dL/dp'_u / (3 |ring(u)|);
// with the records x_u = [u moved] dL/dp'_u / (3 |ring(u)|) (w = [u moved])
// that is m_self x_v plus, per ring triangle, its two other corners' records
// times their w (m_self = |ring(v)| for a moved v, whose own record is
// scaled; 1 otherwise).  Writes v's record for the step before, or dL/dp_v
// itself after the first iteration (`last`).
__global__ void __launch_bounds__(kStepThreads) k_bwd_step(const uint2* __restrict__ ring, int stride,
                                                           const uint8_t* __restrict__ deg,
                                                           const uint8_t* __restrict__ mv,
                                                           const float* __restrict__ bscale,
                                                           const float4* __restrict__ in, int V, int last,
                                                           float4* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int n = deg[v];
    const bool moved = mv[v] != 0;
    const float4 own = in[v];
    const float ms = moved ? (float)n : 1.0f;
    float sx = ms * own.x, sy = ms * own.y, sz = ms * own.z;
#pragma unroll 4
    for (int k = 0; k < n; k++) {
        const uint2 e = ring[(int64_t)k * stride + v];
        const float4 a = in[e.x & kVidMask], b = in[e.y];
        sx = __builtin_fmaf(a.w, a.x, sx);
        sy = __builtin_fmaf(a.w, a.y, sy);
        sz = __builtin_fmaf(a.w, a.z, sz);
        sx = __builtin_fmaf(b.w, b.x, sx);
        sy = __builtin_fmaf(b.w, b.y, sy);
        sz = __builtin_fmaf(b.w, b.z, sz);
    }
    const float sc = last ? 1.0f : bscale[v];
    out[v] = make_float4(sx * sc, sy * sc, sz * sc, moved ? 1.0f : 0.0f);
}

// dL/d(density) per sample: the sample is s0 of its own sign-changing edges
// (+x, +y, +z, vertex ids vid0[i] + 0, 1, ...) and s1 of the sign-changing
// edges of its -x, -y, -z neighbours.  mu = -s0 / (s1 - s0), so
// dmu/ds0 = -s1 / (s1 - s0)^2 and dmu/ds1 = s0 / (s1 - s0)^2; the vertex
// moves along its edge axis by mu (grid units).
__global__ void k_bwd_place(const float* __restrict__ dens, const uint32_t* __restrict__ vid0,
                            const float4* __restrict__ g, int dim, float* __restrict__ out) {
    const uint32_t udim = (uint32_t)dim;
    const uint32_t n3 = udim * udim * udim;  // dim <= 256
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const uint32_t row = i / udim;
    const int z = (int)(i - row * udim), y = (int)(row % udim), x = (int)(row / udim);
    const int64_t st[3] = {(int64_t)dim * dim, (int64_t)dim, 1};
    const int co[3] = {x, y, z};
    const float di = dens[i];
    const bool li = di < 0.0f;
    float acc = 0.0f;
    int k = 0;
    for (int a = 0; a < 3; a++) {  // own edges: this sample is s0
        if (co[a] + 1 >= dim) continue;
        const float dj = dens[i + st[a]];
        if ((dj < 0.0f) == li) continue;
        const float4 q = g[vid0[i] + (uint32_t)k];
        k++;
        const float ds = dj - di;
        const float ga = a == 0 ? q.x : (a == 1 ? q.y : q.z);
        acc = acc + ga * (-dj / (ds * ds));
    }
    for (int a = 0; a < 3; a++) {  // edges of the -a neighbour: this sample is s1
        if (co[a] == 0) continue;
        const int64_t j = i - st[a];
        const float dj = dens[j];
        const bool lj = dj < 0.0f;
        if (lj == li) continue;
        // rank of j's axis-a edge among j's sign-changing edges
        uint32_t r = 0;
        for (int b = 0; b < a; b++) {
            const int cb = b == 0 ? x - (a == 0) : (b == 1 ? y - (a == 1) : z);
            if (cb + 1 < dim && ((dens[j + st[b]] < 0.0f) != lj)) r++;
        }
        const float4 q = g[vid0[j] + r];
        const float ds = di - dj;
        const float ga = a == 0 ? q.x : (a == 1 ? q.y : q.z);
        acc = acc + ga * (dj / (ds * ds));
    }
    out[i] = acc;
}

// ------------------------------------------------------------------------
// nearest-target loss
// ------------------------------------------------------------------------
__device__ __forceinline__ int float_order(float f) {
    const int b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float order_float(int b) { return __int_as_float(b >= 0 ? b : b ^ 0x7FFFFFFF); }

__global__ void k_nn_bbox_init(int* bb) {
    if (threadIdx.x < 3) bb[threadIdx.x] = 0x7FFFFFFF;
    else if (threadIdx.x < 6) bb[threadIdx.x] = (int)0x80000000;
}

// Bounding box of the targets: per-thread, per-wave, per-block reductions,
// then one atomic per block and coordinate (min/max are order-independent).
__global__ void __launch_bounds__(kThreads) k_nn_bbox(const float* __restrict__ t, int64_t n, int* bb) {
    __shared__ float s_r[6][kThreads / 64];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], t[3 * i + a]);
            hi[a] = fmaxf(hi[a], t[3 * i + a]);
        }
    const int wave = threadIdx.x >> 6;
    for (int a = 0; a < 3; a++) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[a] = fminf(lo[a], __shfl_xor(lo[a], off));
            hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off));
        }
        if ((threadIdx.x & 63) == 0) {
            s_r[a][wave] = lo[a];
            s_r[3 + a][wave] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float v = s_r[a][0];
        for (int w = 1; w < kThreads / 64; w++) v = a < 3 ? fminf(v, s_r[a][w]) : fmaxf(v, s_r[a][w]);
        if (a < 3) atomicMin(&bb[a], float_order(v));
        else atomicMax(&bb[a], float_order(v));
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton_key(const int* bb, float x, float y, float z) {
    const float p[3] = {x, y, z};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        const float lo = order_float(bb[a]), hi = order_float(bb[3 + a]);
        const float ext = hi - lo;
        float u = ext > 0.0f ? (p[a] - lo) * (1024.0f / ext) : 0.0f;
        u = fminf(fmaxf(u, 0.0f), 1023.0f);
        q[a] = (uint32_t)u;
    }
    return (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
}

__global__ void k_nn_keys(const float* __restrict__ t, int64_t n, const int* __restrict__ bb,
                          uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = morton_key(bb, t[3 * i], t[3 * i + 1], t[3 * i + 2]);
    vals[i] = (int32_t)i;
}

__global__ void k_nn_gather(const float* __restrict__ t, const int32_t* __restrict__ perm, int64_t n,
                            float4* __restrict__ st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t p = perm[i];
    st[i] = make_float4(t[3 * (int64_t)p], t[3 * (int64_t)p + 1], t[3 * (int64_t)p + 2], __int_as_float(p));
}

// Leaf boxes of the implicit hierarchy: heap-numbered nodes 1 .. 2P-1 over
// P (a power of two) leaves; leaf l is node P + l and holds sorted targets
// [kLeaf l, kLeaf (l+1)); leaves past the targets are empty boxes.
__global__ void k_nn_leaves(const float4* __restrict__ st, int64_t n, int64_t P, float4* __restrict__ lo,
                            float4* __restrict__ hi) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= P) return;
    float4 a = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), b = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    const int64_t e = std::min<int64_t>(n, (l + 1) * kLeaf);
    for (int64_t i = l * kLeaf; i < e; i++) {
        const float4 q = st[i];
        a.x = fminf(a.x, q.x);
        a.y = fminf(a.y, q.y);
        a.z = fminf(a.z, q.z);
        b.x = fmaxf(b.x, q.x);
        b.y = fmaxf(b.y, q.y);
        b.z = fmaxf(b.z, q.z);
    }
    lo[P + l] = a;
    hi[P + l] = b;
}

// Oriented bound of each leaf: the centre c of its targets, the principal
// axes of their covariance (cyclic Jacobi, then re-orthonormalised) and the
// largest |(q - c) . axis| per axis; four float4 per leaf (c, then each axis
// with its extent in w).  A patch of a curved target set is thin along its
// normal, where the axis-aligned box of a tilted patch is not.  Empty leaves
// get a zero record (their axis-aligned box already excludes them).
__global__ void k_nn_obb(const float4* __restrict__ st, int64_t n, int64_t P, float4* __restrict__ ob) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= P) return;
    const int64_t b = l * kLeaf, e = std::min<int64_t>(n, (l + 1) * kLeaf);
    float4* o = ob + 4 * l;
    if (b >= e) {
        for (int k = 0; k < 4; k++) o[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (int64_t i = b; i < e; i++) {
        const float4 q = st[i];
        cx += q.x;
        cy += q.y;
        cz += q.z;
    }
    const float inv = 1.0f / (float)(e - b);
    cx *= inv;
    cy *= inv;
    cz *= inv;
    float a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int64_t i = b; i < e; i++) {
        const float4 q = st[i];
        const float d[3] = {q.x - cx, q.y - cy, q.z - cz};
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) a[r][c] += d[r] * d[c];
    }
    float v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sw = 0; sw < 8; sw++)
#pragma unroll
        for (int pq = 0; pq < 3; pq++) {
            const int p0 = pq == 2 ? 1 : 0, q0 = pq == 0 ? 1 : 2;
            const float apq = a[p0][q0];
            if (!(fabsf(apq) > 1e-30f)) continue;
            const float th = (a[q0][q0] - a[p0][p0]) / (2.0f * apq);
            const float t = (th >= 0.0f ? 1.0f : -1.0f) / (fabsf(th) + sqrtf(th * th + 1.0f));
            const float cs = 1.0f / sqrtf(t * t + 1.0f), sn = t * cs;
#pragma unroll
            for (int k = 0; k < 3; k++) {  // columns p0, q0 of A J
                const float akp = a[k][p0], akq = a[k][q0];
                a[k][p0] = cs * akp - sn * akq;
                a[k][q0] = sn * akp + cs * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {  // rows p0, q0 of J^T (A J)
                const float apk = a[p0][k], aqk = a[q0][k];
                a[p0][k] = cs * apk - sn * aqk;
                a[q0][k] = sn * apk + cs * aqk;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float vkp = v[k][p0], vkq = v[k][q0];
                v[k][p0] = cs * vkp - sn * vkq;
                v[k][q0] = sn * vkp + cs * vkq;
            }
        }
    // axes: the first two eigenvector columns, Gram-Schmidt, and their cross product
    float u[3] = {v[0][0], v[1][0], v[2][0]}, w[3] = {v[0][1], v[1][1], v[2][1]};
    float lu = sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int k = 0; k < 3; k++) u[k] /= lu;
    const float dw = w[0] * u[0] + w[1] * u[1] + w[2] * u[2];
    for (int k = 0; k < 3; k++) w[k] -= dw * u[k];
    const float lw = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int k = 0; k < 3; k++) w[k] /= lw;
    const float nn[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    const float* ax[3] = {u, w, nn};
    float ex[3] = {0.0f, 0.0f, 0.0f};
    for (int64_t i = b; i < e; i++) {
        const float4 q = st[i];
        const float d[3] = {q.x - cx, q.y - cy, q.z - cz};
#pragma unroll
        for (int k = 0; k < 3; k++) ex[k] = fmaxf(ex[k], fabsf(d[0] * ax[k][0] + d[1] * ax[k][1] + d[2] * ax[k][2]));
    }
    o[0] = make_float4(cx, cy, cz, 0.0f);
    for (int k = 0; k < 3; k++) o[1 + k] = make_float4(ax[k][0], ax[k][1], ax[k][2], ex[k]);
}

// Inner boxes of one wide level [level, 2 level).
__global__ void k_nn_level(int64_t level, float4* __restrict__ lo, float4* __restrict__ hi) {
    const int64_t i = level + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * level) return;
    const float4 a = lo[2 * i], b = lo[2 * i + 1], c = hi[2 * i], d = hi[2 * i + 1];
    lo[i] = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
    hi[i] = make_float4(fmaxf(c.x, d.x), fmaxf(c.y, d.y), fmaxf(c.z, d.z), 0.0f);
}

// Inner boxes of the levels from `top` up to the root (one workgroup).
__global__ void __launch_bounds__(1024) k_nn_inner(int64_t top, float4* __restrict__ lo, float4* __restrict__ hi) {
    for (int64_t level = top; level >= 1; level >>= 1) {
        for (int64_t i = level + threadIdx.x; i < 2 * level; i += blockDim.x) {
            const float4 a = lo[2 * i], b = lo[2 * i + 1], c = hi[2 * i], d = hi[2 * i + 1];
            lo[i] = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
            hi[i] = make_float4(fmaxf(c.x, d.x), fmaxf(c.y, d.y), fmaxf(c.z, d.z), 0.0f);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ float dist2(float px, float py, float pz, float4 q) {
    const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
    return (dx * dx + dy * dy) + dz * dz;
}

// Squared distance from p to the box (0 inside, +inf for an empty box);
// never above the distance to any point in the box in FP32 (every step
// rounds monotonically).
__device__ __forceinline__ float box_lb2(float px, float py, float pz, float4 l, float4 h) {
    if (!(l.x <= h.x)) return INFINITY;
    const float dx = fmaxf(fmaxf(l.x - px, px - h.x), 0.0f);
    const float dy = fmaxf(fmaxf(l.y - py, py - h.y), 0.0f);
    const float dz = fmaxf(fmaxf(l.z - pz, pz - h.z), 0.0f);
    return (dx * dx + dy * dy) + dz * dz;
}

// Squared-distance lower bound from p to the targets of a leaf from its
// oriented record: per axis max(0, |(p - c) . axis| - extent - slack), summed
// and scaled down.  The slack (1e-5 of |p - c|_1 + extent per axis, then
// 1 - 1e-5 on the sum) is far above the FP32 errors of the projections, of
// the axes' orthonormality and of the targets' own FP32 distances (each of
// order 1e-7 relative), so the bound stays below the FP32 distance of every
// target of the leaf.
__device__ __forceinline__ float obb_lb2(float px, float py, float pz, const float4* __restrict__ o) {
    const float4 c = o[0], u = o[1], w = o[2], nrm = o[3];
    const float dx = px - c.x, dy = py - c.y, dz = pz - c.z;
    const float l1 = (fabsf(dx) + fabsf(dy)) + fabsf(dz);
    const float tu = fabsf((dx * u.x + dy * u.y) + dz * u.z), tw = fabsf((dx * w.x + dy * w.y) + dz * w.z),
                tn = fabsf((dx * nrm.x + dy * nrm.y) + dz * nrm.z);
    const float su = fmaxf(tu - u.w - 1e-5f * (l1 + u.w), 0.0f);
    const float sw = fmaxf(tw - w.w - 1e-5f * (l1 + w.w), 0.0f);
    const float sn = fmaxf(tn - nrm.w - 1e-5f * (l1 + nrm.w), 0.0f);
    return ((su * su + sw * sw) + sn * sn) * (1.0f - 1e-5f);
}

// fmin with the value DPP control CTRL brings from another lane (+inf where
// the source lane is outside the row or the row is not in ROW_MASK).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_fmin(float v) {
    const int o = __builtin_amdgcn_update_dpp(__float_as_int(INFINITY), __float_as_int(v), CTRL, ROW_MASK, 0xf, false);
    return fminf(v, __int_as_float(o));
}

// Minimum over the 64 lanes of a wave (all lanes active): prefix minima
// within each row of 16 lanes (row_shr 1, 2, 4, 8), then row 0's last lane
// into row 1 and row 2's into row 3 (row_bcast:15), then lane 31 into rows 2
// and 3 (row_bcast:31); lane 63 then holds the minimum.
__device__ __forceinline__ float wave_min(float x) {
    x = dpp_fmin<0x111, 0xf>(x);
    x = dpp_fmin<0x112, 0xf>(x);
    x = dpp_fmin<0x114, 0xf>(x);
    x = dpp_fmin<0x118, 0xf>(x);
    x = dpp_fmin<0x142, 0xa>(x);
    x = dpp_fmin<0x143, 0xc>(x);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// A wave-uniform 64-bit value held in every lane, moved to scalar registers.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Lowest lane of the (non-empty) set m whose value x is smallest.
__device__ __forceinline__ int wave_argmin(float x, uint64_t m, int lane) {
    const float y = (m >> lane) & 1 ? x : INFINITY;
    const float mn = wave_min(y);
    return __ffsll((unsigned long long)(__ballot(y == mn) & m)) - 1;
}

// Nearest target of each vertex, one wave per vertex.  The hierarchy (heap
// nodes 1 .. 2P-1, leaves at depth D = log2 P) is walked in frames: frame 0
// holds the root, and each later frame holds the 2^f descendants, f <=
// kNNFan levels down, of a node popped from the frame above (the first step
// takes D mod kNNFan levels so that the leaf frames are full).  Lane j
// tests descendant j; the frame keeps the set whose box is no farther than
// the best target distance U found so far, and pops its nearest box first.
// A frame of leaves is evaluated min(kLeaf, 64) lanes per leaf (each lane
// taking kLeaf / 64 targets of a larger leaf), 64 / kLeaf leaves per step
// when leaves are smaller than a wave, re-pruned against U between steps; the very first leaf frame starts
// with its nearest leaf alone so that U is finite early.  Pruning drops a
// box only when it is strictly farther than U, an actual target distance,
// so every target at the minimum distance is evaluated and the result is
// the lowest (distance, original index), as a brute-force scan gives.
__global__ void __launch_bounds__(kNNThreads) k_nn_query(const float* __restrict__ pos, int V,
                                                         const float4* __restrict__ st, int64_t n, int64_t P,
                                                         const float4* __restrict__ lo, const float4* __restrict__ hi,
                                                         const float4* __restrict__ ob,
                                                         int32_t* __restrict__ nearest, float* __restrict__ dldp,
                                                         double* __restrict__ d2) {
    const int v = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (v >= V) return;  // uniform across the wave
    const int lane = (int)(threadIdx.x & 63);
    const float px = pos[3 * (int64_t)v], py = pos[3 * (int64_t)v + 1], pz = pos[3 * (int64_t)v + 2];
    int D = 0;
    while ((int64_t)1 << D < P) D++;
    const int f0 = D == 0 ? 0 : D - kNNFan * ((D - 1) / kNNFan);
    float bd = INFINITY;  // this lane's best (distance, original index, sorted position)
    int bi = 0x7FFFFFFF;
    int bs = -1;  // sorted position (n < 2^31, heap indices < 2^30)
    float U = INFINITY;  // best distance over the wave
    // the frames of this wave: per-lane box distances, the live set, heap index of descendant 0
    __shared__ float s_lb[kNNThreads / 64][kNNFrames][64];
    __shared__ unsigned long long s_mk[kNNThreads / 64][kNNFrames];
    __shared__ int s_base[kNNThreads / 64][kNNFrames];
    const int w = (int)(threadIdx.x >> 6);
    s_lb[w][0][lane] = lane == 0 ? box_lb2(px, py, pz, lo[1], hi[1]) : INFINITY;
    s_mk[w][0] = 1;
    s_base[w][0] = 1;
    int depth = 1;
    const int Pi = (int)P, ni = (int)n;
    // a leaf's targets: kLane lanes each take kSub of them (lane, lane + kLane, ...)
    constexpr int kLane = kLeaf < 64 ? kLeaf : 64, kSub = kLeaf / kLane;
    auto eval = [&](int leaf) {
        // this lane's targets of `leaf` (heap index), if any
#pragma unroll
        for (int u = 0; u < kSub; u++) {
            const int i = (leaf - Pi) * kLeaf + u * kLane + (lane & (kLane - 1));
            if (i < ni) {
                const float4 q = st[i];
                const float d = dist2(px, py, pz, q);
                const int id = __float_as_int(q.w);
                if (d < bd || (d == bd && id < bi)) {
                    bd = d;
                    bi = id;
                    bs = i;
                }
            }
        }
    };
    while (depth > 0) {
        const int k = depth - 1;
        const int lk = k == 0 ? 0 : f0 + kNNFan * (k - 1);
        const float lb = s_lb[w][k][lane];
        // the frame's base and live set are the wave's (moved to scalar registers)
        const int b0 = __builtin_amdgcn_readfirstlane(s_base[w][k]);
        unsigned long long m = uniform_u64(s_mk[w][k]) & __ballot(lb <= U);
        if (lk == D) {
            if (U == INFINITY && m) {
                const int j = wave_argmin(lb, m, lane);
                if (lane < kLane) eval(b0 + j);
                m &= ~(1ull << j);
                U = wave_min(bd);
                m &= __ballot(lb <= U);
            }
            while (m) {
                // lane group g (kLane lanes) takes the g-th lowest live leaf:
                // the positions come off the wave-uniform mask on the scalar
                // unit, each group selects its own
                int j = -1;
#pragma unroll
                for (int t = 0; t < 64 / kLane; t++) {
                    const int pos = m ? __ffsll((unsigned long long)m) - 1 : -1;
                    j = lane / kLane == t ? pos : j;
                    m &= m - 1;
                }
                if (j >= 0) eval(b0 + j);
                U = wave_min(bd);
                m &= __ballot(lb <= U);
            }
            depth--;
            continue;
        }
        if (!m) {
            depth--;
            continue;
        }
        const int j = wave_argmin(lb, m, lane);
        s_mk[w][k] = m & ~(1ull << j);
        const int f = k == 0 ? f0 : kNNFan;
        const int c0 = (b0 + j) << f;
        float lc = lane < (1 << f) ? box_lb2(px, py, pz, lo[c0 + lane], hi[c0 + lane]) : INFINITY;
        // a frame of leaves also takes each leaf's oriented bound
        if (f0 + kNNFan * k == D && lane < (1 << f) && lc < INFINITY)
            lc = fmaxf(lc, obb_lb2(px, py, pz, ob + 4 * (int64_t)(c0 + lane - Pi)));
        s_lb[w][k + 1][lane] = lc;
        s_mk[w][k + 1] = __ballot(lane < (1 << f) && lc <= U);
        s_base[w][k + 1] = c0;
        depth++;
    }
    // the wave's lowest (distance, index)
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float od = __shfl_xor(bd, off);
        const int oi = __shfl_xor(bi, off);
        const int os = __shfl_xor(bs, off);
        if (od < bd || (od == bd && oi < bi)) {
            bd = od;
            bi = oi;
            bs = os;
        }
    }
    if (lane != 0) return;
    const float4 q = bs >= 0 ? st[bs] : make_float4(px, py, pz, 0.0f);
    nearest[v] = bs >= 0 ? bi : -1;
    const float sc = 2.0f / (float)V;
    dldp[3 * (int64_t)v] = (px - q.x) * sc;
    dldp[3 * (int64_t)v + 1] = (py - q.y) * sc;
    dldp[3 * (int64_t)v + 2] = (pz - q.z) * sc;
    d2[v] = bs >= 0 ? (double)bd : 0.0;
}

// Deterministic sum: thread t adds x[t], x[t + 1024], ... in order, then a
// fixed tree.
__global__ void __launch_bounds__(1024) k_sum(const double* __restrict__ x, int n, double* __restrict__ out) {
    __shared__ double s[1024];
    const int tid = threadIdx.x;
    double a = 0.0;
    int i = tid;
    // the same order as one add per stride step; the loads of 8 steps are
    // issued together
    for (; i + 7 * 1024 < n; i += 8 * 1024) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = x[i + j * 1024];
#pragma unroll
        for (int j = 0; j < 8; j++) a += v[j];
    }
    for (; i < n; i += 1024) a += x[i];
    s[tid] = a;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (tid < w) s[tid] += s[tid + w];
        __syncthreads();
    }
    if (tid == 0) out[0] = s[0] / (double)n;
}

}  // namespace

struct bmf_tape {
    bmf_ctx* ctx = nullptr;
    int32_t max_dim = 0;
    int64_t cap_v = 0, cap_i = 0;
    int32_t ckpt_every = 1;
    bool valid = false;
    int32_t dim = 0, K = 0, bp = 0, V = 0, T = 0, contains_mesh = 0;
    float* dens = nullptr;
    uint32_t* vid0 = nullptr;
    uint32_t* vedge = nullptr;
    uint8_t* boundary = nullptr;
    uint32_t* idx = nullptr;
    uint2* ring = nullptr;
    float* bscale = nullptr;
    uint8_t* deg = nullptr;
    uint8_t* mv = nullptr;
    float4* p4 = nullptr;
    float4* q4 = nullptr;
    float* pos = nullptr;
    float4* g = nullptr;
    float4* g2 = nullptr;   // ping-pong partner of g in the backward
    float4* r4 = nullptr;   // ping-pong partner of q4 in bmf_tape_positions
    int* err = nullptr;
    hipEvent_t ev[6] = {};  // start / end of the last forward, loss and backward
    bool timed[3] = {false, false, false};
    float4* ckpt = nullptr;
    int64_t ckpt_slots = 0;
    // nearest-target search: hierarchy buffers for up to nn_cap targets and
    // the target set last built into them (nn_n targets, 0 = none)
    int64_t nn_cap = 0;
    int64_t nn_n = 0;
    uint32_t* keys[2] = {nullptr, nullptr};
    int32_t* vals[2] = {nullptr, nullptr};
    float4* st = nullptr;
    float4* blo = nullptr;  // hierarchy boxes, heap-numbered [2P]
    float4* bhi = nullptr;
    float4* obb = nullptr;  // oriented bound of each leaf [4P]
    int* bb = nullptr;
    void* cub_tmp = nullptr;
    size_t cub_bytes = 0;
    double* d2 = nullptr;
    double* loss = nullptr;
    int32_t* nearest_tmp = nullptr;
};

namespace {

// Leaves of the target hierarchy, rounded up to a power of two.
int64_t leaf_pow2(int64_t n) {
    const int64_t leaves = (n + kLeaf - 1) / kLeaf;
    int64_t P = 1;
    while (P < leaves) P <<= 1;
    return P;
}

void free_nn(bmf_tape* t) {
    for (int i = 0; i < 2; i++) {
        dfree(t->keys[i]);
        dfree(t->vals[i]);
    }
    dfree(t->st);
    dfree(t->blo);
    dfree(t->bhi);
    dfree(t->obb);
    if (t->cub_tmp) (void)hipFree(t->cub_tmp);
    t->cub_tmp = nullptr;
    t->cub_bytes = 0;
    t->nn_cap = 0;
    t->nn_n = 0;
}

int reserve_nn(bmf_tape* t, int64_t n) {
    if (n <= t->nn_cap) return BMF_OK;
    free_nn(t);
    const int64_t P = leaf_pow2(n);
    bool ok = true;
    for (int i = 0; i < 2; i++)
        ok = ok && dalloc(&t->keys[i], n) == hipSuccess && dalloc(&t->vals[i], n) == hipSuccess;
    ok = ok && dalloc(&t->st, n) == hipSuccess && dalloc(&t->blo, 2 * P) == hipSuccess &&
         dalloc(&t->bhi, 2 * P) == hipSuccess && dalloc(&t->obb, 4 * P) == hipSuccess;
    if (ok) {
        hipcub::DoubleBuffer<uint32_t> kb(t->keys[0], t->keys[1]);
        hipcub::DoubleBuffer<int32_t> vb(t->vals[0], t->vals[1]);
        size_t bytes = 0;
        ok = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kb, vb, (int)n, 0, 30) == hipSuccess;
        ok = ok && hipMalloc(&t->cub_tmp, std::max<size_t>(bytes, 1)) == hipSuccess;
        t->cub_bytes = bytes;
    }
    if (!ok) {
        free_nn(t);
        return BMF_EHIP;
    }
    t->nn_cap = n;
    return BMF_OK;
}

// Buffers of the nearest-target hierarchy of n targets: Morton keys and
// target ids (double-buffered for the radix sort), the targets in Morton
// order (x, y, z, id), the boxes of the heap-numbered nodes 1 .. 2P-1 and
// the oriented bound of each leaf.
struct NNBuffers {
    uint32_t* keys[2];
    int32_t* vals[2];
    float4* st;
    float4* blo;
    float4* bhi;
    float4* ob;
    void* cub_tmp;
    size_t cub_bytes;
};

NNBuffers tape_nn(bmf_tape* t) {
    return NNBuffers{{t->keys[0], t->keys[1]}, {t->vals[0], t->vals[1]}, t->st, t->blo, t->bhi, t->obb, t->cub_tmp,
                     t->cub_bytes};
}

// Enqueues the hierarchy build of the n targets [n][3] into b: bounding box,
// Morton keys, radix sort of (key, id), sorted targets, leaf and inner boxes.
int nn_build(const float* targets, int64_t n, int* bb, const NNBuffers& b, hipStream_t s) {
    const int64_t P = leaf_pow2(n);
    hipLaunchKernelGGL(k_nn_bbox_init, dim3(1), dim3(64), 0, s, bb);
    hipLaunchKernelGGL(k_nn_bbox, dim3((unsigned)std::min<int64_t>(256, blocks_for(n))), dim3(kThreads), 0, s, targets,
                       n, bb);
    hipLaunchKernelGGL(k_nn_keys, dim3(blocks_for(n)), dim3(kThreads), 0, s, targets, n, bb, b.keys[0], b.vals[0]);
    BMF_HIP(hipGetLastError());
    hipcub::DoubleBuffer<uint32_t> kb(b.keys[0], b.keys[1]);
    hipcub::DoubleBuffer<int32_t> vb(b.vals[0], b.vals[1]);
    size_t bytes = b.cub_bytes;
    BMF_HIP(hipcub::DeviceRadixSort::SortPairs(b.cub_tmp, bytes, kb, vb, (int)n, 0, 30, s));
    hipLaunchKernelGGL(k_nn_gather, dim3(blocks_for(n)), dim3(kThreads), 0, s, targets, vb.Current(), n, b.st);
    hipLaunchKernelGGL(k_nn_leaves, dim3(blocks_for(P)), dim3(kThreads), 0, s, b.st, n, P, b.blo, b.bhi);
    hipLaunchKernelGGL(k_nn_obb, dim3(blocks_for(P)), dim3(kThreads), 0, s, b.st, n, P, b.ob);
    int64_t level = P >> 1;
    for (; level >= 2048; level >>= 1)
        hipLaunchKernelGGL(k_nn_level, dim3(blocks_for(level)), dim3(kThreads), 0, s, level, b.blo, b.bhi);
    if (level >= 1) hipLaunchKernelGGL(k_nn_inner, dim3(1), dim3(1024), 0, s, level, b.blo, b.bhi);
    BMF_HIP(hipGetLastError());
    return BMF_OK;
}

// Enqueues the nearest-target query of the tape's V final positions against
// a hierarchy of n targets, dL/dp, and the loss (FP64, fixed order) into
// loss_out (device).
int nn_query(bmf_tape* t, const NNBuffers& b, int64_t n, float* dldp, int32_t* nearest, double* loss_out,
             hipStream_t s) {
    const int64_t P = leaf_pow2(n);
    hipLaunchKernelGGL(k_nn_query, dim3((unsigned)(((int64_t)t->V * 64 + kNNThreads - 1) / kNNThreads)),
                       dim3(kNNThreads), 0, s, t->pos, t->V, b.st, n, P, b.blo, b.bhi, b.ob,
                       nearest ? nearest : t->nearest_tmp, dldp, t->d2);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, t->d2, t->V, loss_out);
    BMF_HIP(hipGetLastError());
    return BMF_OK;
}

// Runs k smoothing iterations starting from `a` (float4 [V]), ping-ponging
// between a and b, one launch per iteration; the state after iteration m is
// kept in checkpoint slot (from + m + 1) / every when that divides (from >= 0),
// and the final state is written as float3 to pos3.
int smooth_steps(bmf_tape* t, float4* a, float4* b, int k, hipStream_t s, float* pos3, int ckpt_from = -1) {
    float4* cur = a;
    float4* nxt = b;
    if (k == 0) {
        hipLaunchKernelGGL(k_tape_out, dim3(blocks_for(t->V)), dim3(kThreads), 0, s, a, t->V, pos3);
        BMF_HIP(hipGetLastError());
        return BMF_OK;
    }
    for (int m = 0; m < k; m++) {
        float4* ck = nullptr;
        if (ckpt_from >= 0 && (ckpt_from + m + 1) % t->ckpt_every == 0)
            ck = t->ckpt + (int64_t)((ckpt_from + m + 1) / t->ckpt_every) * t->cap_v;
        hipLaunchKernelGGL(k_tape_step, dim3(step_blocks(t->V)), dim3(kStepThreads), 0, s, t->ring, (int)t->cap_v,
                           t->deg, t->mv, cur, t->V, nxt, ck, m == k - 1 ? pos3 : nullptr);
        BMF_HIP(hipGetLastError());
        std::swap(cur, nxt);
    }
    return BMF_OK;
}

// ---- scratch of bmf_loss_nearest_noalloc ---------------------------------
// The hierarchy of up to BMF_NOALLOC_MAX_TARGETS targets lives in one
// file-scope device array, so that entry point allocates nothing: Morton
// keys and target ids (2 x 2 x 1 MiB), the sorted targets (4 MiB), the boxes
// of 2 x 4096 heap nodes of 64-target leaves (2 x 128 KiB), the oriented
// bounds of the 4096 leaves (256 KiB) and the radix sort's temporary storage
// (2 MiB), 10.5 MiB per GPU for the life of the process.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr int64_t kNAMax = BMF_NOALLOC_MAX_TARGETS;
constexpr int64_t kNALeaves = kNAMax / kLeaf;
static_assert(kNAMax % kLeaf == 0 && (kNALeaves & (kNALeaves - 1)) == 0, "leaves of the maximum set: a power of two");
constexpr size_t kNAKeyBytes = sizeof(uint32_t) * kNAMax;
constexpr size_t kNATgtBytes = sizeof(float4) * kNAMax;
constexpr size_t kNABoxBytes = sizeof(float4) * 2 * kNALeaves;
constexpr size_t kNAObbBytes = sizeof(float4) * 4 * kNALeaves;
constexpr size_t kNACubBytes = 2u << 20;
constexpr size_t kScratchBytes = 4 * kNAKeyBytes + kNATgtBytes + 2 * kNABoxBytes + kNAObbBytes + kNACubBytes;
static_assert(kScratchBytes == (42u << 18), "10.5 MiB of scratch with 64-target leaves");
static_assert(kNAKeyBytes % 256 == 0 && kNATgtBytes % 256 == 0 && kNABoxBytes % 256 == 0 && kNAObbBytes % 256 == 0,
              "256-byte aligned parts");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current), looked up
// on the first call for that device.
int scratch_acquire(int device) {
    if (device < 0 || device >= kMaxDevices) return BMF_EINVAL;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return BMF_OK;
    void* ptr = nullptr;
    BMF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return BMF_EHIP;
    g_scratch_dev[device] = ptr;
    return BMF_OK;
}

void* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_dev[device];
}

NNBuffers scratch_nn(void* base) {
    unsigned char* p = static_cast<unsigned char*>(base);
    NNBuffers b;
    b.keys[0] = reinterpret_cast<uint32_t*>(p);
    b.keys[1] = reinterpret_cast<uint32_t*>(p + kNAKeyBytes);
    b.vals[0] = reinterpret_cast<int32_t*>(p + 2 * kNAKeyBytes);
    b.vals[1] = reinterpret_cast<int32_t*>(p + 3 * kNAKeyBytes);
    p += 4 * kNAKeyBytes;
    b.st = reinterpret_cast<float4*>(p);
    p += kNATgtBytes;
    b.blo = reinterpret_cast<float4*>(p);
    b.bhi = reinterpret_cast<float4*>(p + kNABoxBytes);
    p += 2 * kNABoxBytes;
    b.ob = reinterpret_cast<float4*>(p);
    p += kNAObbBytes;
    b.cub_tmp = p;
    b.cub_bytes = kNACubBytes;
    return b;
}

__global__ void k_store_f64(double* __restrict__ out, double v) { out[0] = v; }

}  // namespace

extern "C" {

int bmf_tape_create(bmf_tape** out, bmf_ctx* ctx, int32_t max_dim, int64_t max_vertices, int64_t max_indices,
                    int32_t checkpoint_every) {
    if (!out || !ctx) return BMF_EINVAL;
    *out = nullptr;
    if (max_dim < 2 || max_dim > BMF_MAX_DIM || max_vertices < 1 || max_vertices > kMaxTapeVertices ||
        max_indices < 3 || checkpoint_every < 1)
        return BMF_EINVAL;
    BMF_HIP(hipSetDevice(ctx->device));
    bmf_tape* t = new (std::nothrow) bmf_tape();
    if (!t) return BMF_EHIP;
    t->ctx = ctx;
    t->max_dim = max_dim;
    t->cap_v = max_vertices;
    t->cap_i = max_indices;
    t->ckpt_every = checkpoint_every;
    const int64_t n3 = (int64_t)max_dim * max_dim * max_dim;
    const bool ok = dalloc(&t->dens, n3) == hipSuccess && dalloc(&t->vid0, n3) == hipSuccess &&
                    dalloc(&t->vedge, max_vertices) == hipSuccess && dalloc(&t->boundary, max_vertices) == hipSuccess &&
                    dalloc(&t->idx, max_indices) == hipSuccess &&
                    dalloc(&t->ring, (int64_t)kMaxRing * max_vertices) == hipSuccess &&
                    dalloc(&t->bscale, max_vertices) == hipSuccess &&
                    dalloc(&t->deg, max_vertices) == hipSuccess && dalloc(&t->mv, max_vertices) == hipSuccess &&
                    dalloc(&t->p4, max_vertices) == hipSuccess && dalloc(&t->q4, max_vertices) == hipSuccess &&
                    dalloc(&t->pos, 3 * max_vertices) == hipSuccess &&
                    dalloc(&t->g, max_vertices) == hipSuccess && dalloc(&t->g2, max_vertices) == hipSuccess &&
                    dalloc(&t->r4, max_vertices) == hipSuccess && dalloc(&t->err, 1) == hipSuccess &&
                    dalloc(&t->bb, 8) == hipSuccess && dalloc(&t->d2, max_vertices) == hipSuccess &&
                    dalloc(&t->loss, 1) == hipSuccess && dalloc(&t->nearest_tmp, max_vertices) == hipSuccess;
    bool evs = ok;
    for (int i = 0; i < 6 && evs; i++) evs = hipEventCreate(&t->ev[i]) == hipSuccess;
    if (!ok || !evs) {
        bmf_tape_destroy(t);
        return BMF_EHIP;
    }
    // set-up of bmf_loss_nearest_noalloc: the scratch address on this device
    // (the device is current); a failure here only disables that entry point
    (void)scratch_acquire(ctx->device);
    *out = t;
    return BMF_OK;
}

void bmf_tape_destroy(bmf_tape* t) {
    if (!t) return;
    (void)hipSetDevice(t->ctx->device);
    (void)hipStreamSynchronize(t->ctx->stream);
    dfree(t->dens);
    dfree(t->vid0);
    dfree(t->vedge);
    dfree(t->boundary);
    dfree(t->idx);
    dfree(t->ring);
    dfree(t->bscale);
    dfree(t->deg);
    dfree(t->mv);
    dfree(t->p4);
    dfree(t->q4);
    dfree(t->pos);
    dfree(t->g);
    dfree(t->g2);
    dfree(t->r4);
    for (int i = 0; i < 6; i++)
        if (t->ev[i]) (void)hipEventDestroy(t->ev[i]);
    dfree(t->err);
    dfree(t->ckpt);
    dfree(t->bb);
    dfree(t->d2);
    dfree(t->loss);
    dfree(t->nearest_tmp);
    free_nn(t);
    delete t;
}

int bmf_forward_diff(bmf_ctx* c, const bmf_chunk_desc* d, const float* densities, bmf_tape* t) {
    if (!c || !d || !t || t->ctx != c) return BMF_EINVAL;
    int rc = check_desc(*d);
    if (rc) return rc;
    if (d->dim > t->max_dim) return BMF_ECAPACITY;
    if (!densities && !c->sampler_set) return BMF_ESTATE;
    const int dim = d->dim;
    const int64_t n3 = (int64_t)dim * dim * dim;
    if (n3 > c->A.cap_samples) return BMF_ECAPACITY;
    BMF_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    BMF_HIP(hipStreamSynchronize(s));
    t->valid = false;
    t->timed[0] = false;
    BMF_HIP(hipEventRecord(t->ev[0], s));
    c->have_extract = false;
    c->results_valid = false;
    // the sampler of this call: the context's, or the caller's densities; the
    // context's own is put back on every way out of this call, so later
    // extractions never read the caller's buffer
    struct RestoreSampler {
        bmf_ctx* c;
        bool active = false;
        ~RestoreSampler() {
            if (active) (void)hipMemcpy(c->A.sampler, &c->h_sampler, sizeof(SamplerDev), hipMemcpyHostToDevice);
        }
    } restore{c};
    if (densities) {
        SamplerDev ext{};
        ext.kind = BMF_SAMPLER_EXTERNAL;
        ext.ext = densities;
        restore.active = true;
        BMF_HIP(hipMemcpy(c->A.sampler, &ext, sizeof(SamplerDev), hipMemcpyHostToDevice));
    }
    rc = upload_descs(c, d, 1);
    if (rc) return rc;
    Arenas A = c->A;
    A.ext_base = 0;
    A.mesh_mode = BMF_MESH_TRIANGLES;  // the adjoint is defined on the triangle path
    A.quad_tris = 0;
    A.smoothing_follows = 0;  // the tape smooths its own copy: every output position is written by emit
    A.normals = 0;
    const int wcls = row_word_class(dim);
    // extraction with every octave evaluated (densities exact everywhere),
    // no smoothing: the tape smooths its own copy
    hipError_t e = launch_plan(A, 1, 0, 1, s);
    if (e == hipSuccess)
        e = launch_sample(A, 1, (int)std::min<int64_t>((n3 + A.sample_tile - 1) / A.sample_tile, c->persistent_blocks),
                          0, densities ? (int)BMF_SAMPLER_EXTERNAL : (int)c->h_sampler.kind,
                          !densities && c->h_sampler.noise_dims == 2, s);
    if (e == hipSuccess) e = launch_count(A, 1, dim * kMaxPlaneSplit, wcls, s);
    if (e == hipSuccess) e = launch_refine(A, 1, dim * kMaxPlaneSplit, s);
    if (e == hipSuccess) e = launch_scan(A, 1, 1, s);
    if (e == hipSuccess) e = launch_emit(A, 1, dim * kMaxPlaneSplit, wcls, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->h_results, A.results, sizeof(bmf_chunk_result), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->h_ctr, A.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return BMF_EHIP;
    const bmf_chunk_result r = c->h_results[0];
    if (c->h_ctr->error) return BMF_EHIP;
    if (r.status != BMF_OK) return r.status;
    if (r.vertex_count > t->cap_v || r.index_count > t->cap_i) return BMF_ECAPACITY;
    t->dim = dim;
    t->K = d->iters;
    t->bp = d->boundary_processing;
    t->V = r.vertex_count;
    t->T = r.index_count / 3;
    t->contains_mesh = r.contains_mesh;
    const int64_t slots = t->K / t->ckpt_every + 1;
    if (slots > t->ckpt_slots) {
        dfree(t->ckpt);
        t->ckpt_slots = 0;  // a failed reallocation leaves no slots, so the next forward retries it
        BMF_HIP(dalloc(&t->ckpt, slots * t->cap_v));
        t->ckpt_slots = slots;
    }
    const int V = t->V, T = t->T;
    BMF_HIP(hipMemsetAsync(t->err, 0, sizeof(int), s));
    const bool smooth = t->K > 0 && T > 0;  // per-cell triangle bases and masks exist for chunks with K > 0
    {
        const int64_t most = std::max<int64_t>(n3, std::max<int64_t>(3 * (int64_t)T, V));
        const unsigned blocks = (unsigned)std::min<int64_t>((most + kThreads - 1) / kThreads, 4096);
        hipLaunchKernelGGL(k_tape_take, dim3(blocks), dim3(kThreads), 0, s, A.dens, A.vedge + r.vertex_offset,
                           A.boundary + r.vertex_offset, A.idx + r.index_offset, A.pos + 3 * r.vertex_offset, n3, V,
                           3 * (int64_t)T, dim, t->dens, t->vedge, t->boundary, t->idx, t->p4, t->ckpt,
                           smooth ? nullptr : t->pos, t->vid0);
        BMF_HIP(hipGetLastError());
    }
    if (V > 0) {
        if (smooth) {
            hipLaunchKernelGGL(k_tape_adj, dim3(step_blocks(V)), dim3(kStepThreads), 0, s, t->vedge, t->boundary,
                               A.cell_tb, A.cell_mask, t->idx, dim, V, T, t->bp, (int)t->cap_v, t->ring, t->deg, t->mv,
                               t->bscale, t->err);
            BMF_HIP(hipGetLastError());
            rc = smooth_steps(t, t->p4, t->q4, t->K, s, t->pos, 0);
            if (rc) return rc;
        } else {
            BMF_HIP(hipMemsetAsync(t->deg, 0, V, s));
            BMF_HIP(hipMemsetAsync(t->mv, 0, V, s));
            for (int64_t q = 1; q < slots; q++)
                BMF_HIP(hipMemcpyAsync(t->ckpt + q * t->cap_v, t->p4, sizeof(float4) * V, hipMemcpyDeviceToDevice, s));
        }
    }
    BMF_HIP(hipEventRecord(t->ev[1], s));
    int herr = 0;
    BMF_HIP(hipMemcpyAsync(&herr, t->err, sizeof(int), hipMemcpyDeviceToHost, s));
    BMF_HIP(hipStreamSynchronize(s));
    if (V > 0 && herr) return BMF_EHIP;
    t->valid = true;
    t->timed[0] = true;
    return BMF_OK;
}

int bmf_tape_get(bmf_tape* t, bmf_tape_info* info) {
    if (!t || !info) return BMF_EINVAL;
    if (!t->valid) return BMF_ESTATE;
    info->dim = t->dim;
    info->iters = t->K;
    info->checkpoint_every = t->ckpt_every;
    info->vertex_count = t->V;
    info->index_count = 3 * t->T;
    info->contains_mesh = t->contains_mesh;
    info->positions = t->pos;
    info->indices = t->idx;
    info->boundary = t->boundary;
    info->densities = t->dens;
    return BMF_OK;
}

int bmf_tape_positions(bmf_tape* t, int32_t iteration, float* out) {
    if (!t || !out) return BMF_EINVAL;
    if (!t->valid) return BMF_ESTATE;
    if (iteration < 0 || iteration > t->K) return BMF_EINVAL;
    if (t->V == 0) return BMF_OK;
    hipStream_t s = t->ctx->stream;
    BMF_HIP(hipSetDevice(t->ctx->device));
    const int64_t slot = iteration / t->ckpt_every;
    BMF_HIP(hipMemcpyAsync(t->q4, t->ckpt + slot * t->cap_v, sizeof(float4) * t->V, hipMemcpyDeviceToDevice, s));
    const int rest = t->T > 0 ? iteration - (int)(slot * t->ckpt_every) : 0;
    int rc = smooth_steps(t, t->q4, t->r4, rest, s, out);
    if (rc) return rc;
    BMF_HIP(hipStreamSynchronize(s));
    return BMF_OK;
}

int bmf_loss_nearest(bmf_tape* t, const float* targets, int64_t n, float* dldp, int32_t* nearest, float* loss) {
    if (!t || !targets || n < 1 || n > (int64_t)0x7FFFFFFF || !dldp) return BMF_EINVAL;
    if (!t->valid) return BMF_ESTATE;
    BMF_HIP(hipSetDevice(t->ctx->device));
    hipStream_t s = t->ctx->stream;
    int rc = reserve_nn(t, n);
    if (rc) return rc;
    t->nn_n = 0;
    t->timed[1] = false;
    BMF_HIP(hipEventRecord(t->ev[2], s));
    rc = nn_build(targets, n, t->bb, tape_nn(t), s);
    if (rc) return rc;
    t->nn_n = n;
    if (t->V == 0) {
        BMF_HIP(hipStreamSynchronize(s));
        if (loss) *loss = 0.0f;
        return BMF_OK;
    }
    rc = nn_query(t, tape_nn(t), n, dldp, nearest, t->loss, s);
    if (rc) return rc;
    BMF_HIP(hipEventRecord(t->ev[3], s));
    double l = 0.0;
    BMF_HIP(hipMemcpyAsync(&l, t->loss, sizeof(double), hipMemcpyDeviceToHost, s));
    BMF_HIP(hipStreamSynchronize(s));
    if (loss) *loss = (float)l;
    t->timed[1] = true;
    return BMF_OK;
}

int bmf_tape_set_targets(bmf_tape* t, const float* targets, int64_t n) {
    if (!t || !targets || n < 1 || n > (int64_t)0x7FFFFFFF) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(t->ctx->device));
    hipStream_t s = t->ctx->stream;
    int rc = reserve_nn(t, n);
    if (rc) return rc;
    t->nn_n = 0;
    rc = nn_build(targets, n, t->bb, tape_nn(t), s);
    if (rc) return rc;
    BMF_HIP(hipStreamSynchronize(s));
    t->nn_n = n;
    return BMF_OK;
}

int bmf_tape_loss_nearest(bmf_tape* t, float* dldp, int32_t* nearest, float* loss) {
    if (!t || !dldp) return BMF_EINVAL;
    if (!t->valid || t->nn_n < 1) return BMF_ESTATE;
    BMF_HIP(hipSetDevice(t->ctx->device));
    hipStream_t s = t->ctx->stream;
    if (t->V == 0) {
        if (loss) *loss = 0.0f;
        return BMF_OK;
    }
    t->timed[1] = false;
    BMF_HIP(hipEventRecord(t->ev[2], s));
    int rc = nn_query(t, tape_nn(t), t->nn_n, dldp, nearest, t->loss, s);
    if (rc) return rc;
    BMF_HIP(hipEventRecord(t->ev[3], s));
    double l = 0.0;
    BMF_HIP(hipMemcpyAsync(&l, t->loss, sizeof(double), hipMemcpyDeviceToHost, s));
    BMF_HIP(hipStreamSynchronize(s));
    if (loss) *loss = (float)l;
    t->timed[1] = true;
    return BMF_OK;
}

int bmf_loss_nearest_noalloc(bmf_tape* t, const float* targets, int64_t n, float* dldp, int32_t* nearest,
                             double* loss) {
    if (!t || !targets || n < 1 || !dldp) return BMF_EINVAL;
    if (n > BMF_NOALLOC_MAX_TARGETS) return BMF_ECAPACITY;
    if (!t->valid) return BMF_ESTATE;
    void* base = scratch_on(t->ctx->device);
    if (!base) return BMF_EHIP;  // the address lookup at bmf_tape_create failed
    BMF_HIP(hipSetDevice(t->ctx->device));
    hipStream_t s = t->ctx->stream;
    const NNBuffers b = scratch_nn(base);
    // the radix sort's temporary storage for n pairs must fit its share (a
    // size query: nothing is launched or allocated)
    size_t need = 0;
    {
        hipcub::DoubleBuffer<uint32_t> kb(b.keys[0], b.keys[1]);
        hipcub::DoubleBuffer<int32_t> vb(b.vals[0], b.vals[1]);
        BMF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kb, vb, (int)n, 0, 30, s));
    }
    if (need > b.cub_bytes) return BMF_ECAPACITY;
    int rc = nn_build(targets, n, t->bb, b, s);
    if (rc) return rc;
    if (t->V == 0) {
        if (loss) hipLaunchKernelGGL(k_store_f64, dim3(1), dim3(1), 0, s, loss, 0.0);
        BMF_HIP(hipGetLastError());
        return BMF_OK;
    }
    return nn_query(t, b, n, dldp, nearest, loss ? loss : t->loss, s);
}

int bmf_backward_diff(bmf_tape* t, const float* dldp, float* dldd) {
    if (!t || !dldd || (!dldp && t->V > 0)) return BMF_EINVAL;
    if (!t->valid) return BMF_ESTATE;
    BMF_HIP(hipSetDevice(t->ctx->device));
    hipStream_t s = t->ctx->stream;
    const int64_t n3 = (int64_t)t->dim * t->dim * t->dim;
    if (t->V == 0) {
        BMF_HIP(hipMemsetAsync(dldd, 0, sizeof(float) * n3, s));
        BMF_HIP(hipStreamSynchronize(s));
        return BMF_OK;
    }
    t->timed[2] = false;
    BMF_HIP(hipEventRecord(t->ev[4], s));
    const bool smooth = t->K > 0 && t->T > 0;
    hipLaunchKernelGGL(k_bwd_init, dim3(blocks_for(t->V)), dim3(kThreads), 0, s, dldp, t->V, t->mv, t->bscale,
                       (int)smooth, t->g);
    float4* cur = t->g;
    float4* nxt = t->g2;
    if (smooth)
        for (int m = t->K - 1; m >= 0; m--) {
            hipLaunchKernelGGL(k_bwd_step, dim3(step_blocks(t->V)), dim3(kStepThreads), 0, s, t->ring, (int)t->cap_v,
                               t->deg, t->mv, t->bscale, cur, t->V, (int)(m == 0), nxt);
            std::swap(cur, nxt);
        }
    hipLaunchKernelGGL(k_bwd_place, dim3(blocks_for(n3)), dim3(kThreads), 0, s, t->dens, t->vid0, cur, t->dim, dldd);
    BMF_HIP(hipGetLastError());
    BMF_HIP(hipEventRecord(t->ev[5], s));
    BMF_HIP(hipStreamSynchronize(s));
    t->timed[2] = true;
    return BMF_OK;
}

int bmf_tape_timing(bmf_tape* t, float* ms) {
    if (!t || !ms) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(t->ctx->device));
    for (int k = 0; k < 3; k++) {
        ms[k] = -1.0f;
        if (t->timed[k]) BMF_HIP(hipEventElapsedTime(&ms[k], t->ev[2 * k], t->ev[2 * k + 1]));
    }
    return BMF_OK;
}

}  // extern "C"
