// Output formatting of a batch (C ABI bmf_format in include/bmf.h): the
// meshes of the last extraction written as VBO-ready arrays in device memory,
// the work of WorldOctreeNode::format (WorldOctreeNode.cpp:68-86) ->
// GLChunk::format_data (GLChunk.cpp:278-297, unwind_verts = false) and of
// DMCChunk::copy_verts_and_inds (DMCChunk.cpp:578-591), without the 84-byte
// DualVertex round trip through host memory.
//
//   positions  float [V][3]: grid units, or world units p * scale +
//              overlap_pos (the vertex shader's transform,
//              shaders/main_vs.glsl:34), one FP32 multiply then one add
//   colours    float [V][3]: white (DualVertex::color, DMCChunk.cpp:681) or
//              ColorMapper::generate_colors (ColorMapper.cpp:15-62): a
//              4-octave FBM n at the vertex's grid position (scale 1) mapped
//              by hsl_to_rgb((4n + 1) * 0.5 * 360, 0.72, 1) (:60-119, an HSV
//              conversion); the FBM is the project's simplex FBM with
//              frequency 0.01, lacunarity 2, gain 0.5, seed 1337
//   normals    float [V][3]: the unit normals of the extraction
//   indices    u32: chunk-local, or plus the chunk's first vertex in the
//              formatted batch
// Chunks are taken in batch order when they contain a mesh (contains_mesh,
// the condition of WorldOctreeNode::format and copy_verts_and_inds) and were
// extracted completely.
//
// Two launches: k_format_plan (one workgroup) forms the chunks' offsets in
// the formatted arrays and their 256-element tile counts from the device
// results; k_format (persistent, grid-stride over the tiles) writes the
// vertices and indices.  HBM traffic: 12 B read + 12-36 B written per
// vertex, 4 B read + 4 B written per index.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/bmf.h"
#include "bmf_context.h"
#include "bmf_internal.h"
#include "samplers.h"

namespace bmf {
namespace {

constexpr int kFmtThreads = 256;   // elements per tile, one per thread
constexpr int kPlanThreads = 1024;
constexpr int kColorOctaves = 4;

#define BMF_HIP(call)                          \
    do {                                       \
        hipError_t e_ = (call);                \
        if (e_ != hipSuccess) return BMF_EHIP; \
    } while (0)

struct FmtParams {
    int32_t space;
    int32_t colors;
    int32_t global;
    int32_t normals;
    float fo[kColorOctaves];
    float ao[kColorOctaves];
    float inv_norm;
    uint32_t seed;
    float* pos;
    float* col;
    float* nrm;
    uint32_t* idx;
    int64_t cap_v;
    int64_t cap_i;
};

// ColorMapper::hsl_to_rgb (ColorMapper.cpp:60-119) for s > 0.
__device__ __forceinline__ float3 hsv_to_rgb(float h, float s, float v) {
    float hh = fmodf(fabsf(h), 360.0f);
    hh = hh / 60.0f;
    const int i = (int)hh;
    const float ff = hh - (float)i;
    const float p = v * (1.0f - s);
    const float q = v * (1.0f - (s * ff));
    const float t = v * (1.0f - (s * (1.0f - ff)));
    switch (i) {
        case 0: return make_float3(v, t, p);
        case 1: return make_float3(q, v, p);
        case 2: return make_float3(p, v, t);
        case 3: return make_float3(p, q, v);
        case 4: return make_float3(t, p, v);
        default: return make_float3(v, p, q);
    }
}

// Per chunk j of the batch (table rows of n + 1 entries, the last holding the
// totals): 0 first formatted vertex, 1 first formatted index, 2 first vertex
// tile, 3 first index tile.  Thread t walks a contiguous segment of chunks,
// the segment totals are scanned across the workgroup in LDS.
__global__ void __launch_bounds__(kPlanThreads) k_format_plan(const bmf_chunk_result* __restrict__ R, int n,
                                                               int64_t* __restrict__ tab) {
    __shared__ int64_t s[4][kPlanThreads];
    const int t = threadIdx.x;
    const int seg = (n + kPlanThreads - 1) / kPlanThreads;
    const int j0 = min(n, t * seg), j1 = min(n, j0 + seg);
    int64_t a[4] = {0, 0, 0, 0};
    for (int j = j0; j < j1; j++) {
        const bmf_chunk_result r = R[j];
        const bool take = r.contains_mesh && r.status == BMF_OK;
        const int64_t V = take ? r.vertex_count : 0, I = take ? r.index_count : 0;
        a[0] += V;
        a[1] += I;
        a[2] += (V + kFmtThreads - 1) / kFmtThreads;
        a[3] += (I + kFmtThreads - 1) / kFmtThreads;
    }
    for (int k = 0; k < 4; k++) s[k][t] = a[k];
    __syncthreads();
    for (int off = 1; off < kPlanThreads; off <<= 1) {  // inclusive scan (Hillis-Steele)
        int64_t v[4];
        for (int k = 0; k < 4; k++) v[k] = t >= off ? s[k][t - off] : 0;
        __syncthreads();
        for (int k = 0; k < 4; k++) s[k][t] += v[k];
        __syncthreads();
    }
    int64_t run[4];
    for (int k = 0; k < 4; k++) run[k] = s[k][t] - a[k];  // exclusive prefix of the segment
    for (int j = j0; j < j1; j++) {
        const bmf_chunk_result r = R[j];
        const bool take = r.contains_mesh && r.status == BMF_OK;
        const int64_t V = take ? r.vertex_count : 0, I = take ? r.index_count : 0;
        for (int k = 0; k < 4; k++) tab[(int64_t)k * (n + 1) + j] = run[k];
        run[0] += V;
        run[1] += I;
        run[2] += (V + kFmtThreads - 1) / kFmtThreads;
        run[3] += (I + kFmtThreads - 1) / kFmtThreads;
    }
    if (t == kPlanThreads - 1)
        for (int k = 0; k < 4; k++) tab[(int64_t)k * (n + 1) + n] = s[k][t];
}

// The chunk j of a tile: the largest j with prefix[j] <= tile (prefix
// non-decreasing, prefix[n] > tile); that chunk has a tile.
__device__ __forceinline__ int tile_chunk(const int64_t* __restrict__ prefix, int n, int64_t tile) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kFmtThreads) k_format(const bmf_chunk_result* __restrict__ R, int n,
                                                        const int64_t* __restrict__ tab, const float* __restrict__ pos,
                                                        const uint32_t* __restrict__ idx,
                                                        const float* __restrict__ nrm, FmtParams P) {
    const int64_t* dv = tab;
    const int64_t* di = tab + (n + 1);
    const int64_t* vt = tab + 2 * (int64_t)(n + 1);
    const int64_t* it = tab + 3 * (int64_t)(n + 1);
    if (dv[n] > P.cap_v || di[n] > P.cap_i) return;  // reported by the host: nothing is written
    const int64_t VT = vt[n], IT = it[n];
    for (int64_t tile = blockIdx.x; tile < VT + IT; tile += gridDim.x) {
        if (tile < VT) {
            const int j = tile_chunk(vt, n, tile);
            const bmf_chunk_result r = R[j];
            const int64_t v = (tile - vt[j]) * kFmtThreads + threadIdx.x;
            if (v >= r.vertex_count) continue;
            const int64_t src = r.vertex_offset + v, dst = dv[j] + v;
            const float px = pos[3 * src], py = pos[3 * src + 1], pz = pos[3 * src + 2];
            float* O = P.pos + 3 * dst;
            if (P.space == BMF_SPACE_WORLD) {
                const float mx = px * r.delta, my = py * r.delta, mz = pz * r.delta;
                O[0] = mx + r.overlap_pos[0];
                O[1] = my + r.overlap_pos[1];
                O[2] = mz + r.overlap_pos[2];
            } else {
                O[0] = px;
                O[1] = py;
                O[2] = pz;
            }
            if (P.colors == BMF_COLOR_WHITE) {
                float* C = P.col + 3 * dst;
                C[0] = 1.0f;
                C[1] = 1.0f;
                C[2] = 1.0f;
            } else if (P.colors == BMF_COLOR_NOISE) {
                const float nz = fbm_value(P.fo, P.ao, P.inv_norm, kColorOctaves, P.seed, px, py, pz) * 4.0f;
                const float3 c = hsv_to_rgb(((nz + 1.0f) * 0.5f) * 360.0f, 0.72f, 1.0f);
                float* C = P.col + 3 * dst;
                C[0] = c.x;
                C[1] = c.y;
                C[2] = c.z;
            }
            if (P.normals) {
                float* N = P.nrm + 3 * dst;
                N[0] = nrm[3 * src];
                N[1] = nrm[3 * src + 1];
                N[2] = nrm[3 * src + 2];
            }
        } else {
            const int64_t tt = tile - VT;
            const int j = tile_chunk(it, n, tt);
            const bmf_chunk_result r = R[j];
            const int64_t i = (tt - it[j]) * kFmtThreads + threadIdx.x;
            if (i >= r.index_count) continue;
            P.idx[di[j] + i] = idx[r.index_offset + i] + (P.global ? (uint32_t)dv[j] : 0u);
        }
    }
}

}  // namespace
}  // namespace bmf

using namespace bmf;

extern "C" {

int bmf_format(bmf_ctx* c, const bmf_format_options* o, const bmf_vbo* out, int64_t* vertices, int64_t* indices) {
    if (!c || !o || !out || !out->positions || !out->indices) return BMF_EINVAL;
    if (o->space != BMF_SPACE_GRID && o->space != BMF_SPACE_WORLD) return BMF_EINVAL;
    if (o->colors != BMF_COLOR_NONE && o->colors != BMF_COLOR_WHITE && o->colors != BMF_COLOR_NOISE)
        return BMF_EINVAL;
    if (o->global_indices != 0 && o->global_indices != 1) return BMF_EINVAL;
    if (o->colors != BMF_COLOR_NONE && !out->colors) return BMF_EINVAL;
    if (out->cap_vertices < 0 || out->cap_indices < 0) return BMF_EINVAL;
    if (!c->have_extract) return BMF_ESTATE;
    if (out->normals && !c->A.normals) return BMF_ESTATE;
    BMF_HIP(hipSetDevice(c->device));
    const int n = c->last_n;
    if (!c->d_fmt) {
        BMF_HIP(hipMalloc((void**)&c->d_fmt, sizeof(int64_t) * 4 * ((size_t)c->lim.max_chunks + 1)));
    }
    if (!c->h_fmt) {
        BMF_HIP(hipHostMalloc((void**)&c->h_fmt, sizeof(int64_t) * 2, hipHostMallocDefault));
    }
    FmtParams P{};
    P.space = o->space;
    P.colors = o->colors;
    P.global = o->global_indices;
    P.normals = out->normals != nullptr;
    // FBM tables of the colour noise: the same FP32 product chains as the
    // terrain sampler's (frequency f0 * L^o, amplitude g^o)
    float f = 0.01f, a = 1.0f, norm = 0.0f;
    for (int k = 0; k < kColorOctaves; k++) {
        P.fo[k] = f;
        P.ao[k] = a;
        norm = norm + a;
        f = f * 2.0f;
        a = a * 0.5f;
    }
    P.inv_norm = 1.0f / norm;
    P.seed = 1337u;
    P.pos = out->positions;
    P.col = out->colors;
    P.nrm = out->normals;
    P.idx = out->indices;
    P.cap_v = out->cap_vertices;
    P.cap_i = out->cap_indices;
    hipStream_t s = c->stream;
    const int nn = std::max(n, 0);
    hipLaunchKernelGGL(k_format_plan, dim3(1), dim3(kPlanThreads), 0, s, c->A.results, nn, c->d_fmt);
    hipLaunchKernelGGL(k_format, dim3((unsigned)std::max(1, c->persistent_blocks)), dim3(kFmtThreads), 0, s,
                       c->A.results, nn, c->d_fmt, c->A.pos, c->A.idx, c->A.nrm, P);
    BMF_HIP(hipGetLastError());
    BMF_HIP(hipMemcpyAsync(&c->h_fmt[0], c->d_fmt + nn, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    BMF_HIP(hipMemcpyAsync(&c->h_fmt[1], c->d_fmt + (nn + 1) + nn, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    BMF_HIP(hipStreamSynchronize(s));
    if (vertices) *vertices = c->h_fmt[0];
    if (indices) *indices = c->h_fmt[1];
    if (c->h_fmt[0] > out->cap_vertices || c->h_fmt[1] > out->cap_indices) return BMF_ECAPACITY;
    return BMF_OK;
}

}  // extern "C"
