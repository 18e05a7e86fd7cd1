// Device-side work-queue producer (C ABI bmf_lod_queue / bmf_extract_queued
// in include/bmf.h): the LOD pass of the watcher thread run on the GPU,
// writing the chunk descriptors of the next extraction straight into the
// context's descriptor arena.
//
// Reference: WorldWatcher::check_leaves (WorldWatcher.cpp:155-173) walks the
// renderables list in order while counter < max_gen; a node that needs a
// split (WorldOctree::node_needs_split, WorldOctree.cpp:239-249) adds 8 to
// the counter and is pushed by handle_split_check (:175-188) when it is a
// drawn world leaf not already split, grouped or superseded; otherwise a
// world leaf whose parent needs grouping (node_needs_group, :251-261) has the
// parent pushed by handle_group_check (:189-212) when the parent is a
// superseded inner node, not already grouped, whose 8 children are world
// leaves -- the first such child in list order pushes it.  process_batch
// (:214-238) then turns each pushed node into chunks by its flags: DESTROY
// none, SPLIT its 8 children (WorldOctree::split_node, WorldOctree.cpp:175-210:
// child i at pos + (MCDX, MCDY, MCDZ)[i] * size/2, level + 1, morton
// (code << 3) | (dx | dy << 1 | dz << 2)), GROUP the node itself
// (group_node_1); ChunkGenerator::process_queue gives each chunk dim =
// chunk_resolution (WorldOctree.cpp:263-269) and extract_chunk's overlap
// rule (ChunkGenerator.cpp:98).
//
// One workgroup walks the list in tiles of 1024 nodes in order (pass A: split
// tests, the counter cut-off by a scan of the split tests, group candidates
// marking their parent with the lowest list position by atomicMin; pass B:
// the pushed nodes, their flags, and the dirty list and descriptors placed
// by scans).  Latency bound; the list is the watcher's renderables (hundreds
// to tens of thousands of nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "../../include/bmf.h"
#include "bmf_context.h"
#include "bmf_internal.h"

namespace bmf {
namespace {

constexpr int kLodThreads = 1024;
constexpr uint32_t kDraw = 1, kDestroy = 2, kSplit = 4, kGroup = 8, kDrawChildren = 64, kSuperceded = 256;
constexpr uint32_t kMCDX[8] = {0, 1, 1, 0, 0, 1, 1, 0};  // Tables.hpp:7-9
constexpr uint32_t kMCDY[8] = {0, 0, 0, 0, 1, 1, 1, 1};
constexpr uint32_t kMCDZ[8] = {0, 0, 1, 1, 0, 0, 1, 1};

// pass-A code of each list position
enum : uint8_t { kNotVisited = 0, kVisited = 1, kSplitPush = 2, kGroupCandidate = 3 };

#define BMF_HIP(call)                          \
    do {                                       \
        hipError_t e_ = (call);                \
        if (e_ != hipSuccess) return BMF_EHIP; \
    } while (0)

// |middle - focus| with middle = pos + size * 0.5 (WorldOctreeNode.cpp:27),
// glm::distance's (x*x + y*y) + z*z under the square root.
__device__ __forceinline__ float focus_distance(const bmf_world_node& n, const bmf_lod_params& P) {
    const float h = n.size * 0.5f;
    const float mx = n.pos[0] + h, my = n.pos[1] + h, mz = n.pos[2] + h;
    const float dx = P.focus[0] - mx, dy = P.focus[1] - my, dz = P.focus[2] - mz;
    return sqrtf((dx * dx + dy * dy) + dz * dz);
}

__device__ __forceinline__ bool needs_split(const bmf_world_node& n, const bmf_lod_params& P) {
    if (n.level >= P.max_level) return false;
    if (n.level < P.min_level) return true;
    const float lim = (n.size * P.split_multiplier + P.size_modifier) + n.size * 0.5f;
    return focus_distance(n, P) < lim;
}

__device__ __forceinline__ bool needs_group(const bmf_world_node& n, const bmf_lod_params& P) {
    if (n.level < P.min_level) return false;
    if (n.level > P.max_level) return true;
    const float lim = (n.size * P.group_multiplier + P.size_modifier) + n.size * 0.5f;
    return focus_distance(n, P) > lim;
}

// Exclusive scan over the workgroup (wave prefix by shuffles, then the
// waves' totals); `total` receives the sum.
__device__ __forceinline__ int block_scan(int v, int* s_w, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < kLodThreads / 64; w++) {
            const int t = s_w[w];
            s_w[w] = run;
            run += t;
        }
        s_w[kLodThreads / 64] = run;
    }
    __syncthreads();
    const int out = s_w[wave] + x - v;
    total = s_w[kLodThreads / 64];
    __syncthreads();
    return out;
}

__device__ __forceinline__ bool can_group(const bmf_world_node* __restrict__ nodes, int n_nodes, int p) {
    const bmf_world_node& q = nodes[p];
    if (q.world_leaf || (q.flags & kGroup) || !(q.flags & kSuperceded)) return false;
    for (int i = 0; i < 8; i++) {
        const int ch = q.children[i];
        if (ch < 0 || ch >= n_nodes || !nodes[ch].world_leaf) return false;
    }
    return true;
}

__global__ void __launch_bounds__(kLodThreads) k_lod_queue(const bmf_world_node* __restrict__ nodes, int n_nodes,
                                                          const int32_t* __restrict__ rlist, int n_r, bmf_lod_params P,
                                                          uint32_t* __restrict__ flags_out, int* __restrict__ first_r,
                                                          uint8_t* __restrict__ code, int32_t* __restrict__ dirty,
                                                          bmf_lod_item* __restrict__ items,
                                                          bmf_chunk_desc* __restrict__ descs, int cap,
                                                          int* __restrict__ counts) {
    __shared__ int s_w[kLodThreads / 64 + 1];
    const int t = threadIdx.x;
    for (int i = t; i < n_nodes; i += kLodThreads) {
        flags_out[i] = nodes[i].flags;
        first_r[i] = INT_MAX;
    }
    __syncthreads();
    // ---- pass A: the counter cut-off, split pushes, group candidates ----
    int splits_before = 0;
    for (int base = 0; base < n_r; base += kLodThreads) {
        const int r = base + t;
        const int id = r < n_r ? rlist[r] : -1;
        const bool ok = id >= 0 && id < n_nodes;
        const bmf_world_node* n = ok ? &nodes[id] : nullptr;
        const bool ns = ok && needs_split(*n, P);
        int tile_splits = 0;
        const int before = splits_before + block_scan(ns ? 1 : 0, s_w, tile_splits);
        splits_before += tile_splits;
        uint8_t c = kNotVisited;
        if (r < n_r && 8ll * before < (long long)P.max_gen) {  // the loop test `counter < max_gen`
            c = kVisited;
            if (ns) {
                const uint32_t f = n->flags;
                if (n->world_leaf && !(f & kSplit) && !(f & kDrawChildren) && (f & kDraw) && !(f & kGroup) &&
                    !(f & kSuperceded))
                    c = kSplitPush;
            } else if (ok && n->world_leaf && n->parent >= 0 && n->parent < n_nodes &&
                       needs_group(nodes[n->parent], P)) {
                c = kGroupCandidate;
                atomicMin(&first_r[n->parent], r);
            }
        }
        if (r < n_r) code[r] = c;
    }
    __threadfence();
    __syncthreads();
    // ---- pass B: pushed nodes in list order, their flags and chunks ----
    int dirty_before = 0, chunks_before = 0;
    for (int base = 0; base < n_r; base += kLodThreads) {
        const int r = base + t;
        const uint8_t c = r < n_r ? code[r] : kNotVisited;
        int pushed = -1;
        uint32_t fin = 0;
        if (c == kSplitPush) {
            pushed = rlist[r];
            fin = nodes[pushed].flags | kSplit;
            flags_out[pushed] = fin;
        } else if (c == kGroupCandidate) {
            const int p = nodes[rlist[r]].parent;
            if (__hip_atomic_load(&first_r[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == r &&
                can_group(nodes, n_nodes, p)) {
                pushed = p;
                fin = nodes[p].flags | kGroup;
                flags_out[p] = fin;
            }
        }
        const int nd = pushed < 0 || (fin & kDestroy) ? 0 : ((fin & kSplit) ? 8 : ((fin & kGroup) ? 1 : 0));
        int tile_dirty = 0, tile_chunks = 0;
        const int di = dirty_before + block_scan(pushed >= 0 ? 1 : 0, s_w, tile_dirty);
        const int ci = chunks_before + block_scan(nd, s_w, tile_chunks);
        dirty_before += tile_dirty;
        chunks_before += tile_chunks;
        if (pushed >= 0) dirty[di] = pushed;
        if (nd) {
            const bmf_world_node& q = nodes[pushed];
            for (int k = 0; k < nd; k++) {
                const int slot = ci + k;
                if (slot >= cap) break;
                bmf_chunk_desc d;
                bmf_lod_item it;
                if (nd == 8) {  // child k of the split node
                    const float cs = q.size * 0.5f;
                    d.pos[0] = q.pos[0] + (float)kMCDX[k] * cs;
                    d.pos[1] = q.pos[1] + (float)kMCDY[k] * cs;
                    d.pos[2] = q.pos[2] + (float)kMCDZ[k] * cs;
                    d.size = q.size * 0.5f;
                    it.level = q.level + 1;
                    it.child = k;
                    it.action = BMF_LOD_SPLIT;
                    it.morton = (q.morton << 3) | (uint64_t)(kMCDX[k] | (kMCDY[k] << 1) | (kMCDZ[k] << 2));
                } else {  // the grouped node's own chunk
                    d.pos[0] = q.pos[0];
                    d.pos[1] = q.pos[1];
                    d.pos[2] = q.pos[2];
                    d.size = q.size;
                    it.level = q.level;
                    it.child = -1;
                    it.action = BMF_LOD_GROUP;
                    it.morton = q.morton;
                }
                it.node = pushed;
                d.dim = P.chunk_resolution;
                d.iters = P.process_iters;
                d.boundary_processing = P.boundary_processing;
                d.overlap = (it.level == P.max_level && (!P.boundary_processing || P.process_iters == 0))
                                ? 0.0f
                                : P.overlap + 0.005f * (float)P.process_iters;
                descs[slot] = d;
                items[slot] = it;
            }
        }
    }
    if (t == 0) {
        counts[0] = dirty_before;
        counts[1] = chunks_before;
    }
}

bool finite_f(float x) { return std::isfinite(x); }

}  // namespace
}  // namespace bmf

using namespace bmf;

extern "C" {

int bmf_lod_queue(bmf_ctx* c, const bmf_world_node* nodes, int32_t n_nodes, const int32_t* renderables,
                  int32_t n_renderables, const bmf_lod_params* P, uint32_t* flags_out, int32_t* dirty_out,
                  bmf_lod_item* items_out, int32_t cap_items, int32_t* n_dirty, int32_t* n_chunks) {
    if (!c || !P || n_nodes < 0 || n_renderables < 0 || cap_items < 0) return BMF_EINVAL;
    if (n_nodes > 0 && (!nodes || !flags_out)) return BMF_EINVAL;
    if (n_renderables > 0 && (!renderables || !dirty_out)) return BMF_EINVAL;
    if (cap_items > 0 && !items_out) return BMF_EINVAL;
    if (P->chunk_resolution < 2 || P->chunk_resolution > BMF_MAX_DIM || P->process_iters < 0 || P->process_iters > 64 ||
        (P->boundary_processing != 0 && P->boundary_processing != 1) || P->max_gen < 0 || !finite_f(P->overlap) ||
        P->overlap < 0.0f || !finite_f(P->split_multiplier) || !finite_f(P->group_multiplier) ||
        !finite_f(P->size_modifier) || !finite_f(P->focus[0]) || !finite_f(P->focus[1]) || !finite_f(P->focus[2]))
        return BMF_EINVAL;
    const float ov = P->overlap + 0.005f * (float)P->process_iters;
    if (ov > 1.0f) return BMF_EINVAL;  // the descriptors' overlap bound
    BMF_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    c->queued_n = -1;
    if (n_nodes > c->lod_cap_nodes) {
        if (c->d_lod_first) (void)hipFree(c->d_lod_first);
        c->d_lod_first = nullptr;
        c->lod_cap_nodes = 0;
        BMF_HIP(hipMalloc((void**)&c->d_lod_first, sizeof(int) * (size_t)n_nodes));
        c->lod_cap_nodes = n_nodes;
    }
    if (n_renderables > c->lod_cap_list) {
        if (c->d_lod_code) (void)hipFree(c->d_lod_code);
        c->d_lod_code = nullptr;
        c->lod_cap_list = 0;
        BMF_HIP(hipMalloc((void**)&c->d_lod_code, (size_t)n_renderables));
        c->lod_cap_list = n_renderables;
    }
    if (!c->d_lod_counts) BMF_HIP(hipMalloc((void**)&c->d_lod_counts, sizeof(int) * 2));
    if (!c->h_lod_counts) BMF_HIP(hipHostMalloc((void**)&c->h_lod_counts, sizeof(int) * 2, hipHostMallocDefault));
    // the descriptor arena is written by this call's kernel, after any queued
    // extraction that reads it (stream order)
    const int cap = std::min(cap_items, c->lim.max_chunks);
    hipLaunchKernelGGL(k_lod_queue, dim3(1), dim3(kLodThreads), 0, s, nodes, n_nodes, renderables, n_renderables, *P,
                       flags_out, c->d_lod_first, c->d_lod_code, dirty_out, items_out, c->A.descs, cap,
                       c->d_lod_counts);
    BMF_HIP(hipGetLastError());
    BMF_HIP(hipMemcpyAsync(c->h_lod_counts, c->d_lod_counts, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    BMF_HIP(hipStreamSynchronize(s));
    if (n_dirty) *n_dirty = c->h_lod_counts[0];
    if (n_chunks) *n_chunks = c->h_lod_counts[1];
    if (c->h_lod_counts[1] > cap) return BMF_ECAPACITY;
    c->queued_n = c->h_lod_counts[1];
    c->queued_dim = P->chunk_resolution;
    c->queued_iters = P->process_iters;
    return BMF_OK;
}

}  // extern "C"
