// Internal structures shared by the HIP kernels and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bmf.h"

namespace bmf {

// Samples handled by one work-queue tile of the sampling kernel: 256 threads
// x rounds (4, 8 or 16, chosen at run time: Arenas::sample_tile); a multiple
// of 64 so that every wave owns whole label words.
constexpr int kSampleThreads = 256;
constexpr int kSampleTileDefault = 2048;
constexpr int kRowThreads = 256;      // 4 waves; one x-plane (or row segment of one) per block, one label row per wave
constexpr int kMaxPlaneSplit = 8;     // row segments per x-plane when few planes are sampled
constexpr int kSplitTarget = 512;     // row-kernel work items aimed at before splitting stops
constexpr int kSmoothThreads = 256;   // one vertex per thread
constexpr int kMaxRowWords = 4;       // ceil(BMF_MAX_DIM / 64)
constexpr int kXcds = 8;              // MI355X: 8 XCDs, each with its own L2

enum ChunkState : int32_t {
    kStateAir = 0,      // culled: every sample is air (d < 0 everywhere)
    kStateSolid = 1,    // culled: every sample is solid
    kStateSampled = 2,  // densities and labels computed
    kStateSkipped = 3   // did not fit an arena
};

// Device view of the sampler (Sampler + NoiseSamplerProperties).
struct SamplerDev {
    int32_t kind;
    float world_size;
    float amp, y0;
    int32_t octaves;
    float inv_norm;
    uint32_t seed;
    int32_t n_spheres, n_boxes;
    float fo[BMF_MAX_OCTAVES];  // per-octave frequency f0 * L^o (FP32 product chain)
    float ao[BMF_MAX_OCTAVES];  // per-octave amplitude g^o
    // reach[k] >= the most the FP32 partial sum after k octaves can still
    // move: sum of |a_o| for o >= k plus the rounding slack of those steps
    float reach[BMF_MAX_OCTAVES + 1];
    const float* spheres;       // device, n_spheres x 4
    const float* boxes;         // device, n_boxes x 6
    const float* ext;           // EXTERNAL: caller's densities, chunks packed in descriptor order
    // NOISE: basis, fractal, density map and perturbation (bmf_sampler_desc)
    int32_t noise_dims, noise_basis, noise_fractal;
    float g_scale, y_scale, height;
    int32_t perturb, perturb_octaves;
    float perturb_amp;
    float pf[BMF_MAX_PERTURB_OCTAVES];  // perturbation frequency f * 2^k (FP32 chain)
    float pa[BMF_MAX_PERTURB_OCTAVES];  // GRADIENT: amp * gain^k; OFFSET_XZ: gain^k
    float pinv;                         // OFFSET_XZ: 1 / sum of pa
    // k_plan's chunk culling: n lies in [nlo, nhi] for every sample (cull = 0: no bound)
    float nlo, nhi;
    int32_t cull;
};

// Per-chunk state in device memory.
struct ChunkMeta {
    float opos[3];
    float delta;
    int32_t dim;
    int32_t iters;
    int32_t bp;
    int32_t state;
    int64_t dens_off;   // float index into the density arena
    int64_t lab_off;    // u64 index into the label bitmap arena
    int64_t lab_words;  // u64 words of this chunk's bitmap
    int64_t row_off;    // index into the per-row arrays (rows (x, y) in x-major order)
    uint32_t any_air;   // set by the sampling kernel
    uint32_t any_solid;
    int32_t contains_mesh;
    int32_t status;
    int32_t V;          // vertices
    int32_t T;          // primitives (triangles, or quads in the quad mode)
    int32_t ipp;        // indices per primitive: 3 or 4
    int32_t csg_ns;     // CSG: spheres and boxes that can reach above the floor in this chunk,
    int32_t csg_nb;     //      listed in Arenas::csg_list (sphere ids, then box ids, in order)
    int64_t v_off;      // output offsets
    int64_t i_off;
    int64_t plane_slot; // plane_map entry of this chunk's x-plane 0 (sampled chunks)
};

// Global counters, reset by the plan kernel.
struct Counters {
    unsigned long long sample_head;
    unsigned long long v_bump;
    unsigned long long i_bump;
    unsigned long long need_v;
    unsigned long long need_i;
    int64_t total_tiles;
    int64_t total_samples;
    int64_t total_vtiles;
    int64_t total_ttiles;
    int32_t n_smooth;
    int32_t overflow;
    int32_t error;
    int32_t plane_split[4];    // row segments per x-plane item of row-word class W (index W-1)
    int64_t class_base[4];     // first plane_map entry of class W's planes
    int64_t class_planes[4];   // sampled planes of class W
};

// Device pointers of one context.
struct Arenas {
    ChunkMeta* meta;           // [max_chunks]
    bmf_chunk_desc* descs;     // [max_chunks]
    bmf_chunk_result* results; // [max_chunks]
    int64_t* tile_prefix;      // [max_chunks + 1] sampling tiles
    int64_t* plan_part;        // [40 nb + 64] multi-block plan passes (nb = max_chunks / 1024 rounded up):
                               // block totals and bases, flags, class and group bases
    int64_t* plan_tmp;         // [2 max_chunks] multi-block smoothing plan: per-chunk exclusive tile sums
    unsigned long long* stat_blocks;  // [2][stat_cap] terrain octaves evaluated per workgroup of
                                      // the sampling and the refinement kernels (plain stores, no atomics)
    int32_t stat_cap;
    int2* plane_map;           // [max_chunks * BMF_MAX_DIM] (chunk, x) of each sampled x-plane, grouped by
                               // row-word class (written by k_sample's prologue)
    int32_t* smooth_list;      // [max_chunks] chunks that run the smoothing
    int64_t* smooth_prefix;    // [max_chunks + 1] vertex tiles of the listed chunks
    int64_t* smooth_tprefix;   // [max_chunks + 1] triangle tiles of the listed chunks
    int64_t* smooth_xcd_v;     // [kXcds + 1] vertex-tile range of each XCD group
    int64_t* smooth_xcd_t;     // [kXcds + 1] triangle-tile range of each XCD group
    int2* smooth_vmap;         // [vertex tiles] (chunk, tile within the chunk) of each smoothing vertex tile
    int2* smooth_tmap;         // [triangle tiles] the same for the triangle tiles
    uint32_t* refine;          // [max_samples] per x-plane queue of samples (y*dim + z) to finish,
                               // a region of rows [y0, y1) per row segment
    int32_t row_split;         // row segments per x-plane item, or -1 for the automatic choice
    uint32_t* plane_refine_n;  // [max_planes * kMaxPlaneSplit] queue length of each (x-plane, row segment) item
    float* dens;               // [max_samples]
    uint64_t* labels;          // [max_samples / 64 + max_chunks]
    uint64_t* exact;           // same layout as labels: density of the sample is final
    uint32_t* row_v;           // [max_rows] per label row: vertex count, then base
    uint32_t* row_t;           // [max_rows] per label row: triangle count, then base
    uint32_t* cell_tb;         // [max_samples] first triangle of each surface cell
    uint8_t* cell_mask;        // [max_samples] corner mask of each surface cell
    float* pos;                // [max_vertices * 3]
    float4* pos4;              // [max_vertices] positions during the smoothing
    float4* cen4;              // [max_indices / 3] triangle centroids
    uint8_t* boundary;         // [max_vertices]
    uint32_t* vedge;           // [max_vertices] lattice edge of each vertex: x | y<<8 | z<<16 | axis<<24
    uint4* vring;              // [max_vertices] ring cells of each vertex, packed by the first smoothing step
    uint32_t* idx;             // [max_indices]
    Counters* ctr;
    SamplerDev* sampler;
    int64_t cap_samples, cap_rows, cap_vertices, cap_indices;
    int32_t cap_chunks;
    int64_t ext_base;          // EXTERNAL: offset of this launch's first chunk in the caller's densities
    int32_t sample_tile;       // samples per sampling work-queue tile: 1024, 2048 or 4096
    int32_t smoothing_follows; // the launch runs the smoothing kernels after the emit kernel (which then
                               // leaves the output positions of the vertices they move to them)
    uint16_t* csg_list;        // CSG: [chunk][csg_stride] primitives per chunk (k_plan), or null
    int32_t csg_stride;        // n_spheres + n_boxes
    // mesh options (BMF_MESH_*, BMF_PLACE_*): the quad mode's extra arenas
    int32_t mesh_mode;
    int32_t placement;
    float qef_tol;
    int32_t quad_tris;         // quads written as two triangles each
    uint32_t* ring;            // [kRingSlots][cap_vertices] primitives around each vertex (quad mode)
    uint32_t* ring_n;          // [cap_vertices] ring sizes
    float* qef_err;            // [cap_vertices] QEF residual of each vertex (QEF placement)
    // vertex normals of the triangle path (bmf_mesh_options::normals)
    int32_t normals;
    float4* nrm4;              // [cap_vertices] normal state during the smoothing
    float4* fn4;               // [cap_indices / 3] per-triangle normal of the current step
    float* nrm;                // [cap_vertices * 3] unit normals (output)
};

// Quad mode: a vertex (one polygon of a cell) has at most 7 edges, so at most
// 7 quads, or 14 triangles when quads are written as triangles.
constexpr int kRingSlots = 16;

inline int row_word_class(int dim) { return 1 << (((dim + 63) >> 6) - 1); }

// Launch wrappers (bmf_kernels.hip).
hipError_t launch_plan(const Arenas& A, int n, int allow_cull, int reset_bumps, hipStream_t s);
// adaptive: terrain samples stop adding octaves once their label is decided
// kind: the sampler the device struct holds for this launch (BMF_SAMPLER_*)
// slab: the sampler is a 2D noise field (its columns are evaluated once per tile)
hipError_t launch_sample(const Arenas& A, int n, int blocks, int adaptive, int kind, int slab, hipStream_t s);
// wclasses: bit W-1 set when the batch holds chunks with ceil(dim/64) == W
// (the classes' kernels one after another on s).
hipError_t launch_count(const Arenas& A, int n, int blocks, int wclasses, hipStream_t s);
hipError_t launch_scan(const Arenas& A, int n, int blocks, hipStream_t s);
hipError_t launch_emit(const Arenas& A, int n, int blocks, int wclasses, hipStream_t s);
hipError_t launch_refine(const Arenas& A, int n, int blocks, hipStream_t s);
hipError_t launch_smooth(const Arenas& A, int m, int blocks, hipStream_t s);
// (chunk, tile) of every smoothing vertex and primitive tile, once per frame before the first step
hipError_t launch_smooth_map(const Arenas& A, int blocks, hipStream_t s);
// Quad mode (bmf_dmc.hip): counts after launch_count/launch_refine, emission
// after launch_scan, smoothing over the per-vertex rings.
hipError_t launch_dmc_emit(const Arenas& A, int n, int blocks, int wclasses, hipStream_t s);
hipError_t launch_dmc_smooth(const Arenas& A, int m, int blocks, hipStream_t s);
hipError_t launch_ref_labels(const Arenas& A, uint32_t* words, hipStream_t s);
hipError_t launch_ref_masks(const Arenas& A, uint64_t* masks, hipStream_t s);

}  // namespace bmf
