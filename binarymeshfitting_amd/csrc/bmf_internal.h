// Internal structures shared by the HIP kernels and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bmf.h"

namespace bmf {

// Samples handled by one work-queue tile of the sampling kernel: 256 threads
// x 4 rounds; a multiple of 64 so that every wave owns whole label words.
constexpr int kSampleThreads = 256;
constexpr int kSampleRounds = 4;
constexpr int kSampleTile = kSampleThreads * kSampleRounds;
constexpr int kRowThreads = 256;      // 4 waves, one label row each
constexpr int kSmoothThreads = 256;
constexpr int kMaxRowWords = 4;       // ceil(BMF_MAX_DIM / 64)

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
    const float* spheres;       // device, n_spheres x 4
    const float* boxes;         // device, n_boxes x 6
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
    int64_t row_off;    // index into the per-row arrays
    uint32_t any_air;   // set by the sampling kernel
    uint32_t any_solid;
    int32_t contains_mesh;
    int32_t status;
    int32_t V;          // vertices
    int32_t T;          // triangles
    int64_t v_off;      // output offsets
    int64_t i_off;
};

// Global counters, reset by the plan kernel.
struct Counters {
    unsigned long long sample_head;
    unsigned long long v_bump;
    unsigned long long i_bump;
    unsigned long long need_v;
    unsigned long long need_i;
    int64_t total_tiles;
    int64_t total_rows;
    int64_t total_samples;
    int64_t total_vtiles;
    int64_t total_ttiles;
    int32_t overflow;
    int32_t error;
};

// Device pointers of one context.
struct Arenas {
    ChunkMeta* meta;           // [max_chunks]
    bmf_chunk_desc* descs;     // [max_chunks]
    bmf_chunk_result* results; // [max_chunks]
    int64_t* tile_prefix;      // [max_chunks + 1]
    int64_t* row_prefix;       // [max_chunks + 1]
    int64_t* vtile_prefix;     // [max_chunks + 1]
    int64_t* ttile_prefix;     // [max_chunks + 1]
    float* dens;               // [max_samples]
    uint64_t* labels;          // [max_samples / 64 + max_chunks]
    uint32_t* row_v;           // [max_rows]
    uint32_t* row_t;           // [max_rows]
    float* pos;                // [max_vertices * 3]
    uint8_t* boundary;         // [max_vertices]
    uint32_t* valence;         // [max_vertices]
    uint32_t* adj;             // [max_vertices * BMF_MAX_VALENCE]
    uint32_t* idx;             // [max_indices]
    float* cen;                // [max_indices] (one float3 per triangle)
    Counters* ctr;
    SamplerDev* sampler;
    int64_t cap_samples, cap_rows, cap_vertices, cap_indices;
    int32_t cap_chunks;
};

// Launch wrappers (bmf_kernels.hip).
struct LaunchShape {
    int sample_blocks;
    int row_blocks;
    int scan_blocks;
    int smooth_blocks;
};

hipError_t launch_plan(const Arenas& A, int n, int allow_cull, int reset_bumps, hipStream_t s);
hipError_t launch_sample(const Arenas& A, int n, int blocks, hipStream_t s);
hipError_t launch_count(const Arenas& A, int n, int blocks, hipStream_t s);
hipError_t launch_scan(const Arenas& A, int n, int blocks, hipStream_t s);
hipError_t launch_emit(const Arenas& A, int n, int blocks, hipStream_t s);
hipError_t launch_sort_adj(const Arenas& A, int n, int blocks, hipStream_t s);
hipError_t launch_dual(const Arenas& A, int n, int m, int blocks, hipStream_t s);
hipError_t launch_primal(const Arenas& A, int n, int m, int blocks, hipStream_t s);
hipError_t launch_finalize(const Arenas& A, int n, hipStream_t s);
hipError_t launch_ref_labels(const Arenas& A, uint32_t* words, hipStream_t s);
hipError_t launch_ref_masks(const Arenas& A, uint64_t* masks, hipStream_t s);

}  // namespace bmf
