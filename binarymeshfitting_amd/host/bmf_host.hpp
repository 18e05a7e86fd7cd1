// Headless C++ host layer over the C ABI (include/bmf.h) that mirrors the
// reference's extraction interface: Sampler (Sampler.hpp:24-34) and its
// create_sampler_* factories, DMCChunk's public fields (DMCChunk.hpp:24-60),
// VerticesIndicesBlock (ChunkBlocks.hpp:161-179), DualVertex
// (Vertices.hpp:5-24), WorldProperties (WorldOctree.hpp:19-34) and
// ChunkGenerator::process_queue (ChunkGenerator.cpp:27-147).  A caller that
// used the reference classes fills the same fields and gets the same
// per-chunk results, computed on the GPU.  glm is not required: vec3 is a
// layout-compatible stand-in.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/bmf.h"

namespace bmf_host {

struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct ivec3 {
    int x = 0, y = 0, z = 0;
};

// Vertices.hpp:5-24.  Fields the extraction sets: boundary, index, p, color
// (DMCChunk.cpp:676-689); the rest are zero.
struct DualVertex {
    bool boundary = false;
    uint8_t mask = 0;
    uint32_t index = 0;
    uint8_t valence = 0;
    uint8_t init_valence = 0;
    uint8_t adj_next = 0;
    uint32_t adj_offset = 0;
    uint16_t edge_mask = 0;
    float s = 0;
    ivec3 xyz;
    vec3 p;
    vec3 n;
    vec3 avg;
    vec3 color;
};

// ChunkBlocks.hpp:161-179
struct VerticesIndicesBlock {
    std::vector<DualVertex> vertices;
    std::vector<uint32_t> mesh_indexes;
    void init() {
        vertices.clear();
        mesh_indexes.clear();
    }
};

// Sampler.hpp:24-34: the block function becomes a device sampler descriptor.
struct Sampler {
    float world_size = 256.0f;
    bmf_sampler_desc desc{};
};

namespace NoiseSamplers {
// NoiseSamplerProperties (NoiseSampler.hpp:7-22) restricted to what the FBM
// terrain uses: d = height_amp * fbm(p) - (p.y - base_height).
struct TerrainProperties {
    float amp = 24.0f;
    float base_height = 32.0f;
    int octaves = 5;
    float frequency = 0.02f;
    float lacunarity = 2.0f;
    float gain = 0.5f;
    uint32_t seed = 1337;
};
void create_sampler_terrain_fbm_3d(Sampler* s, const TerrainProperties& p);

// The reference's noise sampler factories (NoiseSampler.hpp:91-149) as NOISE
// samplers with the project's own noise (bmf_noise_preset): the block shapes
// of terrain2d_block, terrain2d_pert_block, terrain3d_block,
// terrain3d_pert_block and windy3d_block.
void create_sampler_terrain_2d(Sampler* s, uint32_t seed = 1337);
void create_sampler_terrain_pert_2d(Sampler* s, uint32_t seed = 1337);
void create_sampler_terrain_3d(Sampler* s, uint32_t seed = 1337);
void create_sampler_terrain_pert_3d(Sampler* s, uint32_t seed = 1337);
void create_sampler_windy_3d(Sampler* s, uint32_t seed = 1337);

// NoiseSamplerProperties (NoiseSampler.hpp:7-22): what terrain2d_pert_block
// reads from the world (g_scale, height, and the perturbation's octaves,
// amplitude, frequency and gain, in FastNoiseSIMD's frequency-scaled units);
// the defaults are the world's (WorldOctree.cpp:49-54).
struct NoiseSamplerProperties {
    float g_scale = 0.25f;
    float height = 75.0f;
    int octaves = 13;
    float amp = 0.87f;
    float frequency = 0.585f;
    float gain = 0.488f;
};
// terrain2d_pert_block called with `properties` (NoiseSampler.cpp:150-158).
void apply_properties(Sampler* s, const NoiseSamplerProperties& p);
}  // namespace NoiseSamplers

namespace ImplicitFunctions {
// ImplicitFunctions::create_sampler(sphere | cuboid | torus_z | plane_y)
// (ImplicitSampler.hpp:51-59); kind is a BMF_SAMPLER_* implicit kind.
Sampler create_sampler(int kind, float world_size);
}  // namespace ImplicitFunctions

// WorldProperties (WorldOctree.hpp:19-34) with the reference defaults
// (WorldOctree.cpp:20-33).
struct WorldProperties {
    int max_level = 7;
    int min_level = 1;
    int process_iters = 0;
    int chunk_resolution = 32;
    float overlap = 0.035f;
    bool boundary_processing = false;
    // The reference's SMOOTH_NORMALS build switch (MeshProcessor<3>'s
    // smooth_normals, off on its live path): fills DualVertex::n.
    bool smooth_normals = false;
};

enum GenerationStage { GENERATION_STAGES_GENERATING = 0, GENERATION_STAGES_NEEDS_UPLOAD = 1, GENERATION_STAGES_DONE = 2 };

// DMCChunk's public state (DMCChunk.hpp:24-60) after extraction.
class DMCChunk {
public:
    int id = 0;
    int level = 0;
    uint32_t dim = 32;
    float size = 0;
    vec3 pos;
    bool contains_mesh = false;
    Sampler sampler;
    VerticesIndicesBlock* vi = nullptr;
    vec3 overlap_pos;
    vec3 bound_start;
    float bound_size = 0;
    float scale = 0;

    // DMCChunk::init (DMCChunk.cpp:58-77); dim comes from the world's
    // chunk_resolution (WorldOctree.cpp:263-269).
    void init(vec3 pos, float size, int level, const Sampler& sampler);
    ~DMCChunk();
};

// The slice of WorldOctreeNode the generator touches.
struct WorldOctreeNode {
    DMCChunk* chunk = nullptr;
    int level = 0;
    int generation_stage = GENERATION_STAGES_GENERATING;
};

// ChunkGenerator (ChunkGenerator.hpp:15-57).  process_queue runs the whole
// batch through one bmf_extract call.
class ChunkGenerator {
public:
    ~ChunkGenerator();
    // Creates the GPU context (device arenas) and sets the sampler.  Returns a BMF_* status.
    int init(const WorldProperties& props, const Sampler& sampler, int device = 0, int max_chunks = 4096);
    // ChunkGenerator::process_queue / extract_chunk: fills every node's chunk
    // (contains_mesh, geometry, vi) and generation_stage.  Returns a BMF_*
    // status; on BMF_ECAPACITY the arenas are grown to the reported needs and
    // the batch is run again once.
    int process_queue(std::vector<WorldOctreeNode*>& batch);
    // WorldOctreeNode::format -> GLChunk::format_data over the last batch
    // (WorldOctreeNode.cpp:68-86, GLChunk.cpp:278-297, DMCChunk.cpp:578-591),
    // on the device into `out`: bmf_format.  Returns a BMF_* status.
    int format_batch(const bmf_format_options& options, const bmf_vbo& out, int64_t* vertices, int64_t* indices);
    // The watcher's LOD pass over a node table and renderables list in device
    // memory (WorldWatcher::check_leaves / process_batch), then the extraction
    // of the generate batch it queued on the device: bmf_lod_queue +
    // bmf_extract_queued.  Returns a BMF_* status.
    int process_lod(const bmf_world_node* nodes, int32_t n_nodes, const int32_t* renderables, int32_t n_renderables,
                    const bmf_lod_params& params, uint32_t* flags_out, int32_t* dirty_out, bmf_lod_item* items_out,
                    int32_t cap_items, int32_t* n_dirty, int32_t* n_chunks);
    int mode = BMF_MODE_BATCHED;
    WorldProperties properties;

private:
    int create_context(int64_t max_samples, int64_t max_vertices, int64_t max_indices);
    int run(std::vector<WorldOctreeNode*>& batch, int64_t* need_v, int64_t* need_i);
    bmf_ctx* ctx_ = nullptr;
    Sampler sampler_;
    int device_ = 0;
    int max_chunks_ = 0;
    int64_t max_samples_ = 0, max_vertices_ = 0, max_indices_ = 0;
    std::vector<bmf_chunk_desc> descs_;
    std::vector<bmf_chunk_result> results_;
};

}  // namespace bmf_host
