// Host layer over the C ABI mirroring the reference's ChunkGenerator /
// DMCChunk / Sampler interface (see bmf_host.hpp).
#include "bmf_host.hpp"

#include <algorithm>

namespace bmf_host {

void NoiseSamplers::create_sampler_terrain_fbm_3d(Sampler* s, const TerrainProperties& p) {
    s->desc = bmf_sampler_desc{};
    s->desc.kind = BMF_SAMPLER_TERRAIN;
    s->desc.world_size = s->world_size;
    s->desc.amp = p.amp;
    s->desc.y0 = p.base_height;
    s->desc.freq = p.frequency;
    s->desc.lacunarity = p.lacunarity;
    s->desc.gain = p.gain;
    s->desc.octaves = p.octaves;
    s->desc.seed = p.seed;
}

namespace {
void noise_preset(Sampler* s, int32_t which, uint32_t seed) {
    bmf_sampler_desc d{};
    if (bmf_noise_preset(which, s->world_size, seed, &d) == BMF_OK) s->desc = d;
}
}  // namespace

void NoiseSamplers::create_sampler_terrain_2d(Sampler* s, uint32_t seed) { noise_preset(s, BMF_PRESET_TERRAIN2D, seed); }
void NoiseSamplers::create_sampler_terrain_pert_2d(Sampler* s, uint32_t seed) {
    noise_preset(s, BMF_PRESET_TERRAIN2D_PERT, seed);
}
void NoiseSamplers::create_sampler_terrain_3d(Sampler* s, uint32_t seed) { noise_preset(s, BMF_PRESET_TERRAIN3D, seed); }
void NoiseSamplers::create_sampler_terrain_pert_3d(Sampler* s, uint32_t seed) {
    noise_preset(s, BMF_PRESET_TERRAIN3D_PERT, seed);
}
void NoiseSamplers::create_sampler_windy_3d(Sampler* s, uint32_t seed) { noise_preset(s, BMF_PRESET_WINDY3D, seed); }

void NoiseSamplers::apply_properties(Sampler* s, const NoiseSamplerProperties& p) {
    // the perturbation acts on c = g_scale p here, before the noise frequency
    const float f = s->desc.freq;
    s->desc.g_scale = p.g_scale;
    s->desc.height = p.height;
    s->desc.perturb = BMF_PERTURB_GRADIENT;
    s->desc.perturb_octaves = p.octaves;
    s->desc.perturb_amp = p.amp / f;
    s->desc.perturb_freq = p.frequency * f;
    s->desc.perturb_gain = p.gain;
}

Sampler ImplicitFunctions::create_sampler(int kind, float world_size) {
    Sampler s;
    s.world_size = world_size;
    s.desc.kind = kind;
    s.desc.world_size = world_size;
    return s;
}

void DMCChunk::init(vec3 p, float sz, int lvl, const Sampler& smp) {
    pos = p;
    size = sz;
    level = lvl;
    sampler = smp;
    contains_mesh = false;
    delete vi;
    vi = nullptr;
}

DMCChunk::~DMCChunk() { delete vi; }

ChunkGenerator::~ChunkGenerator() {
    if (ctx_) bmf_destroy(ctx_);
}

int ChunkGenerator::create_context(int64_t max_samples, int64_t max_vertices, int64_t max_indices) {
    if (ctx_) bmf_destroy(ctx_);
    ctx_ = nullptr;
    bmf_limits lim{max_chunks_, max_samples, max_vertices, max_indices};
    int rc = bmf_create(&ctx_, device_, &lim);
    if (rc) return rc;
    max_samples_ = max_samples;
    max_vertices_ = max_vertices;
    max_indices_ = max_indices;
    rc = bmf_set_sampler(ctx_, &sampler_.desc);
    if (rc) return rc;
    bmf_mesh_options o{};
    o.mesh = BMF_MESH_TRIANGLES;
    o.qef_tolerance = 0.1f;
    o.normals = properties.smooth_normals ? 1 : 0;
    return bmf_set_mesh_options(ctx_, &o);
}

int ChunkGenerator::init(const WorldProperties& props, const Sampler& sampler, int device, int max_chunks) {
    properties = props;
    sampler_ = sampler;
    device_ = device;
    max_chunks_ = max_chunks;
    const int64_t d = std::max(2, props.chunk_resolution);
    return create_context((int64_t)max_chunks * d * d * d / 8 + d * d * d, 1 << 22, 3 << 22);
}

int ChunkGenerator::run(std::vector<WorldOctreeNode*>& batch, int64_t* need_v, int64_t* need_i) {
    const WorldProperties& P = properties;
    const int n = (int)batch.size();
    descs_.resize(n);
    int64_t samples = 0;
    for (int i = 0; i < n; i++) {
        const DMCChunk* c = batch[i]->chunk;
        // ChunkGenerator.cpp:98
        const float overlap = (batch[i]->level == P.max_level && (!P.boundary_processing || P.process_iters == 0))
                                  ? 0.0f
                                  : P.overlap + 0.005f * (float)P.process_iters;
        descs_[i] = bmf_chunk_desc{{c->pos.x, c->pos.y, c->pos.z}, c->size, (int32_t)c->dim, overlap,
                                   P.process_iters, P.boundary_processing ? 1 : 0};
        samples += (int64_t)c->dim * c->dim * c->dim;
    }
    if (samples > max_samples_) {  // every chunk may need sampling: size the density arena for it
        int rc = create_context(samples, max_vertices_, max_indices_);
        if (rc) return rc;
    }
    int rc = bmf_extract(ctx_, descs_.data(), n, mode);
    if (rc) return rc;
    results_.resize(n);
    return bmf_result(ctx_, results_.data(), n, need_v, need_i);
}

int ChunkGenerator::process_queue(std::vector<WorldOctreeNode*>& batch) {
    const int n = (int)batch.size();
    if (n == 0) return BMF_OK;
    if (n > max_chunks_) return BMF_EINVAL;
    int64_t need_v = 0, need_i = 0;
    int rc = run(batch, &need_v, &need_i);
    if (rc == BMF_ECAPACITY) {
        rc = create_context(max_samples_, std::max(max_vertices_, need_v), std::max(max_indices_, need_i));
        if (rc) return rc;
        rc = run(batch, &need_v, &need_i);
    }
    if (rc) return rc;
    std::vector<float> p, nrm;
    std::vector<uint8_t> b;
    for (int i = 0; i < n; i++) {
        DMCChunk* c = batch[i]->chunk;
        const bmf_chunk_result& r = results_[i];
        c->contains_mesh = r.contains_mesh != 0;
        c->scale = r.delta;
        c->overlap_pos = vec3(r.overlap_pos[0], r.overlap_pos[1], r.overlap_pos[2]);
        // DMCChunk.cpp:100-101
        c->bound_size = c->size * (1.0f + descs_[i].overlap * 2.0f) * 0.5f;
        c->bound_start = vec3(c->overlap_pos.x + c->bound_size, c->overlap_pos.y + c->bound_size,
                              c->overlap_pos.z + c->bound_size);
        delete c->vi;
        c->vi = nullptr;
        if (r.contains_mesh) {
            // the reference allocates vi for every chunk with a mesh flag
            // (DMCChunk.cpp:170-177), even when it stays empty
            c->vi = new VerticesIndicesBlock();
            c->vi->init();
            if (r.vertex_count) {
                p.resize(3 * (size_t)r.vertex_count);
                b.resize(r.vertex_count);
                c->vi->mesh_indexes.resize(r.index_count);
                int64_t nv = 0, ni = 0;
                rc = bmf_mesh_to_host(ctx_, i, p.data(), c->vi->mesh_indexes.data(), b.data(), r.vertex_count,
                                      r.index_count, &nv, &ni);
                if (rc) return rc;
                if (properties.smooth_normals) {
                    nrm.resize(3 * (size_t)r.vertex_count);
                    rc = bmf_normals_to_host(ctx_, i, nrm.data(), r.vertex_count);
                    if (rc) return rc;
                }
                c->vi->vertices.resize(r.vertex_count);
                for (int v = 0; v < r.vertex_count; v++) {  // DMCChunk::calculate_dual_vertex (DMCChunk.cpp:676-689)
                    DualVertex& dv = c->vi->vertices[v];
                    dv.index = (uint32_t)v;
                    dv.p = vec3(p[3 * v], p[3 * v + 1], p[3 * v + 2]);
                    dv.color = vec3(1, 1, 1);
                    dv.boundary = b[v] != 0;
                    if (properties.smooth_normals) dv.n = vec3(nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2]);
                }
            }
        }
        // ChunkGenerator.cpp:137-143
        batch[i]->generation_stage = c->vi ? GENERATION_STAGES_NEEDS_UPLOAD : GENERATION_STAGES_DONE;
    }
    return BMF_OK;
}

int ChunkGenerator::format_batch(const bmf_format_options& options, const bmf_vbo& out, int64_t* vertices,
                                 int64_t* indices) {
    if (!ctx_) return BMF_ESTATE;
    return bmf_format(ctx_, &options, &out, vertices, indices);
}

int ChunkGenerator::process_lod(const bmf_world_node* nodes, int32_t n_nodes, const int32_t* renderables,
                                int32_t n_renderables, const bmf_lod_params& params, uint32_t* flags_out,
                                int32_t* dirty_out, bmf_lod_item* items_out, int32_t cap_items, int32_t* n_dirty,
                                int32_t* n_chunks) {
    if (!ctx_) return BMF_ESTATE;
    int rc = bmf_lod_queue(ctx_, nodes, n_nodes, renderables, n_renderables, &params, flags_out, dirty_out, items_out,
                           cap_items, n_dirty, n_chunks);
    if (rc) return rc;
    rc = bmf_extract_queued(ctx_, mode);
    if (rc) return rc;
    results_.resize((size_t)*n_chunks);
    return bmf_result(ctx_, results_.data(), *n_chunks, nullptr, nullptr);
}

}  // namespace bmf_host
