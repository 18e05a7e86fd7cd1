// Runs ChunkGenerator::process_queue over a batch of terrain chunks (the
// reference's world-update path, headless) and prints one JSON line per
// chunk: contains_mesh, vertex and index counts, and the mesh checksum
// (sum over the u32 stream of position bit patterns then indices of
// word_k * (2k + 1), times 0x9E3779B97F4A7C15, modulo 2^64).
// With smooth_normals = 1 it also prints the same checksum over the normals.
// Every node sits at `level` (default max_level = 7): below max_level the
// overlap rule of ChunkGenerator.cpp:98 gives 0.035 + 0.005 * iters.
// The sampler is the FBM terrain, or `world` (the reference world's
// terrain2d_pert sampler with its noise_properties, WorldOctree.cpp:40-54)
// or `pert3d` (create_sampler_terrain_pert_3d, DebugScene.cpp:242), whose
// chunk rows start n_side / 2 chunks below y = 0.
//   bmf_host_demo <n_side> <resolution> <iters> <mode: 0 plain, 1 batched> [smooth_normals] [level] [sampler]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "bmf_host.hpp"

using namespace bmf_host;

int main(int argc, char** argv) {
    const int n_side = argc > 1 ? std::atoi(argv[1]) : 2;
    const int res = argc > 2 ? std::atoi(argv[2]) : 33;
    const int iters = argc > 3 ? std::atoi(argv[3]) : 2;
    const int mode = argc > 4 ? std::atoi(argv[4]) : BMF_MODE_BATCHED;
    const bool normals = argc > 5 && std::atoi(argv[5]) != 0;
    WorldProperties props;
    const int level = argc > 6 ? std::atoi(argv[6]) : props.max_level;
    props.smooth_normals = normals;
    props.chunk_resolution = res;
    props.process_iters = iters;
    const char* kind = argc > 7 ? argv[7] : "fbm";
    Sampler sampler;
    float y_base = 0.0f;
    if (!std::strcmp(kind, "world")) {
        NoiseSamplers::create_sampler_terrain_pert_2d(&sampler);
        NoiseSamplers::apply_properties(&sampler, NoiseSamplers::NoiseSamplerProperties{});
        y_base = -(float)(n_side / 2);
    } else if (!std::strcmp(kind, "pert3d")) {
        NoiseSamplers::create_sampler_terrain_pert_3d(&sampler);
        y_base = -(float)(n_side / 2);
    } else {
        NoiseSamplers::create_sampler_terrain_fbm_3d(&sampler, NoiseSamplers::TerrainProperties{});
    }
    ChunkGenerator gen;
    gen.mode = mode;
    int rc = gen.init(props, sampler);
    if (rc) {
        std::fprintf(stderr, "init: %s\n", bmf_strerror(rc));
        return 1;
    }
    const float size = (float)(res - 1);
    std::vector<std::unique_ptr<DMCChunk>> chunks;
    std::vector<std::unique_ptr<WorldOctreeNode>> nodes;
    std::vector<WorldOctreeNode*> batch;
    for (int i = 0; i < n_side; i++)
        for (int j = 0; j < n_side; j++)
            for (int k = 0; k < n_side; k++) {
                chunks.emplace_back(new DMCChunk());
                chunks.back()->init(vec3(size * i, size * ((float)j + y_base), size * k), size, level, sampler);
                chunks.back()->dim = (uint32_t)res;
                nodes.emplace_back(new WorldOctreeNode());
                nodes.back()->chunk = chunks.back().get();
                nodes.back()->level = level;
                batch.push_back(nodes.back().get());
            }
    rc = gen.process_queue(batch);
    if (rc) {
        std::fprintf(stderr, "process_queue: %s\n", bmf_strerror(rc));
        return 1;
    }
    for (size_t c = 0; c < batch.size(); c++) {
        const DMCChunk* ch = batch[c]->chunk;
        uint64_t h = 0, k = 0;
        size_t nv = ch->vi ? ch->vi->vertices.size() : 0, ni = ch->vi ? ch->vi->mesh_indexes.size() : 0;
        for (size_t v = 0; v < nv; v++) {
            const float* f = &ch->vi->vertices[v].p.x;
            for (int q = 0; q < 3; q++, k++) {
                uint32_t b;
                std::memcpy(&b, &f[q], 4);
                h += (uint64_t)b * (2 * k + 1);
            }
        }
        for (size_t q = 0; q < ni; q++, k++) h += (uint64_t)ch->vi->mesh_indexes[q] * (2 * k + 1);
        h *= 0x9E3779B97F4A7C15ull;
        uint64_t hn = 0, kn = 0;
        for (size_t v = 0; v < nv; v++) {
            const float* f = &ch->vi->vertices[v].n.x;
            for (int q = 0; q < 3; q++, kn++) {
                uint32_t b;
                std::memcpy(&b, &f[q], 4);
                hn += (uint64_t)b * (2 * kn + 1);
            }
        }
        hn *= 0x9E3779B97F4A7C15ull;
        std::printf("{\"chunk\": %zu, \"contains_mesh\": %d, \"V\": %zu, \"I\": %zu, \"stage\": %d, \"checksum\": %llu, "
                    "\"normals\": %llu}\n",
                    c, (int)ch->contains_mesh, nv, ni, batch[c]->generation_stage, (unsigned long long)h,
                    (unsigned long long)hn);
    }
    return 0;
}
