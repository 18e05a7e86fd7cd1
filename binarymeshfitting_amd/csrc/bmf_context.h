// The context object behind the C ABI, shared by the extraction entry
// points (bmf_capi.hip) and the differentiable path (bmf_diff.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/bmf.h"
#include "bmf_internal.h"

namespace bmf {

struct GraphEntry {
    int n = -1;
    int maxk = -1;
    int wclasses = 0;
    hipGraphExec_t exec = nullptr;
};

}  // namespace bmf

struct bmf_ctx {
    int device = 0;
    int cu_count = 256;
    hipStream_t stream = nullptr;
    bmf_limits lim{};
    bmf::Arenas A{};
    bmf_chunk_desc* h_descs[2] = {nullptr, nullptr};  // pinned, double-buffered uploads
    hipEvent_t ev_upload[2] = {nullptr, nullptr};
    int upload_slot = 0;
    hipEvent_t ev_stage[BMF_NUM_STAGES + 1] = {};
    bmf_chunk_result* h_results = nullptr;  // pinned
    bmf::Counters* h_ctr = nullptr;         // pinned
    bmf::SamplerDev h_sampler{};
    float* d_spheres = nullptr;
    float* d_boxes = nullptr;
    bool sampler_set = false;
    int last_n = 0;
    bool have_extract = false;
    bool results_valid = false;
    std::vector<bmf::GraphEntry> graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int persistent_blocks = 2048;  // default workgroups of the persistent kernels
    int stat_nsample = 0, stat_nrefine = 0;  // workgroups of the last pipeline's sampling / refinement
    bool stats_batch = false;                // that pipeline ran the whole batch (BATCHED or profiled)
    bmf_tuning tune{};             // BATCHED launch shapes (bmf_set_tuning / bmf_autotune)
    bmf_mesh_options mesh{0, 0, 0.1f, 0, 0};  // bmf_set_mesh_options
    int64_t* d_fmt = nullptr;  // bmf_format: per-chunk offsets and tile counts [4][max_chunks + 1]
    int64_t* h_fmt = nullptr;  // pinned: formatted vertex and index totals
    // bmf_lod_queue: scratch, and the descriptors it left in the descriptor
    // arena for bmf_extract_queued (queued_n < 0: none)
    int* d_lod_first = nullptr;    // [lod_cap_nodes] lowest list position of a group candidate child
    uint8_t* d_lod_code = nullptr; // [lod_cap_list] pass-A code of each list position
    int lod_cap_nodes = 0, lod_cap_list = 0;
    int* d_lod_counts = nullptr;   // dirty nodes, chunks
    int* h_lod_counts = nullptr;   // pinned
    int queued_n = -1, queued_dim = 0, queued_iters = 0;
    // bmf_stitch scratch: triangles per candidate point, tile totals / offsets,
    // (root, deepest level) and the total
    uint8_t* d_stitch_counts = nullptr;
    int64_t* d_stitch_tiles = nullptr;
    int64_t* d_stitch_misc = nullptr;
    int64_t* h_stitch_total = nullptr;  // pinned
    int64_t stitch_cap_cand = 0;
};

namespace bmf {

// Smoothing launches of a batch: the largest K, at least 1 with normals.
int max_steps(const bmf_ctx* c, const bmf_chunk_desc* chunks, int n);
// Copies n descriptors to the device arena on the context stream.
int upload_descs(bmf_ctx* c, const bmf_chunk_desc* chunks, int n);
// BMF_OK or BMF_EINVAL for one descriptor (host only).
int check_desc(const bmf_chunk_desc& d);
// Validates a sampler descriptor and fills its device view (host only; the
// CSG primitive pointers are left for the caller to set).
int build_sampler(const bmf_sampler_desc* s, SamplerDev* out);
// Destroys the context's captured frames (they hold arena pointers).
void drop_graphs(bmf_ctx* c);

}  // namespace bmf
