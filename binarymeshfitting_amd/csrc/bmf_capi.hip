// C ABI of libbmf.so (declared in include/bmf.h): contexts, device arenas,
// the PLAIN and BATCHED execution paths and stage inspection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/bmf.h"
#include "bmf_context.h"
#include "bmf_internal.h"

using namespace bmf;

namespace {

#define BMF_HIP(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return BMF_EHIP;         \
    } while (0)

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

bool finite(float v) { return std::isfinite(v); }

template <class T>
hipError_t dalloc(T** p, int64_t count) {
    return hipMalloc((void**)p, (size_t)std::max<int64_t>(count, 1) * sizeof(T));
}

void free_arenas(bmf_ctx* c) {
    Arenas& A = c->A;
    void* ptrs[] = {A.meta,    A.descs,     A.results,  A.tile_prefix, A.plan_part, A.plan_tmp, A.stat_blocks, A.plane_map, A.smooth_list, A.smooth_prefix,
                    A.smooth_tprefix, A.smooth_xcd_v, A.smooth_xcd_t, A.smooth_vmap, A.smooth_tmap, A.refine, A.plane_refine_n, A.cen4,
                    A.dens,    A.labels,    A.exact,    A.row_v,       A.row_t,        A.cell_tb,   A.cell_mask,
                    A.pos,     A.pos4,      A.boundary, A.vedge,       A.vring,        A.idx,          A.ctr,
                    A.sampler, c->d_spheres, c->d_boxes, A.ring, A.ring_n, A.qef_err, A.nrm4, A.fn4, A.nrm,
                    A.csg_list};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    A = Arenas{};
    c->d_spheres = c->d_boxes = nullptr;
}

// The whole device pipeline for chunks [0, n) of the uploaded descriptors.
// With `ev` (BMF_NUM_STAGES + 1 events) the stage boundaries are recorded.
// Workgroups of the batched row scan: one per chunk up to two per CU; the
// kernel strides over the batch and skips the chunks that were not sampled,
// so a large batch with few sampled chunks does not dispatch a workgroup of
// 1024 threads per chunk.
int scan_blocks(const bmf_ctx* c, int n) { return std::min(n, 2 * c->cu_count); }

int enqueue_pipeline(bmf_ctx* c, const Arenas& A, int n, int maxk, int wclasses, int allow_cull, int reset,
                     int sample_blocks, int row_blocks, int scan_blocks, int smooth_blocks, hipEvent_t* ev = nullptr) {
    hipStream_t s = c->stream;
#define BMF_MARK(i) \
    if (ev) BMF_HIP(hipEventRecord(ev[i], s))
    BMF_MARK(0);
    BMF_HIP(launch_plan(A, n, allow_cull, reset, s));
    BMF_MARK(1);
    BMF_HIP(launch_sample(A, n, sample_blocks, 1, c->h_sampler.kind, c->h_sampler.noise_dims == 2, s));
    c->stat_nsample = sample_blocks;
    c->stat_nrefine = row_blocks;
    BMF_MARK(2);
    const bool quads = A.mesh_mode == BMF_MESH_QUADS;
    // the row-word classes of a mixed batch one after another on the one
    // stream: a context is one stream (one hardware queue), so several
    // contexts on a GPU overlap their batches
    BMF_HIP(launch_count(A, n, row_blocks, wclasses, s));
    BMF_HIP(launch_refine(A, n, row_blocks, s));
    BMF_MARK(3);
    BMF_HIP(launch_scan(A, n, scan_blocks, s));
    BMF_MARK(4);
    if (quads) BMF_HIP(launch_dmc_emit(A, n, row_blocks, wclasses, s));
    else BMF_HIP(launch_emit(A, n, row_blocks, wclasses, s));
    BMF_MARK(5);
    for (int m = 0; m < maxk; m++) {
        if (quads) BMF_HIP(launch_dmc_smooth(A, m, smooth_blocks, s));
        else BMF_HIP(launch_smooth(A, m, smooth_blocks, s));
    }
    BMF_MARK(6);
    BMF_MARK(7);
#undef BMF_MARK
    return BMF_OK;
}

}  // namespace

namespace bmf {

// Host descriptors -> device through one of two pinned staging buffers; a
// buffer is refilled only after its previous copy has executed, so frames
// can be enqueued back to back without a host synchronisation.
int upload_descs(bmf_ctx* c, const bmf_chunk_desc* chunks, int n) {
    c->queued_n = -1;  // the arena no longer holds bmf_lod_queue's descriptors
    const int slot = c->upload_slot;
    c->upload_slot ^= 1;
    BMF_HIP(hipEventSynchronize(c->ev_upload[slot]));
    std::memcpy(c->h_descs[slot], chunks, sizeof(bmf_chunk_desc) * (size_t)n);
    BMF_HIP(hipMemcpyAsync(c->A.descs, c->h_descs[slot], sizeof(bmf_chunk_desc) * (size_t)n,
                           hipMemcpyHostToDevice, c->stream));
    BMF_HIP(hipEventRecord(c->ev_upload[slot], c->stream));
    return BMF_OK;
}

int max_steps(const bmf_ctx* c, const bmf_chunk_desc* chunks, int n) {
    int k = 0;
    for (int i = 0; i < n; i++) k = std::max(k, chunks[i].iters);
    if (c->A.normals && n > 0) k = std::max(k, 1);
    return k;
}

int check_desc(const bmf_chunk_desc& d) {
    if (d.dim < 2 || d.dim > BMF_MAX_DIM) return BMF_EINVAL;
    if (!finite(d.size) || !(d.size > 0.0f)) return BMF_EINVAL;
    if (!finite(d.overlap) || d.overlap < 0.0f || d.overlap > 1.0f) return BMF_EINVAL;
    if (!finite(d.pos[0]) || !finite(d.pos[1]) || !finite(d.pos[2])) return BMF_EINVAL;
    if (d.iters < 0 || d.iters > 64) return BMF_EINVAL;
    if (d.boundary_processing != 0 && d.boundary_processing != 1) return BMF_EINVAL;
    return BMF_OK;
}

int build_sampler(const bmf_sampler_desc* s, SamplerDev* out) {
    if (!s || !out) return BMF_EINVAL;
    if (s->kind < BMF_SAMPLER_TERRAIN || s->kind > BMF_SAMPLER_NOISE) return BMF_EINVAL;
    if (s->kind == BMF_SAMPLER_EXTERNAL && !s->densities) return BMF_EINVAL;
    if (s->kind == BMF_SAMPLER_CSG &&
        (s->n_spheres < 0 || s->n_boxes < 0 || (s->n_spheres && !s->spheres) || (s->n_boxes && !s->boxes)))
        return BMF_EINVAL;
    SamplerDev d{};
    d.kind = s->kind;
    d.world_size = s->world_size;
    d.amp = s->amp;
    d.y0 = s->y0;
    d.seed = s->seed;
    if (s->kind == BMF_SAMPLER_TERRAIN || s->kind == BMF_SAMPLER_NOISE) {
        if (s->octaves < 1 || s->octaves > BMF_MAX_OCTAVES) return BMF_EINVAL;
        if (!finite(s->amp) || !finite(s->y0) || !finite(s->freq) || !finite(s->lacunarity) || !finite(s->gain))
            return BMF_EINVAL;
        d.octaves = s->octaves;
        // per-octave constants: FP32 product chains, normaliser = sum of amplitudes
        float f = s->freq, a = 1.0f, norm = 0.0f;
        for (int o = 0; o < s->octaves; o++) {
            d.fo[o] = f;
            d.ao[o] = a;
            norm = norm + a;
            f = f * s->lacunarity;
            a = a * s->gain;
        }
        if (!(norm > 0.0f) || !finite(norm)) return BMF_EINVAL;
        d.inv_norm = 1.0f / norm;
        // adaptive-octave bounds (k_sample decide_label)
        double total = 0.0;
        for (int o = 0; o < s->octaves; o++) total += std::fabs((double)d.ao[o]);
        const double ulp_total = total * std::ldexp(1.0, -23);
        for (int k = 0; k <= s->octaves; k++) {
            double r = 0.0;
            for (int o = k; o < s->octaves; o++) r += std::fabs((double)d.ao[o]);
            r += (double)(s->octaves - k + 2) * ulp_total;
            float rf = (float)r;
            if ((double)rf < r) rf = std::nextafter(rf, INFINITY);
            d.reach[k] = rf;
        }
    }
    if (s->kind == BMF_SAMPLER_NOISE) {
        const bool dims_ok = s->noise_dims == 2 || s->noise_dims == 3;
        const bool kinds_ok = (s->noise_basis == BMF_NOISE_VALUE || s->noise_basis == BMF_NOISE_SIMPLEX) &&
                              (s->noise_fractal == BMF_FRACTAL_FBM || s->noise_fractal == BMF_FRACTAL_RIGID) &&
                              s->perturb >= BMF_PERTURB_NONE && s->perturb <= BMF_PERTURB_OFFSET_XZ;
        if (!dims_ok || !kinds_ok) return BMF_EINVAL;
        if (s->perturb == BMF_PERTURB_OFFSET_XZ && s->noise_dims != 3) return BMF_EINVAL;
        const int po = s->perturb == BMF_PERTURB_NONE ? 0 : s->perturb_octaves;
        if (po < 0 || po > BMF_MAX_PERTURB_OCTAVES) return BMF_EINVAL;
        if (!finite(s->g_scale) || !finite(s->y_scale) || !finite(s->height) || !finite(s->perturb_amp) ||
            !finite(s->perturb_freq) || !finite(s->perturb_gain))
            return BMF_EINVAL;
        d.noise_dims = s->noise_dims;
        d.noise_basis = s->noise_basis;
        d.noise_fractal = s->noise_fractal;
        d.g_scale = s->g_scale;
        d.y_scale = s->y_scale;
        d.height = s->height;
        d.perturb = s->perturb;
        d.perturb_octaves = po;
        d.perturb_amp = s->perturb_amp;
        // perturbation constants: FP32 chains (frequency doubles per octave)
        float f = s->perturb_freq, a = s->perturb == BMF_PERTURB_GRADIENT ? s->perturb_amp : 1.0f, norm = 0.0f;
        for (int k = 0; k < po; k++) {
            d.pf[k] = f;
            d.pa[k] = a;
            norm = norm + a;
            f = f * 2.0f;
            a = a * s->perturb_gain;
        }
        d.pinv = norm > 0.0f ? 1.0f / norm : 0.0f;
        if (s->perturb == BMF_PERTURB_OFFSET_XZ && po > 0 && !(norm > 0.0f && finite(norm))) return BMF_EINVAL;
        // bounds of n for the chunk culling: FBM is clamped to [-1, 1]; the
        // rigid sum t_0 - sum a_o t_o (t in [0, 1]) is smallest, step by step
        // in FP32 (every fma monotone), at t_0 = 0 and t_o = 1 when a_o >= 0
        d.cull = 1;
        if (s->noise_fractal == BMF_FRACTAL_FBM) {
            d.nlo = -1.0f;
            d.nhi = 1.0f;
        } else {
            float r = 0.0f;
            for (int o = 1; o < s->octaves; o++) {
                if (!(d.ao[o] >= 0.0f)) d.cull = 0;
                r = std::fmaf(-1.0f, d.ao[o], r);
            }
            d.nlo = r;
            d.nhi = 1.0f;
        }
    }
    if (s->kind == BMF_SAMPLER_CSG) {
        d.n_spheres = s->n_spheres;
        d.n_boxes = s->n_boxes;
    }
    if (s->kind == BMF_SAMPLER_EXTERNAL) d.ext = s->densities;
    *out = d;
    return BMF_OK;
}

void drop_graphs(bmf_ctx* c) {
    for (auto& g : c->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

}  // namespace bmf

extern "C" {

const char* bmf_strerror(int code) {
    switch (code) {
        case BMF_OK: return "ok";
        case BMF_EINVAL: return "invalid argument";
        case BMF_ECAPACITY: return "arena capacity exceeded";
        case BMF_EHIP: return "HIP runtime error";
        case BMF_ENODEV: return "no usable GPU";
        case BMF_ESTATE: return "call out of order";
    }
    return "unknown error";
}

const char* bmf_version(void) { return "bmf 0.1 (gfx950)"; }

int bmf_validate_chunks(const bmf_chunk_desc* chunks, int n, const bmf_limits* lim) {
    if (n < 0 || (n > 0 && !chunks)) return BMF_EINVAL;
    if (lim && n > lim->max_chunks) return BMF_EINVAL;
    for (int i = 0; i < n; i++) {
        int rc = check_desc(chunks[i]);
        if (rc) return rc;
    }
    return BMF_OK;
}

int bmf_create(bmf_ctx** out, int device, const bmf_limits* lim) {
    if (!out || !lim) return BMF_EINVAL;
    *out = nullptr;
    if (lim->max_chunks < 1 || lim->max_chunks > (1 << 20) || lim->max_samples < 8 || lim->max_vertices < 1 ||
        lim->max_indices < 3)
        return BMF_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BMF_ENODEV;
    if (device < 0 || device >= ndev) return BMF_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return BMF_EHIP;
    bmf_ctx* c = new (std::nothrow) bmf_ctx();
    if (!c) return BMF_EHIP;
    c->device = device;
    c->lim = *lim;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->cu_count = prop.multiProcessorCount;
    c->persistent_blocks = c->cu_count * 8;
    c->tune.sample_tile = kSampleTileDefault;
    c->tune.sample_blocks = c->persistent_blocks;
    // ten per CU: the best of 4..20 per CU for count + emit on a batch of
    // 512 64^3-cell chunks (fewer leave workgroups with a third plane item,
    // more only add empty workgroups); the autotuner tries the others
    c->tune.row_blocks = c->cu_count * 10;
    c->tune.smooth_blocks = 4 * c->persistent_blocks;
    c->tune.row_split = -1;
    Arenas& A = c->A;
    const int64_t nc = lim->max_chunks;
    A.cap_chunks = (int32_t)nc;
    A.cap_samples = lim->max_samples;
    A.cap_rows = lim->max_samples / 2 + nc;
    A.cap_vertices = lim->max_vertices;
    A.cap_indices = lim->max_indices;
    A.sample_tile = kSampleTileDefault;
    A.smoothing_follows = 1;
    A.stat_cap = 64 * c->cu_count;
    A.row_split = -1;
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && dalloc(&A.meta, nc) == hipSuccess && dalloc(&A.descs, nc) == hipSuccess &&
         dalloc(&A.results, nc) == hipSuccess && dalloc(&A.tile_prefix, nc + 1) == hipSuccess &&
         dalloc(&A.plan_part, 40 * ((nc + 1023) / 1024) + 64) == hipSuccess &&
         dalloc(&A.plan_tmp, 2 * nc) == hipSuccess &&
         dalloc(&A.stat_blocks, 2 * (int64_t)(64 * c->cu_count)) == hipSuccess &&
         dalloc(&A.plane_map, nc * BMF_MAX_DIM) == hipSuccess && dalloc(&A.smooth_list, nc) == hipSuccess &&
         dalloc(&A.smooth_prefix, nc + 1) == hipSuccess && dalloc(&A.smooth_tprefix, nc + 1) == hipSuccess &&
         dalloc(&A.smooth_xcd_v, kXcds + 1) == hipSuccess && dalloc(&A.smooth_xcd_t, kXcds + 1) == hipSuccess &&
         dalloc(&A.smooth_vmap, A.cap_vertices / kSmoothThreads + nc + 1) == hipSuccess &&
         dalloc(&A.smooth_tmap, A.cap_indices / 3 / kSmoothThreads + nc + 1) == hipSuccess &&
         dalloc(&A.refine, A.cap_samples) == hipSuccess && dalloc(&A.plane_refine_n, nc * BMF_MAX_DIM * kMaxPlaneSplit) == hipSuccess &&
         dalloc(&A.cen4, A.cap_indices / 3 + 1) == hipSuccess &&
         dalloc(&A.dens, A.cap_samples) == hipSuccess && dalloc(&A.labels, A.cap_samples / 64 + nc) == hipSuccess &&
         dalloc(&A.exact, A.cap_samples / 64 + nc) == hipSuccess &&
         dalloc(&A.row_v, A.cap_rows) == hipSuccess && dalloc(&A.row_t, A.cap_rows) == hipSuccess &&
         dalloc(&A.cell_tb, A.cap_samples) == hipSuccess && dalloc(&A.cell_mask, A.cap_samples) == hipSuccess &&
         dalloc(&A.pos, 3 * A.cap_vertices) == hipSuccess && dalloc(&A.pos4, A.cap_vertices) == hipSuccess &&
         dalloc(&A.boundary, A.cap_vertices) == hipSuccess && dalloc(&A.vedge, A.cap_vertices) == hipSuccess &&
         dalloc(&A.vring, A.cap_vertices) == hipSuccess &&
         dalloc(&A.idx, A.cap_indices) == hipSuccess && dalloc(&A.ctr, 1) == hipSuccess &&
         dalloc(&A.sampler, 1) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&c->h_descs[0], sizeof(bmf_chunk_desc) * nc) == hipSuccess &&
         hipHostMalloc((void**)&c->h_descs[1], sizeof(bmf_chunk_desc) * nc) == hipSuccess &&
         hipHostMalloc((void**)&c->h_results, sizeof(bmf_chunk_result) * nc) == hipSuccess &&
         hipHostMalloc((void**)&c->h_ctr, sizeof(Counters)) == hipSuccess;
    ok = ok && hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
         hipEventCreate(&c->ev_upload[0]) == hipSuccess && hipEventCreate(&c->ev_upload[1]) == hipSuccess;
    for (int i = 0; ok && i <= BMF_NUM_STAGES; i++) ok = hipEventCreate(&c->ev_stage[i]) == hipSuccess;
    if (ok) {
        // counters start defined (plan resets them before use)
        std::memset(c->h_ctr, 0, sizeof(Counters));
        ok = hipMemcpy(A.ctr, c->h_ctr, sizeof(Counters), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        bmf_destroy(c);
        return BMF_EHIP;
    }
    *out = c;
    return BMF_OK;
}

void bmf_destroy(bmf_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    free_arenas(c);
    for (int i = 0; i < 2; i++)
        if (c->h_descs[i]) (void)hipHostFree(c->h_descs[i]);
    if (c->h_results) (void)hipHostFree(c->h_results);
    if (c->h_ctr) (void)hipHostFree(c->h_ctr);
    if (c->d_fmt) (void)hipFree(c->d_fmt);
    if (c->h_fmt) (void)hipHostFree(c->h_fmt);
    if (c->d_lod_first) (void)hipFree(c->d_lod_first);
    if (c->d_lod_code) (void)hipFree(c->d_lod_code);
    if (c->d_lod_counts) (void)hipFree(c->d_lod_counts);
    if (c->h_lod_counts) (void)hipHostFree(c->h_lod_counts);
    if (c->d_stitch_counts) (void)hipFree(c->d_stitch_counts);
    if (c->d_stitch_tiles) (void)hipFree(c->d_stitch_tiles);
    if (c->d_stitch_misc) (void)hipFree(c->d_stitch_misc);
    if (c->h_stitch_total) (void)hipHostFree(c->h_stitch_total);
    hipEvent_t evs[] = {c->ev0, c->ev1, c->ev_upload[0], c->ev_upload[1]};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i <= BMF_NUM_STAGES; i++)
        if (c->ev_stage[i]) (void)hipEventDestroy(c->ev_stage[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int bmf_noise_preset(int32_t which, float world_size, uint32_t seed, bmf_sampler_desc* out) {
    if (!out || which < BMF_PRESET_TERRAIN2D || which > BMF_PRESET_WORLD) return BMF_EINVAL;
    bmf_sampler_desc d{};
    d.kind = BMF_SAMPLER_NOISE;
    d.world_size = world_size;
    d.seed = seed;
    // FastNoiseSIMD's defaults, which the block functions keep unless they set them
    const float freq = 0.01f;
    d.freq = freq;
    d.lacunarity = 2.0f;
    d.gain = 0.5f;
    d.octaves = 3;
    d.noise_dims = 3;
    d.noise_basis = BMF_NOISE_VALUE;
    d.noise_fractal = BMF_FRACTAL_FBM;
    d.g_scale = 1.0f;
    d.y_scale = 1.0f;
    d.height = 1.0f;
    d.perturb = BMF_PERTURB_NONE;
    d.perturb_gain = 0.5f;
    auto gradient = [&](int octaves, float amp, float f, float gain) {
        d.perturb = BMF_PERTURB_GRADIENT;
        d.perturb_octaves = octaves;
        d.perturb_amp = amp / freq;
        d.perturb_freq = f * freq;
        d.perturb_gain = gain;
    };
    switch (which) {
        case BMF_PRESET_TERRAIN2D:  // NoiseSampler.cpp:115-125, 139
            d.noise_dims = 2;
            d.octaves = 12;
            d.height = 64.0f;
            break;
        case BMF_PRESET_TERRAIN2D_PERT:  // :150-166 without properties
            d.noise_dims = 2;
            d.g_scale = 0.25f;
            d.height = 64.0f;
            gradient(20, 1.0f, 0.5f, 0.5f);
            break;
        case BMF_PRESET_WORLD:  // WorldOctree.cpp:47-54 into terrain2d_pert_block
            d.noise_dims = 2;
            d.g_scale = 0.25f;
            d.height = 75.0f;
            gradient(13, 0.87f, 0.585f, 0.488f);
            break;
        case BMF_PRESET_TERRAIN3D:  // :198-207, 219: rigid value noise, dy scaled by ym
            d.octaves = 4;
            d.noise_fractal = BMF_FRACTAL_RIGID;
            d.g_scale = 0.15f;
            d.y_scale = 0.5f;
            break;
        case BMF_PRESET_TERRAIN3D_PERT:  // :231-243, 255
            d.octaves = 8;
            d.noise_basis = BMF_NOISE_SIMPLEX;
            d.noise_fractal = BMF_FRACTAL_RIGID;
            d.g_scale = 0.15f;
            d.height = 48.0f;
            gradient(3, 1.0f, 0.05f, 0.5f);
            break;
        case BMF_PRESET_WINDY3D:  // :267-305, 313: offsets from a 1-octave field at p * 0.01
            d.octaves = 4;
            d.noise_basis = BMF_NOISE_SIMPLEX;
            d.g_scale = 0.25f;
            d.y_scale = 0.5f;
            d.height = world_size * 0.5f;
            d.perturb = BMF_PERTURB_OFFSET_XZ;
            d.perturb_octaves = 1;
            d.perturb_amp = 1.0f;
            d.perturb_freq = 0.01f * freq;
            break;
    }
    *out = d;
    return BMF_OK;
}

int bmf_set_sampler(bmf_ctx* c, const bmf_sampler_desc* s) {
    if (!c || !s) return BMF_EINVAL;
    SamplerDev d{};
    int rc = build_sampler(s, &d);
    if (rc) return rc;
    BMF_HIP(hipSetDevice(c->device));
    // the new CSG buffers are allocated and filled before anything of the
    // current sampler is released: a call that fails leaves it in place
    float* spheres = nullptr;
    float* boxes = nullptr;
    uint16_t* list = nullptr;
    int32_t stride = 0;
    auto release_new = [&]() {
        if (spheres) (void)hipFree(spheres);
        if (boxes) (void)hipFree(boxes);
        if (list) (void)hipFree(list);
    };
    if (s->kind == BMF_SAMPLER_CSG) {
        bool ok = dalloc(&spheres, 4 * (int64_t)s->n_spheres) == hipSuccess &&
                  dalloc(&boxes, 6 * (int64_t)s->n_boxes) == hipSuccess;
        if (ok && s->n_spheres)
            ok = hipMemcpy(spheres, s->spheres, sizeof(float) * 4 * s->n_spheres, hipMemcpyHostToDevice) == hipSuccess;
        if (ok && s->n_boxes)
            ok = hipMemcpy(boxes, s->boxes, sizeof(float) * 6 * s->n_boxes, hipMemcpyHostToDevice) == hipSuccess;
        // per-chunk culled primitive lists (u16 ids; without them every
        // sample evaluates every primitive)
        const int64_t np = (int64_t)s->n_spheres + s->n_boxes;
        if (ok && np > 0 && s->n_spheres < 65536 && s->n_boxes < 65536) {
            ok = dalloc(&list, np * c->A.cap_chunks) == hipSuccess;
            stride = (int32_t)np;
        }
        if (!ok) {
            release_new();
            return BMF_EHIP;
        }
        d.spheres = spheres;
        d.boxes = boxes;
    }
    // queued work may still read the current sampler's buffers
    if (hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(c->A.sampler, &d, sizeof(SamplerDev), hipMemcpyHostToDevice) != hipSuccess) {
        // the device copy of the sampler is now unknown: no extraction until
        // a later call succeeds
        release_new();
        c->sampler_set = false;
        drop_graphs(c);
        return BMF_EHIP;
    }
    if (c->d_spheres) (void)hipFree(c->d_spheres);
    if (c->d_boxes) (void)hipFree(c->d_boxes);
    if (c->A.csg_list) (void)hipFree(c->A.csg_list);
    c->d_spheres = spheres;
    c->d_boxes = boxes;
    c->A.csg_list = list;
    c->A.csg_stride = stride;
    c->h_sampler = d;
    // captured frames hold the arena pointers (the CSG lists among them)
    drop_graphs(c);
    c->sampler_set = true;
    return BMF_OK;
}

// The extraction of n descriptors: `chunks` (host) uploaded first, or, when
// null, the n descriptors already in the arena, all with uni's dim and iters.
static int extract_impl(bmf_ctx* c, const bmf_chunk_desc* chunks, int n, int mode, const bmf_chunk_desc* uni) {
    BMF_HIP(hipSetDevice(c->device));
    c->results_valid = false;
    c->last_n = n;
    c->have_extract = true;
    c->stats_batch = mode == BMF_MODE_BATCHED;  // PLAIN frames run one pipeline per chunk
    if (n == 0) return BMF_OK;
    int rc = BMF_OK;
    if (chunks) {
        rc = upload_descs(c, chunks, n);
        if (rc) return rc;
    }
    auto desc = [&](int j) -> const bmf_chunk_desc& { return chunks ? chunks[j] : *uni; };
    int maxk = chunks ? max_steps(c, chunks, n) : max_steps(c, uni, 1), wclasses = 0;
    for (int i = 0; i < (chunks ? n : 1); i++) wclasses |= row_word_class(desc(i).dim);
    const Arenas& A = c->A;
    const int pb = c->persistent_blocks;
    if (mode == BMF_MODE_PLAIN) {
        // one launch sequence per chunk; scratch arenas reused from offset 0,
        // output arenas bump-allocated across the chunks (ChunkGenerator.cpp:91
        // loop body, one chunk at a time)
        BMF_HIP(hipEventRecord(c->ev0, c->stream));
        int64_t ext_base = 0;
        for (int j = 0; j < n; j++) {
            const bmf_chunk_desc& dj = desc(j);
            Arenas Aj = A;
            Aj.ext_base = ext_base;
            ext_base += (int64_t)dj.dim * dj.dim * dj.dim;
            Aj.descs = A.descs + j;
            Aj.meta = A.meta + j;
            Aj.results = A.results + j;
            const int64_t dim = dj.dim;
            const int64_t ns = dim * dim * dim;
            const int sb = (int)std::min<int64_t>(ceil_div(ns, A.sample_tile), pb);
            rc = enqueue_pipeline(c, Aj, 1, max_steps(c, &dj, 1), row_word_class(dj.dim), 1, j == 0, sb,
                                  (int)dim * kMaxPlaneSplit, 1,
                                  std::min(pb, 1024));
            if (rc) return rc;
        }
        BMF_HIP(hipEventRecord(c->ev1, c->stream));
        return BMF_OK;
    }
    // BATCHED: persistent kernels over the whole batch, captured once per
    // (n, max iterations) in a hipGraph and replayed.
    GraphEntry* ge = nullptr;
    for (auto& g : c->graphs)
        if (g.n == n && g.maxk == maxk && g.wclasses == wclasses) ge = &g;
    if (!ge) {
        hipGraph_t graph = nullptr;
        BMF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_pipeline(c, A, n, maxk, wclasses, 1, 1, c->tune.sample_blocks, c->tune.row_blocks,
                              scan_blocks(c, n), c->tune.smooth_blocks);
        hipGraph_t captured = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &captured);
        if (rc) {
            if (captured) (void)hipGraphDestroy(captured);
            return rc;
        }
        if (e != hipSuccess) return BMF_EHIP;
        graph = captured;
        GraphEntry g;
        g.n = n;
        g.maxk = maxk;
        g.wclasses = wclasses;
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return BMF_EHIP;
        c->graphs.push_back(g);
        ge = &c->graphs.back();
    }
    BMF_HIP(hipEventRecord(c->ev0, c->stream));
    BMF_HIP(hipGraphLaunch(ge->exec, c->stream));
    BMF_HIP(hipEventRecord(c->ev1, c->stream));
    return BMF_OK;
}

int bmf_extract(bmf_ctx* c, const bmf_chunk_desc* chunks, int n, int mode) {
    if (!c) return BMF_EINVAL;
    if (!c->sampler_set) return BMF_ESTATE;
    if (mode != BMF_MODE_PLAIN && mode != BMF_MODE_BATCHED) return BMF_EINVAL;
    int rc = bmf_validate_chunks(chunks, n, &c->lim);
    if (rc) return rc;
    return extract_impl(c, chunks, n, mode, nullptr);
}

int bmf_extract_queued(bmf_ctx* c, int mode) {
    if (!c) return BMF_EINVAL;
    if (!c->sampler_set || c->queued_n < 0) return BMF_ESTATE;
    if (mode != BMF_MODE_PLAIN && mode != BMF_MODE_BATCHED) return BMF_EINVAL;
    bmf_chunk_desc uni{};
    uni.dim = c->queued_dim;
    uni.iters = c->queued_iters;
    return extract_impl(c, nullptr, c->queued_n, mode, &uni);
}

int bmf_sync(bmf_ctx* c) {
    if (!c) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(c->device));
    BMF_HIP(hipStreamSynchronize(c->stream));
    return BMF_OK;
}

static int fetch_results(bmf_ctx* c) {
    if (c->results_valid) return BMF_OK;
    if (!c->have_extract) return BMF_ESTATE;
    const int n = c->last_n;
    if (n > 0)
        BMF_HIP(hipMemcpyAsync(c->h_results, c->A.results, sizeof(bmf_chunk_result) * n, hipMemcpyDeviceToHost,
                               c->stream));
    BMF_HIP(hipMemcpyAsync(c->h_ctr, c->A.ctr, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    BMF_HIP(hipStreamSynchronize(c->stream));
    c->results_valid = true;
    return BMF_OK;
}

int bmf_result(bmf_ctx* c, bmf_chunk_result* out, int n, int64_t* need_v, int64_t* need_i) {
    if (!c) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(c->device));
    int rc = fetch_results(c);
    if (rc) return rc;
    if (out) {
        if (n < 0 || n > c->last_n) return BMF_EINVAL;
        std::memcpy(out, c->h_results, sizeof(bmf_chunk_result) * (size_t)n);
    }
    if (need_v) *need_v = (int64_t)c->h_ctr->need_v;
    if (need_i) *need_i = (int64_t)c->h_ctr->need_i;
    if (c->h_ctr->error) return BMF_EHIP;
    return c->h_ctr->overflow ? BMF_ECAPACITY : BMF_OK;
}

int bmf_mesh_view_get(bmf_ctx* c, bmf_mesh_view* v) {
    if (!c || !v) return BMF_EINVAL;
    int rc = fetch_results(c);
    if (rc) return rc;
    v->positions = c->A.pos;
    v->indices = c->A.idx;
    v->boundary = c->A.boundary;
    v->normals = c->A.normals ? c->A.nrm : nullptr;
    v->total_vertices = std::min<int64_t>((int64_t)c->h_ctr->v_bump, c->A.cap_vertices);
    v->total_indices = std::min<int64_t>((int64_t)c->h_ctr->i_bump, c->A.cap_indices);
    return BMF_OK;
}

int bmf_mesh_to_host(bmf_ctx* c, int chunk, float* positions, uint32_t* indices, uint8_t* boundary, int64_t cap_v,
                     int64_t cap_i, int64_t* nv, int64_t* ni) {
    if (!c) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(c->device));
    int rc = fetch_results(c);
    if (rc) return rc;
    if (chunk < 0 || chunk >= c->last_n) return BMF_EINVAL;
    const bmf_chunk_result& r = c->h_results[chunk];
    if (nv) *nv = r.vertex_count;
    if (ni) *ni = r.index_count;
    if (r.status != BMF_OK) return r.status;
    if (cap_v < r.vertex_count || cap_i < r.index_count) return BMF_ECAPACITY;
    if (r.vertex_count) {
        if (positions)
            BMF_HIP(hipMemcpyAsync(positions, c->A.pos + 3 * r.vertex_offset, sizeof(float) * 3 * r.vertex_count,
                                   hipMemcpyDeviceToHost, c->stream));
        if (boundary)
            BMF_HIP(hipMemcpyAsync(boundary, c->A.boundary + r.vertex_offset, r.vertex_count, hipMemcpyDeviceToHost,
                                   c->stream));
    }
    if (r.index_count && indices)
        BMF_HIP(hipMemcpyAsync(indices, c->A.idx + r.index_offset, sizeof(uint32_t) * r.index_count,
                               hipMemcpyDeviceToHost, c->stream));
    BMF_HIP(hipStreamSynchronize(c->stream));
    return BMF_OK;
}

// Runs plan + sample (+ count/scan for the flag) for one chunk with culling
// off, so the density block and labels are materialised.
static int stage_one(bmf_ctx* c, const bmf_chunk_desc* d, bool scan) {
    if (!c || !d) return BMF_EINVAL;
    if (!c->sampler_set) return BMF_ESTATE;
    int rc = check_desc(*d);
    if (rc) return rc;
    const int64_t ns = (int64_t)d->dim * d->dim * d->dim;
    if (ns > c->A.cap_samples) return BMF_ECAPACITY;
    BMF_HIP(hipSetDevice(c->device));
    BMF_HIP(hipStreamSynchronize(c->stream));
    c->have_extract = false;
    c->results_valid = false;
    rc = upload_descs(c, d, 1);
    if (rc) return rc;
    const Arenas& A = c->A;
    BMF_HIP(launch_plan(A, 1, 0, 1, c->stream));
    BMF_HIP(launch_sample(A, 1, (int)std::min<int64_t>(ceil_div(ns, A.sample_tile), c->persistent_blocks), 0,
                          c->h_sampler.kind, c->h_sampler.noise_dims == 2, c->stream));
    if (scan) {
        BMF_HIP(launch_count(A, 1, d->dim, row_word_class(d->dim), c->stream));
        BMF_HIP(launch_scan(A, 1, 1, c->stream));
    }
    return BMF_OK;
}

int bmf_sample_block(bmf_ctx* c, const bmf_chunk_desc* d, float* dens) {
    int rc = stage_one(c, d, false);
    if (rc) return rc;
    if (!dens) return BMF_EINVAL;
    const int64_t ns = (int64_t)d->dim * d->dim * d->dim;
    BMF_HIP(hipMemcpyAsync(dens, c->A.dens, sizeof(float) * ns, hipMemcpyDeviceToHost, c->stream));
    BMF_HIP(hipStreamSynchronize(c->stream));
    return BMF_OK;
}

int bmf_label_block(bmf_ctx* c, const bmf_chunk_desc* d, uint32_t* words, int32_t* contains) {
    int rc = stage_one(c, d, true);
    if (rc) return rc;
    if (!words) return BMF_EINVAL;
    const int64_t nw = (int64_t)d->dim * d->dim * ((d->dim + 31) / 32);
    uint32_t* tmp = nullptr;
    BMF_HIP(dalloc(&tmp, nw));
    hipError_t e = launch_ref_labels(c->A, tmp, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(words, tmp, sizeof(uint32_t) * nw, hipMemcpyDeviceToHost, c->stream);
    ChunkMeta m{};
    if (e == hipSuccess) e = hipMemcpyAsync(&m, c->A.meta, sizeof(ChunkMeta), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return BMF_EHIP;
    if (contains) *contains = m.contains_mesh;
    return BMF_OK;
}

int bmf_cell_masks(bmf_ctx* c, const bmf_chunk_desc* d, uint64_t* masks) {
    int rc = stage_one(c, d, false);
    if (rc) return rc;
    if (!masks) return BMF_EINVAL;
    const int64_t nw = (int64_t)d->dim * d->dim * ((d->dim + 7) / 8);
    uint64_t* tmp = nullptr;
    BMF_HIP(dalloc(&tmp, nw));
    hipError_t e = launch_ref_masks(c->A, tmp, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(masks, tmp, sizeof(uint64_t) * nw, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    return e == hipSuccess ? BMF_OK : BMF_EHIP;
}

int bmf_last_timing(bmf_ctx* c, float* total_ms) {
    if (!c || !total_ms) return BMF_EINVAL;
    if (!c->have_extract) return BMF_ESTATE;
    BMF_HIP(hipSetDevice(c->device));
    BMF_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    BMF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *total_ms = ms;
    return BMF_OK;
}

int bmf_get_frame_stats(bmf_ctx* c, bmf_frame_stats* out) {
    if (!c || !out) return BMF_EINVAL;
    if (!c->results_valid || !c->stats_batch) return BMF_ESTATE;
    out->samples = c->h_ctr->total_samples;
    // per-workgroup counts of the last pipeline's two kernels (plain stores there)
    const int ns = std::min(c->stat_nsample, c->A.stat_cap), nr = std::min(c->stat_nrefine, c->A.stat_cap);
    std::vector<unsigned long long> h((size_t)std::max(ns, nr));
    BMF_HIP(hipSetDevice(c->device));
    BMF_HIP(hipStreamSynchronize(c->stream));
    int64_t sum[2] = {0, 0};
    for (int k = 0; k < 2; k++) {
        const int m = k == 0 ? ns : nr;
        if (m <= 0) continue;
        BMF_HIP(hipMemcpy(h.data(), c->A.stat_blocks + (size_t)k * c->A.stat_cap, sizeof(unsigned long long) * m,
                          hipMemcpyDeviceToHost));
        for (int b = 0; b < m; b++) sum[k] += (int64_t)h[b];
    }
    out->octave_evals = sum[0];
    out->refine_evals = sum[1];
    return BMF_OK;
}

int bmf_profile_stages(bmf_ctx* c, const bmf_chunk_desc* chunks, int n, float* stage_ms) {
    if (!c || !stage_ms) return BMF_EINVAL;
    if (!c->sampler_set) return BMF_ESTATE;
    int rc = bmf_validate_chunks(chunks, n, &c->lim);
    if (rc) return rc;
    if (n == 0) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(c->device));
    c->results_valid = false;
    c->last_n = n;
    c->have_extract = true;
    rc = upload_descs(c, chunks, n);
    if (rc) return rc;
    int maxk = max_steps(c, chunks, n), wclasses = 0;
    for (int i = 0; i < n; i++) wclasses |= row_word_class(chunks[i].dim);
    BMF_HIP(hipEventRecord(c->ev0, c->stream));
    rc = enqueue_pipeline(c, c->A, n, maxk, wclasses, 1, 1, c->tune.sample_blocks, c->tune.row_blocks,
                          scan_blocks(c, n), c->tune.smooth_blocks, c->ev_stage);
    if (rc) return rc;
    c->stats_batch = true;
    BMF_HIP(hipEventRecord(c->ev1, c->stream));
    BMF_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < BMF_NUM_STAGES; i++)
        BMF_HIP(hipEventElapsedTime(&stage_ms[i], c->ev_stage[i], c->ev_stage[i + 1]));
    return BMF_OK;
}

int bmf_set_mesh_options(bmf_ctx* c, const bmf_mesh_options* o) {
    if (!c || !o) return BMF_EINVAL;
    if (o->mesh != BMF_MESH_TRIANGLES && o->mesh != BMF_MESH_QUADS) return BMF_EINVAL;
    if (o->placement != BMF_PLACE_MASS_POINT && o->placement != BMF_PLACE_QEF) return BMF_EINVAL;
    if (!finite(o->qef_tolerance) || o->qef_tolerance < 0.0f) return BMF_EINVAL;
    if (o->quads_as_triangles != 0 && o->quads_as_triangles != 1) return BMF_EINVAL;
    if (o->normals != 0 && o->normals != 1) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(c->device));
    BMF_HIP(hipStreamSynchronize(c->stream));
    Arenas& A = c->A;
    if (o->mesh == BMF_MESH_QUADS && !A.ring) {
        if (dalloc(&A.ring, (int64_t)kRingSlots * A.cap_vertices) != hipSuccess ||
            dalloc(&A.ring_n, A.cap_vertices) != hipSuccess || dalloc(&A.qef_err, A.cap_vertices) != hipSuccess) {
            if (A.ring) (void)hipFree(A.ring);
            if (A.ring_n) (void)hipFree(A.ring_n);
            if (A.qef_err) (void)hipFree(A.qef_err);
            A.ring = A.ring_n = nullptr;
            A.qef_err = nullptr;
            return BMF_EHIP;
        }
    }
    if (o->normals && !A.nrm) {
        if (dalloc(&A.nrm4, A.cap_vertices) != hipSuccess || dalloc(&A.fn4, A.cap_indices / 3 + 1) != hipSuccess ||
            dalloc(&A.nrm, 3 * A.cap_vertices) != hipSuccess) {
            if (A.nrm4) (void)hipFree(A.nrm4);
            if (A.fn4) (void)hipFree(A.fn4);
            if (A.nrm) (void)hipFree(A.nrm);
            A.nrm4 = A.fn4 = nullptr;
            A.nrm = nullptr;
            return BMF_EHIP;
        }
    }
    c->mesh = *o;
    A.normals = o->normals;
    A.mesh_mode = o->mesh;
    A.placement = o->placement;
    A.qef_tol = o->qef_tolerance;
    A.quad_tris = o->quads_as_triangles;
    drop_graphs(c);
    c->have_extract = false;
    c->results_valid = false;
    return BMF_OK;
}

int bmf_normals_to_host(bmf_ctx* c, int chunk, float* normals, int64_t cap_v) {
    if (!c || !normals) return BMF_EINVAL;
    if (!c->A.normals) return BMF_ESTATE;
    BMF_HIP(hipSetDevice(c->device));
    int rc = fetch_results(c);
    if (rc) return rc;
    if (chunk < 0 || chunk >= c->last_n) return BMF_EINVAL;
    const bmf_chunk_result& r = c->h_results[chunk];
    if (r.status != BMF_OK) return r.status;
    if (cap_v < r.vertex_count) return BMF_ECAPACITY;
    if (!r.vertex_count) return BMF_OK;
    BMF_HIP(hipMemcpyAsync(normals, c->A.nrm + 3 * r.vertex_offset, sizeof(float) * 3 * r.vertex_count,
                           hipMemcpyDeviceToHost, c->stream));
    BMF_HIP(hipStreamSynchronize(c->stream));
    return BMF_OK;
}

int bmf_get_mesh_options(bmf_ctx* c, bmf_mesh_options* o) {
    if (!c || !o) return BMF_EINVAL;
    *o = c->mesh;
    return BMF_OK;
}

int bmf_qef_error_to_host(bmf_ctx* c, int chunk, float* errors, int64_t cap_v) {
    if (!c || !errors) return BMF_EINVAL;
    BMF_HIP(hipSetDevice(c->device));
    int rc = fetch_results(c);
    if (rc) return rc;
    if (chunk < 0 || chunk >= c->last_n) return BMF_EINVAL;
    const bmf_chunk_result& r = c->h_results[chunk];
    if (r.status != BMF_OK) return r.status;
    if (cap_v < r.vertex_count) return BMF_ECAPACITY;
    if (!r.vertex_count) return BMF_OK;
    if (c->A.qef_err && c->mesh.mesh == BMF_MESH_QUADS && c->mesh.placement == BMF_PLACE_QEF) {
        BMF_HIP(hipMemcpyAsync(errors, c->A.qef_err + r.vertex_offset, sizeof(float) * r.vertex_count,
                               hipMemcpyDeviceToHost, c->stream));
        BMF_HIP(hipStreamSynchronize(c->stream));
    } else {
        std::memset(errors, 0, sizeof(float) * (size_t)r.vertex_count);
    }
    return BMF_OK;
}

int bmf_get_tuning(bmf_ctx* c, bmf_tuning* t) {
    if (!c || !t) return BMF_EINVAL;
    *t = c->tune;
    return BMF_OK;
}

int bmf_set_tuning(bmf_ctx* c, const bmf_tuning* t) {
    if (!c || !t) return BMF_EINVAL;
    bmf_tuning n = c->tune;
    if (t->sample_tile) {
        if (t->sample_tile != 1024 && t->sample_tile != 2048 && t->sample_tile != 4096) return BMF_EINVAL;
        n.sample_tile = t->sample_tile;
    }
    const int32_t cap = 64 * c->cu_count;  // at most 64 workgroups per CU for any kernel
    if (t->sample_blocks < 0 || t->row_blocks < 0 || t->smooth_blocks < 0) return BMF_EINVAL;
    if (t->sample_blocks > cap || t->row_blocks > cap || t->smooth_blocks > cap) return BMF_EINVAL;
    if (t->sample_blocks) n.sample_blocks = t->sample_blocks;
    if (t->row_blocks) n.row_blocks = t->row_blocks;
    if (t->smooth_blocks) n.smooth_blocks = (t->smooth_blocks + kXcds - 1) / kXcds * kXcds;
    if (t->row_split) {
        if (t->row_split != -1 && t->row_split != 1 && t->row_split != 2 && t->row_split != 4 && t->row_split != 8)
            return BMF_EINVAL;
        n.row_split = t->row_split;
    }
    BMF_HIP(hipSetDevice(c->device));
    BMF_HIP(hipStreamSynchronize(c->stream));
    c->tune = n;
    c->A.sample_tile = n.sample_tile;
    c->A.row_split = n.row_split;
    drop_graphs(c);
    return BMF_OK;
}

int bmf_autotune(bmf_ctx* c, const bmf_chunk_desc* chunks, int n, int reps, bmf_tuning* best) {
    if (!c || reps < 1 || n < 1) return BMF_EINVAL;
    if (!c->sampler_set) return BMF_ESTATE;
    int rc = bmf_validate_chunks(chunks, n, &c->lim);
    if (rc) return rc;
    BMF_HIP(hipSetDevice(c->device));
    BMF_HIP(hipStreamSynchronize(c->stream));
    c->results_valid = false;
    c->have_extract = false;
    rc = upload_descs(c, chunks, n);
    if (rc) return rc;
    int maxk = max_steps(c, chunks, n), wclasses = 0;
    for (int i = 0; i < n; i++) wclasses |= row_word_class(chunks[i].dim);
    // stage times (ms, the median of reps runs: a candidate with one lucky
    // run does not win) of the pipeline under tuning t
    float st[BMF_NUM_STAGES];
    auto run_n = [&](const bmf_tuning& t, int nrep, float* out) -> int {
        Arenas A = c->A;
        A.sample_tile = t.sample_tile;
        A.row_split = t.row_split;
        std::vector<float> ms(static_cast<size_t>(nrep) * BMF_NUM_STAGES);
        for (int r = 0; r < nrep; r++) {
            int e = enqueue_pipeline(c, A, n, maxk, wclasses, 1, 1, t.sample_blocks, t.row_blocks,
                                     scan_blocks(c, n), t.smooth_blocks, c->ev_stage);
            if (e) return e;
            BMF_HIP(hipStreamSynchronize(c->stream));
            for (int i = 0; i < BMF_NUM_STAGES; i++)
                BMF_HIP(hipEventElapsedTime(&ms[(size_t)i * nrep + r], c->ev_stage[i], c->ev_stage[i + 1]));
        }
        for (int i = 0; i < BMF_NUM_STAGES; i++) {
            float* v = ms.data() + (size_t)i * nrep;
            std::nth_element(v, v + nrep / 2, v + nrep);
            out[i] = v[nrep / 2];
        }
        return BMF_OK;
    };
    auto run = [&](const bmf_tuning& t, float* out) { return run_n(t, reps, out); };
    const bmf_tuning start = c->tune;
    bmf_tuning cur = c->tune;
    const int cu = c->cu_count;
    // sampling: tile x persistent workgroups (stage 1)
    {
        float bt = 1e30f;
        bmf_tuning b = cur;
        const int tiles[3] = {1024, 2048, 4096};
        const int mult[4] = {2, 4, 8, 16};
        for (int ti : tiles)
            for (int m : mult) {
                bmf_tuning t = cur;
                t.sample_tile = ti;
                t.sample_blocks = m * cu;
                rc = run(t, st);
                if (rc) return rc;
                if (st[1] < bt) {
                    bt = st[1];
                    b = t;
                }
            }
        cur.sample_tile = b.sample_tile;
        cur.sample_blocks = b.sample_blocks;
    }
    // plane kernels: count (+ refine) and emit (stages 2 and 4), workgroups x row segments
    {
        float bt = 1e30f;
        int bb = cur.row_blocks, bs = cur.row_split;
        const int mult[7] = {1, 2, 4, 8, 10, 12, 16};
        const int splits[5] = {-1, 1, 2, 4, 8};  // -1: per row-word class from its plane count
        for (int sp : splits)
            for (int m : mult) {
                bmf_tuning t = cur;
                t.row_blocks = m * cu;
                t.row_split = sp;
                rc = run(t, st);
                if (rc) return rc;
                if (st[2] + st[4] < bt) {
                    bt = st[2] + st[4];
                    bb = t.row_blocks;
                    bs = t.row_split;
                }
            }
        cur.row_blocks = bb;
        cur.row_split = bs;
    }
    // smoothing (stage 5)
    if (maxk > 0) {
        float bt = 1e30f;
        int bb = cur.smooth_blocks;
        // from 64 workgroups (a few small meshes) to 32 per CU
        const int cand[7] = {8 * kXcds, cu, 2 * cu, 4 * cu, 8 * cu, 16 * cu, 32 * cu};
        for (int b : cand) {
            bmf_tuning t = cur;
            t.smooth_blocks = (b + kXcds - 1) / kXcds * kXcds;
            rc = run(t, st);
            if (rc) return rc;
            if (st[5] < bt) {
                bt = st[5];
                bb = t.smooth_blocks;
            }
        }
        cur.smooth_blocks = bb;
    }
    // confirmation: the tuned and the starting shapes over the whole frame,
    // alternately (median of 3 * reps runs each); the tuned ones are kept
    // only when they are at least 1 % faster
    {
        float frame[2] = {0.0f, 0.0f};
        const bmf_tuning pair[2] = {start, cur};
        for (int k = 0; k < 2; k++) {
            rc = run_n(pair[k], 3 * reps, st);
            if (rc) return rc;
            for (int i = 0; i < BMF_NUM_STAGES; i++) frame[k] += st[i];
        }
        if (!(frame[1] < 0.99f * frame[0])) cur = start;
    }
    rc = bmf_set_tuning(c, &cur);
    if (rc) return rc;
    if (best) *best = cur;
    return BMF_OK;
}

}  // extern "C"
