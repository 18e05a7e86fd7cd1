/*
 * bmf.h -- C ABI of the MI355X chunk-extraction library (libbmf.so).
 *
 * Drop-in boundary for the per-chunk extraction path of BinaryMeshFitting:
 * the reference's ChunkGenerator::process_queue (ChunkGenerator.cpp:27-60)
 * runs ChunkGenerator::extract_chunk (ChunkGenerator.cpp:80-147) over a batch
 * of chunks: DMCChunk::label_grid (DMCChunk.cpp:79-166) -> label_edges
 * (:168-508) -> polygonize (:514-576) -> MeshProcessor<3> smoothing
 * (MeshProcessor.cpp:25-306).  Here one bmf_extract call takes the batch's
 * chunk descriptors and runs that whole path on the GPU.
 *
 * Conventions (same as the reference):
 *   - dim = samples per axis; cells emitted = (dim-1)^3; 2 <= dim <= 256.
 *   - sample (i,j,k) sits at overlap_pos + (i,j,k)*delta, with
 *     delta = size*(1+2*overlap)/(dim-1), overlap_pos = pos - size*overlap.
 *   - a sample is "air" (label bit set) when its density is < 0; densities
 *     are positive inside the solid.
 *   - vertex positions are in grid units (world = p*delta + overlap_pos);
 *     indices are chunk-local triangle lists (3 per triangle).
 *
 * All functions return BMF_OK (0) or a positive BMF_E* code; nothing throws
 * or aborts across the boundary.  A context is thread-compatible, not
 * thread-safe (one per host thread, like the reference's single watcher
 * thread).  Work is ordered on the context's HIP stream.
 */
#ifndef BMF_H
#define BMF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMF_OK 0
#define BMF_EINVAL 1       /* bad argument or descriptor */
#define BMF_ECAPACITY 2    /* an arena was too small; bmf_result reports what was needed */
#define BMF_EHIP 3         /* HIP runtime error */
#define BMF_ENODEV 4       /* no usable GPU */
#define BMF_ESTATE 5       /* call out of order (e.g. mesh access before extraction) */

#define BMF_MAX_DIM 256
#define BMF_MAX_OCTAVES 16

typedef struct bmf_ctx bmf_ctx;

/* Arena sizes fixed at creation (the reference's ResourceAllocator pools,
 * ResourceAllocator.hpp:34-62, become preallocated device arenas). */
typedef struct {
    int32_t max_chunks;     /* descriptors per bmf_extract call */
    int64_t max_samples;    /* sum of dim^3 over the sampled chunks of one call */
    int64_t max_vertices;   /* output vertices over all chunks of one call */
    int64_t max_indices;    /* output indices over all chunks of one call */
} bmf_limits;

/* Samplers (Sampler.hpp:24-34 and its create_sampler_* factories). */
enum {
    BMF_SAMPLER_TERRAIN = 0, /* d = amp*fbm(p) - (p.y - y0); fbm = FBM of 3D simplex noise */
    BMF_SAMPLER_SPHERE = 1,  /* ImplicitFunctions::sphere, ImplicitSampler.cpp:46-50 */
    BMF_SAMPLER_CUBOID = 2,  /* ImplicitFunctions::cuboid, ImplicitSampler.cpp:52-60 */
    BMF_SAMPLER_TORUS_Z = 3, /* ImplicitFunctions::torus_z, ImplicitSampler.cpp:27-34 */
    BMF_SAMPLER_PLANE_Y = 4, /* ImplicitFunctions::plane_y, ImplicitSampler.cpp:62-65 */
    BMF_SAMPLER_CSG = 5,     /* max-union of spheres and boxes, floored at -16 */
    BMF_SAMPLER_EXTERNAL = 6, /* densities supplied by the caller in device memory (any
                                 host sampler's Sampler::block output, Sampler.hpp:20-34) */
    BMF_SAMPLER_NOISE = 7     /* the NoiseSamplers block shapes (NoiseSampler.cpp:113-318)
                                 with the project's own noise: see the noise_* fields and
                                 bmf_noise_preset */
};

/* NOISE sampler: basis, fractal and coordinate perturbation. */
enum { BMF_NOISE_VALUE = 0, BMF_NOISE_SIMPLEX = 1 };
enum { BMF_FRACTAL_FBM = 0, BMF_FRACTAL_RIGID = 1 };
enum { BMF_PERTURB_NONE = 0, BMF_PERTURB_GRADIENT = 1, BMF_PERTURB_OFFSET_XZ = 2 };
#define BMF_MAX_PERTURB_OCTAVES 24

typedef struct {
    int32_t kind;
    float world_size;                 /* Sampler::world_size (implicit radii derive from it) */
    float amp, y0;                    /* terrain */
    float freq, lacunarity, gain;     /* terrain FBM */
    int32_t octaves;                  /* 1..BMF_MAX_OCTAVES */
    uint32_t seed;
    const float* spheres;             /* CSG, host memory: n_spheres x (cx, cy, cz, r) */
    int32_t n_spheres;
    const float* boxes;               /* CSG, host memory: n_boxes x (cx, cy, cz, hx, hy, hz) */
    int32_t n_boxes;
    const float* densities;           /* EXTERNAL, device memory: the dim^3 FP32 [x][y][z] blocks
                                         of the chunks of each later bmf_extract, packed in
                                         descriptor order; must stay valid while work is queued */
    /* NOISE: at sample position p, c = g_scale p, warped by the perturbation; n = the
     * fractal (octaves, freq, lacunarity, gain, seed above) of the basis at c; density
     * d = -((p.y g_scale) y_scale) - height n.  noise_dims 2 evaluates the noise over
     * (x, z) only (a height field broadcast along y, as terrain2d_block's one-row
     * NOISE_BLOCK), 3 over the volume. */
    int32_t noise_dims;               /* 2 or 3 */
    int32_t noise_basis;              /* BMF_NOISE_* */
    int32_t noise_fractal;            /* BMF_FRACTAL_* */
    float g_scale, y_scale, height;
    int32_t perturb;                  /* BMF_PERTURB_*; OFFSET_XZ needs noise_dims 3 */
    int32_t perturb_octaves;          /* 0..BMF_MAX_PERTURB_OCTAVES */
    float perturb_amp, perturb_freq, perturb_gain;
} bmf_sampler_desc;

/* Fills `out` (every field) with the NOISE sampler that stands for one of the
 * reference's block samplers, with the settings those functions pass to
 * FastNoiseSIMD carried over (its defaults where they set none: frequency
 * 0.01, 3 fractal octaves, gain 0.5, lacunarity 2, 3 perturb octaves).  A
 * perturbation given there in the noise's frequency-scaled coordinates
 * (frequency F, amplitude A) acts here on c, so it is carried as F * 0.01 and
 * A / 0.01.  world_size is Sampler::world_size (windy3d's `resolution`). */
enum {
    BMF_PRESET_TERRAIN2D = 0,      /* NoiseSamplers::terrain2d_block, NoiseSampler.cpp:113-146 */
    BMF_PRESET_TERRAIN2D_PERT = 1, /* terrain2d_pert_block, :148-194 (WorldOctree.cpp:40's default) */
    BMF_PRESET_TERRAIN3D = 2,      /* terrain3d_block, :196-227 */
    BMF_PRESET_TERRAIN3D_PERT = 3, /* terrain3d_pert_block, :229-263 */
    BMF_PRESET_WINDY3D = 4,        /* windy3d_block, :265-318 */
    BMF_PRESET_WORLD = 5           /* terrain2d_pert_block with the world's noise_properties,
                                      WorldOctree.cpp:40-54 */
};
int bmf_noise_preset(int32_t which, float world_size, uint32_t seed, bmf_sampler_desc* out);

/* One chunk: DMCChunk fields (DMCChunk.hpp:24-60) plus the per-chunk
 * WorldProperties the generator applies (ChunkGenerator.cpp:85-98). */
typedef struct {
    float pos[3];                 /* DMCChunk::pos */
    float size;                   /* DMCChunk::size */
    int32_t dim;                  /* DMCChunk::dim (samples per axis) */
    float overlap;                /* result of the overlap rule, ChunkGenerator.cpp:98 */
    int32_t iters;                /* smoothing iterations K (process_iters) */
    int32_t boundary_processing;  /* WorldProperties::boundary_processing */
} bmf_chunk_desc;

enum {
    BMF_MODE_PLAIN = 0,   /* one launch sequence per chunk, host loop over chunks */
    BMF_MODE_BATCHED = 1  /* whole batch: persistent kernels + device work queue, captured in a hipGraph */
};

/* Per-chunk result (after bmf_result). */
typedef struct {
    int64_t vertex_offset;   /* first vertex of the chunk in the view's arrays */
    int64_t index_offset;    /* first index of the chunk */
    int32_t vertex_count;
    int32_t index_count;     /* 3 x triangles */
    int32_t contains_mesh;   /* DMCChunk::contains_mesh (DMCChunk.cpp:159-162) */
    int32_t status;          /* BMF_OK, or BMF_ECAPACITY when the chunk did not fit */
    int32_t sampled;         /* 1 when densities were evaluated, 0 when the chunk was
                                classified uniform without sampling */
    float delta;             /* DMCChunk::scale */
    float overlap_pos[3];    /* DMCChunk::overlap_pos */
} bmf_chunk_result;

/* Device pointers into the context's output arenas; valid until the next
 * bmf_extract on the context. */
typedef struct {
    const float* positions;  /* [vertices][3], grid units */
    const uint32_t* indices; /* chunk-local vertex ids */
    const uint8_t* boundary; /* DualVertex::boundary (DMCChunk.cpp:673) */
    const float* normals;    /* [vertices][3] unit normals, or NULL when not enabled */
    int64_t total_vertices;
    int64_t total_indices;
} bmf_mesh_view;

/* ---- lifetime ---------------------------------------------------------- */
/* A context owns its arenas, one stream and its captured frames; several
 * contexts on one device share nothing, and their batches run concurrently
 * (a stream of batches alternated over two overlaps on the GPU). */
int bmf_create(bmf_ctx** out, int device, const bmf_limits* limits);
void bmf_destroy(bmf_ctx* ctx);
const char* bmf_strerror(int code);
const char* bmf_version(void);

/* Replaces Sampler::block + NoiseSamplerProperties (Sampler.hpp:24-34,
 * NoiseSampler.hpp:7-22); the sampler applies to later extractions. */
int bmf_set_sampler(bmf_ctx* ctx, const bmf_sampler_desc* sampler);

/* Checks descriptors on the host only (no GPU needed). */
int bmf_validate_chunks(const bmf_chunk_desc* chunks, int n, const bmf_limits* limits);

/* ---- the path ---------------------------------------------------------- */
/* ChunkGenerator::process_queue -> extract_chunk over n chunks (host
 * descriptor array).  Enqueues work on the context stream and returns. */
int bmf_extract(bmf_ctx* ctx, const bmf_chunk_desc* chunks, int n, int mode);

/* Waits for the last bmf_extract and copies per-chunk results to `out`
 * (n entries, may be NULL).  Returns BMF_ECAPACITY if any chunk did not fit;
 * `need_vertices/need_indices` (may be NULL) then hold the totals needed. */
int bmf_result(bmf_ctx* ctx, bmf_chunk_result* out, int n, int64_t* need_vertices, int64_t* need_indices);

int bmf_mesh_view_get(bmf_ctx* ctx, bmf_mesh_view* view);

/* Copies one chunk's mesh to host memory (the VerticesIndicesBlock of
 * ChunkBlocks.hpp:161-179, as SoA).  Returns BMF_ECAPACITY if cap_* is too
 * small (counts are still written). */
int bmf_mesh_to_host(bmf_ctx* ctx, int chunk, float* positions, uint32_t* indices, uint8_t* boundary,
                     int64_t cap_vertices, int64_t cap_indices, int64_t* vertices, int64_t* indices_out);

int bmf_sync(bmf_ctx* ctx);

/* Stage inspection for one chunk, to host memory.  Densities: dim^3 FP32
 * [x][y][z] (Sampler::block output).  Labels: the reference's u32 words
 * [x][y][ceil(dim/32)] of DMCChunk::label_grid.  Masks: the reference's
 * u64 cell-mask words [x][y][ceil(dim/8)] of DMCChunk::label_edges. */
int bmf_sample_block(bmf_ctx* ctx, const bmf_chunk_desc* chunk, float* densities);
int bmf_label_block(bmf_ctx* ctx, const bmf_chunk_desc* chunk, uint32_t* words, int32_t* contains_mesh);
int bmf_cell_masks(bmf_ctx* ctx, const bmf_chunk_desc* chunk, uint64_t* masks);

/* Device time of the last bmf_extract (HIP events on the context stream,
 * ms).  Replaces the reference's clock() batch timer (WorldWatcher.cpp:69-73). */
int bmf_last_timing(bmf_ctx* ctx, float* total_ms);

/* Work done by the last extraction: the samples of its sampled chunks and,
 * for the terrain sampler, the noise octaves that its adaptive sampling
 * evaluated, summed over samples (every octave of every sample, as the
 * reference's block sampler computes, would be samples * octaves), and those
 * the refinement of early-labelled edge endpoints evaluated (0 for the other
 * samplers).  Valid after bmf_result of a BATCHED extraction (or after
 * bmf_profile_stages); BMF_ESTATE after a PLAIN one, whose chunks run one
 * pipeline each.  Counterpart of the reference's per-stage counters in its
 * timing printout (DMCChunk.cpp:785-840). */
typedef struct {
    int64_t samples;
    int64_t octave_evals;
    int64_t refine_evals;
} bmf_frame_stats;
int bmf_get_frame_stats(bmf_ctx* ctx, bmf_frame_stats* out);

/* Runs the BATCHED path once without a graph, with HIP events between the
 * stages on the context stream, and returns the stage times in ms:
 * [0] plan, [1] sample, [2] count, [3] scan, [4] emit, [5] smoothing (all
 * iterations), [6] finalize (no kernel since the chunk results are written
 * by the plan and scan stages: the event pair's own overhead, a few us).
 * The per-stage counterpart of the reference's
 * stage timers (DMCChunk.cpp:785-840, ChunkGenerator.cpp:153-193). */
#define BMF_NUM_STAGES 7
int bmf_profile_stages(bmf_ctx* ctx, const bmf_chunk_desc* chunks, int n, float* stage_ms);

/* ---- mesh options ------------------------------------------------------- */
/* Topology and vertex placement of the extraction (per context).
 * TRIANGLES is the reference's live path: marching-cubes triangles with one
 * vertex per sign-changing edge (DMCChunk.cpp:440-689).  QUADS is the dual
 * marching cubes of the reference's polygon table (Tables.hpp:76-354, no
 * live caller there): one vertex per EdgeTable polygon of each cell, one
 * quad per sign-changing edge whose four cells lie in the chunk, wound like
 * the triangles (right-hand normal toward the solid side), smoothed as
 * MeshProcessor<4> (or, written as triangles, as MeshProcessor<3> after
 * MeshProcessor::flush_to_tris, MeshProcessor.cpp:73-91). */
enum { BMF_MESH_TRIANGLES = 0, BMF_MESH_QUADS = 1 };
/* QUADS vertex placement: the mass point of the polygon's edge points, or
 * the QEF of those points with unit normals from the sampler's analytic
 * gradient (qef_simd.h:392-579: FP32 Jacobi SVD, 5 sweeps, singular values
 * kept in [tol, 1/tol]; CSG sampler only). */
enum { BMF_PLACE_MASS_POINT = 0, BMF_PLACE_QEF = 1 };

typedef struct {
    int32_t mesh;                /* BMF_MESH_* */
    int32_t placement;           /* BMF_PLACE_* (QUADS only) */
    float qef_tolerance;         /* truncation threshold of the QEF pseudo-inverse */
    int32_t quads_as_triangles;  /* QUADS: write (a,b,c), (c,d,a) per quad */
    int32_t normals;             /* also compute unit vertex normals (below), either mesh */
} bmf_mesh_options;

/* Vertex normals.  The reference computes them only with smooth_normals
 * (MeshProcessor.cpp:159-216, 271-287), off on its live path
 * (ChunkGenerator.cpp:116); here they follow that branch of MeshProcessor<N>:
 * per smoothing step m the face normal while m == 0, m < min(K/2-3, 10) or
 * m < 3 -- for a triangle -cross(normalize(p0-p1), normalize(p0-p2)), for a
 * quad -normalize((n1 + n2) * 0.5) with n1 = cross(normalize(p0-p1),
 * normalize(p0-p2)), n2 = cross(normalize(p2-p3), normalize(p2-p0)) and the
 * reference's NaN fallbacks (MeshProcessor.cpp:187-205) -- else the sum of
 * the corners' normals; per primal step the ring average, normalised when
 * m == 3 or (m == 0 and K <= 3), kept when its y is not 0, for every vertex
 * with a ring (boundary vertices included; their positions stay pinned);
 * K = 0 runs one such step without moving positions; the output is the final
 * normal divided by its length (0 for a zero normal or a vertex in no
 * polygon).  QUADS written as triangles use the triangle face normal. */
int bmf_normals_to_host(bmf_ctx* ctx, int chunk, float* normals, int64_t cap_vertices);

int bmf_set_mesh_options(bmf_ctx* ctx, const bmf_mesh_options* options);
int bmf_get_mesh_options(bmf_ctx* ctx, bmf_mesh_options* options);

/* Per-vertex QEF residual |ATb - ATA x|^2 (qef_simd.h:434-440) of one
 * chunk's vertices (QUADS with BMF_PLACE_QEF; zeros otherwise). */
int bmf_qef_error_to_host(bmf_ctx* ctx, int chunk, float* errors, int64_t cap_vertices);

/* ---- output formatting ------------------------------------------------- */
/* The meshes of the last bmf_extract as VBO-ready arrays in device memory:
 * WorldOctreeNode::format -> GLChunk::format_data (WorldOctreeNode.cpp:68-86,
 * GLChunk.cpp:278-297) for every chunk of the batch, concatenated in batch
 * order as DMCChunk::copy_verts_and_inds does (DMCChunk.cpp:578-591).  A
 * chunk is included when it contains a mesh (contains_mesh, the condition of
 * both) and its status is BMF_OK; its first vertex / index in the arrays is
 * the sum of the vertex / index counts of the included chunks before it
 * (bmf_result gives the counts). */
enum { BMF_SPACE_GRID = 0,   /* positions in grid units: the reference's VBO (its vertex
                                shader applies p * scale + overlap_pos, shaders/main_vs.glsl:34) */
       BMF_SPACE_WORLD = 1   /* positions p * scale + overlap_pos (one FP32 multiply, then an add) */
};
enum { BMF_COLOR_NONE = 0,   /* no colours written */
       BMF_COLOR_WHITE = 1,  /* DualVertex::color = (1, 1, 1) (DMCChunk.cpp:681) */
       BMF_COLOR_NOISE = 2   /* ColorMapper::generate_colors (ColorMapper.cpp:15-62): n = 4 * FBM at
                                the vertex's grid position (the project's simplex FBM, frequency 0.01,
                                lacunarity 2, gain 0.5, 4 octaves, seed 1337; the reference's
                                FastNoiseSIMD is not part of its tree), colour
                                hsl_to_rgb((n + 1) * 0.5 * 360, 0.72, 1) (:60-119, an HSV map) */
};
typedef struct {
    int32_t space;           /* BMF_SPACE_* */
    int32_t colors;          /* BMF_COLOR_* */
    int32_t global_indices;  /* 1: indices plus the chunk's first vertex in the arrays; 0: chunk-local */
} bmf_format_options;

typedef struct {
    float* positions;        /* device [cap_vertices][3] */
    float* colors;           /* device [cap_vertices][3]; may be NULL with BMF_COLOR_NONE */
    float* normals;          /* device [cap_vertices][3] unit normals, or NULL (needs mesh option normals) */
    uint32_t* indices;       /* device [cap_indices] */
    int64_t cap_vertices;
    int64_t cap_indices;
} bmf_vbo;

/* Formats the last extraction into `out` on the context stream, waits, and
 * writes the totals to vertices / indices (may be NULL).  BMF_ECAPACITY (and
 * nothing written) when they exceed the capacities; BMF_ESTATE before any
 * extraction or when normals are asked for without the mesh option. */
int bmf_format(bmf_ctx* ctx, const bmf_format_options* options, const bmf_vbo* out, int64_t* vertices,
               int64_t* indices);

/* ---- device work-queue producer ---------------------------------------- */
/* The watcher thread's LOD pass on the GPU: WorldWatcher::check_leaves
 * (WorldWatcher.cpp:155-212) over the renderables list with
 * WorldOctree::node_needs_split / node_needs_group (WorldOctree.cpp:239-261),
 * then process_batch (WorldWatcher.cpp:214-238, split_node / group_node_1)
 * and ChunkGenerator::process_queue's chunk set-up (create_chunk,
 * WorldOctree.cpp:263-269; the overlap rule, ChunkGenerator.cpp:98).  The
 * descriptors of the resulting generate batch are written straight into the
 * context's descriptor arena; bmf_extract_queued then extracts them with no
 * host upload. */
#define BMF_NODE_DRAW 1u             /* NODE_FLAGS (WorldOctreeNode.hpp:15-27) */
#define BMF_NODE_DESTROY 2u
#define BMF_NODE_SPLIT 4u
#define BMF_NODE_GROUP 8u
#define BMF_NODE_DIRTY 32u
#define BMF_NODE_DRAW_CHILDREN 64u
#define BMF_NODE_SUPERCEDED 256u

/* The fields of WorldOctreeNode (WorldOctreeNode.hpp) the pass reads; one
 * entry per node of the world octree, referenced by index. */
typedef struct {
    float pos[3];            /* OctreeNode::pos */
    float size;              /* OctreeNode::size */
    int32_t level;           /* OctreeNode::level */
    uint32_t flags;          /* BMF_NODE_* */
    int32_t world_leaf;      /* world_leaf_flag */
    int32_t parent;          /* node index, -1 for the root */
    int32_t children[8];     /* node indices, -1 when absent */
    uint64_t morton;         /* morton_code */
} bmf_world_node;

typedef struct {
    float focus[3];          /* WorldWatcher::focus_pos */
    float split_multiplier;  /* WorldProperties (WorldOctree.cpp:20-33) */
    float group_multiplier;
    float size_modifier;
    int32_t max_level;
    int32_t min_level;
    int32_t max_gen;         /* the cap of check_leaves (400 in WorldWatcher::update) */
    int32_t chunk_resolution;/* dim of every chunk */
    int32_t process_iters;
    int32_t boundary_processing;
    float overlap;           /* base overlap of the rule */
} bmf_lod_params;

enum { BMF_LOD_SPLIT = 1, BMF_LOD_GROUP = 2 };

/* One chunk of the generate batch, in batch order. */
typedef struct {
    int32_t node;            /* the split node (the chunk is its child `child`) or the grouped node */
    int32_t child;           /* 0..7 (the child's corner, Tables.hpp:7-9), -1 for a grouped node */
    int32_t level;           /* the chunk's level */
    int32_t action;          /* BMF_LOD_SPLIT or BMF_LOD_GROUP */
    uint64_t morton;         /* the chunk node's morton code ((parent << 3) | corner for a child) */
} bmf_lod_item;

/* nodes [n_nodes], renderables [n_renderables] (node indices in list order),
 * flags_out [n_nodes] (every node's flags after the pass: SPLIT / GROUP set on
 * the pushed nodes), dirty_out [n_renderables] (the pushed nodes in push
 * order) and items_out [cap_items] are device memory.  Waits, then writes the
 * counts; BMF_ECAPACITY when the generate batch exceeds cap_items or the
 * context's max_chunks (n_chunks then holds the size needed). */
int bmf_lod_queue(bmf_ctx* ctx, const bmf_world_node* nodes, int32_t n_nodes, const int32_t* renderables,
                  int32_t n_renderables, const bmf_lod_params* params, uint32_t* flags_out, int32_t* dirty_out,
                  bmf_lod_item* items_out, int32_t cap_items, int32_t* n_dirty, int32_t* n_chunks);

/* bmf_extract of the descriptors the last bmf_lod_queue left in the arena
 * (BMF_ESTATE when there are none: a later bmf_extract or forward pass
 * replaced them). */
int bmf_extract_queued(bmf_ctx* ctx, int mode);

/* ---- seams between chunks ---------------------------------------------- */
/* The stitcher: marching-cubes triangles of the dual cells that straddle the
 * boundaries between the chunks of a world octree, the output of
 * WorldStitcher::stitch_all (WorldStitcher.cpp:26-49, stitch_cell :184-239,
 * the face / edge recursions :241-460, stitch_indexes :491-572) over the
 * chunks' boundary-shell octrees (DMCChunk::generate_octree,
 * DMCChunk.cpp:700-778).  Every world leaf is a chunk of
 * chunk_resolution^3 unit nodes (a power of two) placed as the shell octree
 * places them (pos + MCD * size/2 + MCD * size/4 + ..., FP32 in that order);
 * a node's value is the context's sampler at that position (the reference's
 * read of the chunk's densities is commented out there, DMCChunk.cpp:755,
 * which leaves every value 0 and the seams empty).  The dual cells are those
 * of the reference's recursion (around every point where the octant
 * neighbourhoods fall in more than one chunk), enumerated point by point;
 * triangles are written in a fixed order: the chunks of `leaves` in order,
 * then the lattice points of each chunk's closed cube on its boundary in
 * x-major order (each point by the finest of its chunks, the lowest node
 * index among equals), then the triangle table's order.  Chunk meshes
 * extracted on the same lattice (bmf_chunk_desc.size = node size *
 * (dim - 1) / dim) close against the seams.  `nodes` / `leaves` (every world
 * leaf, in the output order) / `triangles` ([cap_triangles][3][3] FP32, a
 * triangle soup) are device memory; nodes need correct level and morton
 * fields.  Waits; BMF_ECAPACITY (nothing written) when the count exceeds
 * cap_triangles.  The EXTERNAL sampler has no values off the chunks' sample
 * grids and is rejected. */
typedef struct {
    int32_t chunk_resolution;  /* unit nodes per chunk edge: a power of two, 2..256 */
} bmf_stitch_params;

int bmf_stitch(bmf_ctx* ctx, const bmf_world_node* nodes, int32_t n_nodes, const int32_t* leaves, int32_t n_leaves,
               const bmf_stitch_params* params, float* triangles, int64_t cap_triangles, int64_t* n_triangles);

/* ---- launch shapes ------------------------------------------------------ */
/* Tile and workgroup shapes of the BATCHED path (the PLAIN path sizes its
 * launches per chunk).  A field left 0 keeps its current value. */
typedef struct {
    int32_t sample_tile;    /* samples per work-queue tile of the sampling kernel: 1024, 2048, 4096 */
    int32_t sample_blocks;  /* persistent workgroups of the sampling kernel */
    int32_t row_blocks;     /* persistent workgroups of the plane kernels (count, refine, emit) */
    int32_t smooth_blocks;  /* workgroups of the smoothing kernels (rounded up to a multiple of 8) */
    int32_t row_split;      /* row segments per x-plane item of the plane kernels: 1, 2, 4 or 8; -1
                               doubles the split (up to 8) while sampled planes x split < 512 */
} bmf_tuning;

int bmf_get_tuning(bmf_ctx* ctx, bmf_tuning* tuning);
int bmf_set_tuning(bmf_ctx* ctx, const bmf_tuning* tuning);

/* Autotuner: runs the BATCHED pipeline on these chunks with candidate
 * shapes (HIP events around each stage, the median of `reps` runs per
 * candidate), takes the fastest shape of each stage, then times those shapes
 * against the context's starting ones over the whole frame and keeps them
 * only when they are at least 1 % faster; the result goes to the context and
 * to `best` (may be NULL).  Outputs of the runs are discarded. */
int bmf_autotune(bmf_ctx* ctx, const bmf_chunk_desc* chunks, int n, int reps, bmf_tuning* best);

/* ---- differentiable extraction --------------------------------------- */
/* One chunk's density-positioned vertices plus K smoothing iterations as a
 * function of its dim^3 densities, with the topology (labels, cells,
 * triangles) frozen by the forward pass.  The reference has no adjoint; the
 * forward is ChunkGenerator::extract_chunk for one chunk (ChunkGenerator.cpp:
 * 80-147) and the backward is its transpose: smoothing (MeshProcessor.cpp:
 * 130-306) is linear in the positions for a fixed topology, and vertex
 * placement (DMCChunk.cpp:657-674) depends on the two densities of its edge
 * through mu = (0 - s0) / (s1 - s0).
 *
 * A tape belongs to one context and uses its stream and arenas for the
 * forward extraction; it owns copies of everything the backward reads, so
 * later bmf_extract calls on the context do not disturb it. */
typedef struct bmf_tape bmf_tape;

typedef struct {
    int32_t dim;
    int32_t iters;            /* K */
    int32_t checkpoint_every; /* smoothing states kept every this many iterations */
    int32_t vertex_count;
    int32_t index_count;
    int32_t contains_mesh;
    const float* positions;   /* device [V][3]: positions after K iterations, grid units */
    const uint32_t* indices;  /* device [index_count] */
    const uint8_t* boundary;  /* device [V] */
    const float* densities;   /* device [dim^3]: the densities the forward used */
} bmf_tape_info;

/* max_dim bounds the chunk's samples per axis; max_vertices (at most 2^30)
 * / max_indices bound its mesh; checkpoint_every >= 1.  The tape holds
 * about 280 bytes per vertex of max_vertices (ring records, positions,
 * gradients), 4 per index, 8 per sample of max_dim^3, and 16 per vertex
 * per kept smoothing state. */
int bmf_tape_create(bmf_tape** out, bmf_ctx* ctx, int32_t max_dim, int64_t max_vertices, int64_t max_indices,
                    int32_t checkpoint_every);
void bmf_tape_destroy(bmf_tape* tape);

/* Forward pass for one chunk.  `densities` is a device pointer to its dim^3
 * FP32 block, or NULL to evaluate the context's sampler at every sample
 * (all octaves, no early exit).  Positions after iterations 0, c, 2c, ...
 * (c = checkpoint_every) are kept on the tape.  Synchronises the stream. */
int bmf_forward_diff(bmf_ctx* ctx, const bmf_chunk_desc* chunk, const float* densities, bmf_tape* tape);
int bmf_tape_get(bmf_tape* tape, bmf_tape_info* info);

/* Positions after smoothing iteration k (0 <= k <= K) into a device buffer
 * [V][3], recomputed from the last checkpoint at or before k; bit-identical
 * to the forward's own state at k. */
int bmf_tape_positions(bmf_tape* tape, int32_t iteration, float* positions);

/* Loss L = (1/V) sum_v |p_v - t(v)|^2 over the tape's final positions, t(v)
 * the nearest of the n device targets [n][3] (squared FP32 distance
 * ((dx*dx + dy*dy) + dz*dz); ties go to the lower target index).  Writes
 * dL/dp [V][3] (device), the chosen target per vertex (device, may be NULL)
 * and L to host memory (may be NULL). */
int bmf_loss_nearest(bmf_tape* tape, const float* targets, int64_t n_targets, float* dloss_dpos, int32_t* nearest,
                     float* loss);

/* The same loss split in two for a fixed target set (an optimisation loop
 * over the densities with the targets unchanged): bmf_tape_set_targets
 * builds the search hierarchy of the n device targets [n][3] once (the tape
 * keeps its own sorted copy; the caller's buffer may be reused after the
 * call returns), and bmf_tape_loss_nearest then evaluates the loss of the
 * tape's current forward against that set, with the outputs of
 * bmf_loss_nearest.  The set stays until the next bmf_tape_set_targets or
 * bmf_loss_nearest on the tape (which builds its own); forward passes keep
 * it.  bmf_tape_loss_nearest returns BMF_ESTATE when no set was built. */
int bmf_tape_set_targets(bmf_tape* tape, const float* targets, int64_t n_targets);
int bmf_tape_loss_nearest(bmf_tape* tape, float* dloss_dpos, int32_t* nearest, float* loss);

/* bmf_loss_nearest for callers that cannot allocate device memory (code
 * under HIP graph capture, tight optimisation loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The search hierarchy is
 * built in a static device array of the library (10.5 MiB per GPU, held until
 * the process ends; its address is looked up once per device by
 * bmf_tape_create) instead of the tape's growable buffers, so the call
 * allocates nothing; n_targets above BMF_NOALLOC_MAX_TARGETS returns
 * BMF_ECAPACITY.  Same outputs as bmf_loss_nearest bit for bit, except that
 * the FP64 loss goes to DEVICE memory `loss` (may be NULL) and nothing is
 * synchronised: the work is enqueued on the context stream, and the targets
 * must stay valid until it completes.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (tapes of contexts with
 * different streams, without ordering between them) race. */
#define BMF_NOALLOC_MAX_TARGETS 262144
int bmf_loss_nearest_noalloc(bmf_tape* tape, const float* targets, int64_t n_targets, float* dloss_dpos,
                             int32_t* nearest, double* loss);

/* Device time in ms (HIP events on the context stream) of the tape's last
 * forward pass [0], loss [1] (bmf_loss_nearest or bmf_tape_loss_nearest) and
 * backward pass [2]; -1 for a call that has not completed. */
int bmf_tape_timing(bmf_tape* tape, float* ms);

/* Backward pass: dL/d(density) [dim^3] (device) from dL/d(final positions)
 * [V][3] (device).  Samples that are not an endpoint of a sign-changing
 * edge get 0. */
int bmf_backward_diff(bmf_tape* tape, const float* dloss_dpos, float* dloss_ddensity);

#ifdef __cplusplus
}
#endif
#endif
