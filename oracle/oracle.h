/*
 * CPU oracle for the chunk-extraction path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's per-chunk extraction
 * (Lin20/BinaryMeshFitting), used by tests/ as the parity checker and by
 * bench.py as the CPU baseline.  Nothing in the product links or calls it.
 * Each function cites the reference code it restates.  Compiled with
 * -ffp-contract=off (no FMA contraction, the reference builds with
 * /fp:strict, BinaryMeshFitting.vcxproj:138); the noise sampler uses
 * explicit fmaf() at fixed places, mirrored exactly by the HIP kernels.
 *
 * Parity pins (tests/test_oracle_*.py): every function here is checked
 * against an independent formulation -- numpy brute force over corners,
 * edges, cells and adjacency, closed forms of the smoothing operator,
 * the reference-held MC table, exhaustive single-cell cases, an FP64 twin,
 * and special cases of the QEF solver.  The FBM noise has no reference
 * counterpart (the reference calls the absent FastNoiseSIMD library), so it
 * is pinned to its own published definition (Gustavson 2005 simplex noise)
 * by a brute-force all-lattice-points evaluation in numpy.
 */
#ifndef BMF_ORACLE_H
#define BMF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_SAMPLER_TERRAIN = 0, /* d = amp*fbm(p) - (p.y - y0) */
    OR_SAMPLER_SPHERE = 1,  /* ImplicitSampler.cpp:46-50 */
    OR_SAMPLER_CUBOID = 2,  /* ImplicitSampler.cpp:52-60 */
    OR_SAMPLER_TORUS_Z = 3, /* ImplicitSampler.cpp:27-34 */
    OR_SAMPLER_PLANE_Y = 4, /* ImplicitSampler.cpp:62-65 */
    OR_SAMPLER_CSG = 5,     /* union of spheres and boxes, positive inside */
    OR_SAMPLER_NOISE = 7    /* noise fields: NoiseSampler.cpp:113-318 shapes, own noise */
};
enum { OR_NOISE_VALUE = 0, OR_NOISE_SIMPLEX = 1 };
enum { OR_FRACTAL_FBM = 0, OR_FRACTAL_RIGID = 1 };
enum { OR_PERTURB_NONE = 0, OR_PERTURB_GRADIENT = 1, OR_PERTURB_OFFSET_XZ = 2 };

typedef struct {
    int32_t kind;
    float world_size;   /* Sampler::world_size (Sampler.hpp:25) */
    float amp, y0;      /* terrain: d = amp*fbm - (y - y0) */
    float freq, lacunarity, gain;
    int32_t octaves;
    uint32_t seed;
    const float* spheres; /* CSG: n_spheres x (cx, cy, cz, r) */
    int32_t n_spheres;
    const float* boxes;   /* CSG: n_boxes x (cx, cy, cz, hx, hy, hz) */
    int32_t n_boxes;
    /* NOISE (oracle_noise.c) */
    int32_t noise_dims, noise_basis, noise_fractal;
    float g_scale, y_scale, height;
    int32_t perturb, perturb_octaves;
    float perturb_amp, perturb_freq, perturb_gain;
} or_sampler_t;

typedef struct {
    float pos[3];   /* DMCChunk::pos */
    float size;     /* DMCChunk::size */
    int32_t dim;    /* samples per axis (DMCChunk::dim) */
    float overlap;  /* ChunkGenerator.cpp:98 result */
    int32_t iters;  /* WorldProperties::process_iters */
    int32_t boundary_processing;
} or_chunk_t;

typedef struct {
    float* pos;          /* V x 3, grid units */
    uint8_t* boundary;   /* V */
    uint32_t* edge;      /* V: packed (x,y,z,axis) of the lattice edge: ((x*dim+y)*dim+z)*3+axis */
    uint32_t* idx;       /* nI triangle indices */
    int64_t V, nI;
    int32_t contains_mesh;
    float delta;         /* DMCChunk::scale */
    float opos[3];       /* DMCChunk::overlap_pos */
} or_mesh_t;

/* ---- geometry (DMCChunk.cpp:94-101) ---- */
void or_chunk_geometry(const or_chunk_t* c, float* delta, float opos[3]);

/* ---- samplers: fill dim^3 [x][y][z] at opos + i*delta ---- */
float or_snoise3_f32(float x, float y, float z, uint32_t seed);
double or_snoise3_f64(double x, double y, double z, uint32_t seed);
uint32_t or_grad_index(int32_t i, int32_t j, int32_t k, uint32_t seed);
void or_fbm_params(const or_sampler_t* s, float* fo, float* ao, float* inv_norm);
/* Noise fields (oracle_noise.c). */
float or_value3_f32(float x, float y, float z, uint32_t seed);
float or_value2_f32(float x, float z, uint32_t seed);
double or_value3_f64(double x, double y, double z, uint32_t seed);
double or_value2_f64(double x, double z, uint32_t seed);
float or_lattice_value(int32_t i, int32_t j, int32_t k, uint32_t seed);
float or_noise_field_f32(const or_sampler_t* s, float x, float y, float z);
double or_noise_field_f64(const or_sampler_t* s, double x, double y, double z);
/* n (the fractal after the perturbation) at the column / point, FP32. */
float or_noise_fractal_f32(const or_sampler_t* s, float x, float y, float z);
void or_noise_block_f32(const or_sampler_t* s, int dim, const float opos[3], float delta, float* out);
float or_terrain_value_f32(const or_sampler_t* s, float x, float y, float z);
/* The sampler's FP32 value at one point (every octave of the terrain). */
float or_sample_value_f32(const or_sampler_t* s, float x, float y, float z);
/* The terrain's FBM alone (normalised, clamped), with or_fbm_params' tables. */
float or_fbm_f32(const or_sampler_t* s, const float* fo, const float* ao, float inv_norm, float x, float y, float z);
double or_terrain_value_f64(const or_sampler_t* s, double x, double y, double z);
void or_sample_block_f32(const or_sampler_t* s, int dim, const float opos[3], float delta, float* out);
void or_sample_block_f64(const or_sampler_t* s, int dim, const double opos[3], double delta, double* out);
float or_csg_value_f32(const or_sampler_t* s, float x, float y, float z, float grad[3]);
double or_csg_value_f64(const or_sampler_t* s, double x, double y, double z, double grad[3]);

/* ---- DMCChunk::label_grid threshold (DMCChunk.cpp:113-162) ---- */
int or_label_grid(const float* dens, int dim, uint32_t* words);

/* ---- DMCChunk::label_edges mask passes (DMCChunk.cpp:200-438) ----
 * masks must hold dim*dim*ceil(dim/8) + 4 words. */
void or_cell_masks(const uint32_t* words, int dim, uint64_t* masks);

/* ---- label_edges cell loop + polygonize (DMCChunk.cpp:440-576, 593-689) ----
 * Writes up to capV vertices and capI indices; returns 0, or -1 when a
 * capacity is too small (V/nI then hold the required counts). */
int or_mesh_from_masks(const float* dens, const uint64_t* masks, int dim,
                       float* pos, uint8_t* boundary, uint32_t* edge, uint8_t* valence, int64_t capV,
                       uint32_t* idx, int64_t capI, int64_t* V, int64_t* nI);

/* ---- MeshProcessor<3> smoothing as driven by ChunkGenerator.cpp:112-124 ---- */
int or_smooth_f32(float* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI,
                  int iters, int boundary_processing);
int or_smooth_f64(double* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI,
                  int iters, int boundary_processing);

/* ---- ChunkGenerator::extract_chunk for one chunk (ChunkGenerator.cpp:80-147) ----
 * dens_out (dim^3) is optional.  The mesh arrays are malloc'd; free with or_free_mesh. */
int or_extract_chunk(const or_sampler_t* s, const or_chunk_t* c, or_mesh_t* out, float* dens_out);
void or_free_mesh(or_mesh_t* m);

/* FP64 twin of vertex placement + smoothing on a fixed topology (the FP32
 * mesh's edges and indices): positions from FP64 densities. */
int or_positions_f64(const or_sampler_t* s, const or_chunk_t* c, const or_mesh_t* topo, double* pos_out);

/* ---- batch driver: OpenMP parallel for over chunks (ChunkGenerator.cpp:91) ----
 * Returns the number of chunks processed; per-chunk V, nI, contains_mesh
 * and an order-sensitive checksum of positions and indices. */
int or_extract_batch(const or_sampler_t* s, const or_chunk_t* chunks, int n, int threads,
                     int64_t* V, int64_t* nI, int32_t* contains, uint64_t* checksum);

/* ---- QEF (qef_simd.h:392-579): FP32 with exact 1/sqrt, and FP64 twin ---- */
float or_qef_solve_f32(const float* p, const float* n, int count, float tol, float out[3]);
double or_qef_solve_f64(const double* p, const double* n, int count, double tol, double out[3]);

/* ---- differentiable extraction, FP64 (oracle_diff.c) ----
 * Topology frozen: edge[] / boundary[] / idx[] of an or_mesh_t. */
void or_place_f64(const double* dens, int dim, const uint32_t* edge, int64_t V, double* pos);
void or_forward_f64(const double* dens, int dim, const uint32_t* edge, const uint8_t* boundary, int64_t V,
                    const uint32_t* idx, int64_t nI, int K, int bp, double* pos);
void or_jvp_f64(const double* dens, int dim, const uint32_t* edge, const uint8_t* boundary, int64_t V,
                const uint32_t* idx, int64_t nI, int K, int bp, const double* ddens, double* dpos);
void or_adjoint_f64(const double* dens, int dim, const uint32_t* edge, const uint8_t* boundary, int64_t V,
                    const uint32_t* idx, int64_t nI, int K, int bp, const double* dldp, double* dldd);
typedef struct or_kdtree or_kdtree;
or_kdtree* or_kd_build(const double* pts, int64_t n);
void or_kd_free(or_kdtree* t);
void or_kd_nearest(const or_kdtree* t, const double* q, int64_t m, int64_t* out, double* d2, int threads);
/* FP32 distances ((dx*dx + dy*dy) + dz*dz, each step rounded), ties to the lower index */
void or_kd_nearest_f32(const or_kdtree* t, const float* q, int64_t m, int64_t* out, float* d2, int threads);
double or_loss_nearest_f64(const double* pos, int64_t V, const double* targets, const int64_t* nearest,
                           double* dldp);
void or_fd_grad_f64(const double* dens, int dim, const uint32_t* edge, const uint8_t* boundary, int64_t V,
                    const uint32_t* idx, int64_t nI, int K, int bp, const or_kdtree* kd, const int64_t* samples,
                    int m, double h, int threads, double* out, int32_t* kink);

/* ---- dual-marching-cubes quad mode (oracle_dmc.c) ----
 * polygons of EdgeTable[mask] (Tables.hpp:96-354): edges[] concatenated,
 * polygon p = edges[start[p] .. start[p+1]); returns the polygon count. */
int or_dmc_polygons(int mask, int* edges, int* start);
int or_dmc_poly_of_edge(int mask, int local_edge);
/* placement 0 = mass point, 1 = QEF with the CSG sampler's unit gradients.
 * code[v] = x | y<<8 | z<<16 | polygon<<24 of the vertex's cell.  Returns 0,
 * -1 when a capacity is too small (counts still written), -2 on a table
 * inconsistency. */
int or_dmc_mesh(const or_sampler_t* s, const float* dens, int dim, const float opos[3], float delta, int placement,
                float tol, int as_triangles, float* pos, uint8_t* boundary, uint32_t* code, float* qerr,
                int64_t capV, uint32_t* idx, int64_t capI, int64_t* V, int64_t* nI);
int or_smooth_n_f32(float* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int N,
                    int iters, int boundary_processing);
int or_smooth_n_f64(double* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int N,
                    int iters, int boundary_processing);
/* Triangle-path smoothing with vertex normals (the smooth_normals branch of
 * MeshProcessor<3>, MeshProcessor.cpp:159-216, 271-287, as defined in
 * oracle_dmc.c); positions as or_smooth_*, normals [V][3] unit or zero. */
int or_smooth_normals_f32(float* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int iters,
                          int boundary_processing, float* normals);
int or_smooth_normals_f64(double* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int iters,
                          int boundary_processing, double* normals);
/* N-gon (N = 3 or 4) smoothing with vertex normals: MeshProcessor<N>'s
 * smooth_normals branch, quad face normals of MeshProcessor.cpp:187-205
 * (oracle_dmc.c). */
int or_smooth_normals_n_f32(float* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int N,
                            int iters, int boundary_processing, float* normals);
int or_smooth_normals_n_f64(double* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int N,
                            int iters, int boundary_processing, double* normals);
void or_dmc_positions_f64(const or_sampler_t* s, const double* dens, int dim, const double opos[3], double delta,
                          int placement, double tol, const uint32_t* code, int64_t V, double* pos, double* qerr);


/* ---- output formatting and vertex colours (oracle_format.c) ---- */
void or_hsv_to_rgb_f32(float h, float s, float v, float out[3]);
float or_color_noise_f32(float x, float y, float z);
void or_vertex_color_f32(const float p[3], float rgb[3]);
void or_format_batch(int n, const int64_t* V, const int64_t* I, const int32_t* contains_mesh, const float* delta,
                     const float* opos, const int64_t* voff, const int64_t* ioff, const float* pos,
                     const float* nrm, const uint32_t* idx, int space, int colors, int global, float* out_pos,
                     float* out_col, float* out_nrm, uint32_t* out_idx, int64_t* total_v, int64_t* total_i);


/* ---- the watcher's LOD pass (oracle_lod.c); layouts of bmf_world_node,
 * bmf_lod_params and bmf_lod_item ---- */
typedef struct {
    float pos[3];
    float size;
    int32_t level;
    uint32_t flags;
    int32_t world_leaf;
    int32_t parent;
    int32_t children[8];
    uint64_t morton;
} or_world_node_t;
typedef struct {
    float focus[3];
    float split_multiplier, group_multiplier, size_modifier;
    int32_t max_level, min_level, max_gen, chunk_resolution, process_iters, boundary_processing;
    float overlap;
} or_lod_params_t;
typedef struct {
    int32_t node, child, level, action;
    uint64_t morton;
} or_lod_item_t;
int or_node_needs_split(const or_world_node_t* n, const or_lod_params_t* P);
int or_node_needs_group(const or_world_node_t* n, const or_lod_params_t* P);
int or_lod_queue(const or_world_node_t* nodes, int n_nodes, const int32_t* rlist, int n_r, const or_lod_params_t* P,
                 uint32_t* flags, int32_t* dirty, or_lod_item_t* items, or_chunk_t* descs, int cap, int* n_dirty,
                 int* n_chunks);


/* ---- seams between the chunks of a world octree (oracle_seams.c) ---- */
int64_t or_stitch_recursive(const or_sampler_t* s, const or_world_node_t* nodes, int n_nodes, int dim, float* out,
                            int64_t cap);
int64_t or_stitch_points(const or_sampler_t* s, const or_world_node_t* nodes, int n_nodes, const int32_t* leaves,
                         int n_leaves, int dim, float* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
