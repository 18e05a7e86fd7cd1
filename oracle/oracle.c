/*
 * CPU oracle for the chunk-extraction path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, the smoke check and bench.py's CPU baseline load this code;
 * the product never does.  See oracle.h for what pins each function.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off, no fast-math).
 */
#include "oracle.h"
#include "../include/bmf_mc_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Geometry: DMCChunk::label_grid, DMCChunk.cpp:94-101                       */
/* ------------------------------------------------------------------------ */
void or_chunk_geometry(const or_chunk_t* c, float* delta, float opos[3])
{
    /* delta = size * (1.0f + overlap * 2.0f) / (float)(dim - 1) */
    float t = c->overlap * 2.0f;
    t = 1.0f + t;
    t = c->size * t;
    *delta = t / (float)(c->dim - 1);
    /* overlap_pos = pos - size * overlap */
    float so = c->size * c->overlap;
    opos[0] = c->pos[0] - so;
    opos[1] = c->pos[1] - so;
    opos[2] = c->pos[2] - so;
}

/* ------------------------------------------------------------------------ */
/* FBM simplex noise (the project's own sampler; no reference counterpart:   */
/* the reference's noise comes from the absent FastNoiseSIMD library,        */
/* NoiseSampler.cpp:196-227 is the sampler family it belongs to).            */
/*                                                                          */
/* 3D simplex noise after S. Gustavson, "Simplex noise demystified" (2005):  */
/* skew F3 = 1/3, unskew G3 = 1/6, kernel (r2 - |d|^2)^4 with r2 = 0.5 (the  */
/* continuous variant), gradient set and selection of K. Perlin, "Improving  */
/* Noise" (2002), lattice hash of our own (24-bit multiplies + xor-shifts),  */
/* output scaled by 76.8 and clamped to [-1, 1] so that |noise| <= 1 holds   */
/* exactly.  FBM sums octaves o = 0..O-1 with frequency f0*L^o, amplitude    */
/* g^o and seed + o, then normalises by the amplitude sum and clamps.        */
/* ------------------------------------------------------------------------ */
#define NZ_P1 0x5BD1E9u
#define NZ_P2 0x27D4EBu
#define NZ_P3 0x165667u
#define NZ_M1 0x9E3779u
#define NZ_M2 0x85EBCBu
#define NZ_SCALE 76.8f
#define NZ_R2 0.5f

static inline uint32_t mul24(uint32_t a, uint32_t b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }

uint32_t or_grad_index(int32_t i, int32_t j, int32_t k, uint32_t seed)
{
    uint32_t h = mul24((uint32_t)i, NZ_P1) ^ mul24((uint32_t)j, NZ_P2) ^ mul24((uint32_t)k, NZ_P3) ^ seed;
    h ^= h >> 13;
    h = mul24(h ^ (h >> 24), NZ_M1);
    h ^= h >> 11;
    h = mul24(h ^ (h >> 24), NZ_M2);
    h ^= h >> 16;
    return (h >> 4) & 15u;
}

static inline float grad_dot_f32(uint32_t h, float x, float y, float z)
{
    float u = h < 8u ? x : y;
    float v = h < 4u ? y : ((h == 12u || h == 14u) ? x : z);
    float su = (h & 1u) ? -u : u;
    float sv = (h & 2u) ? -v : v;
    return su + sv;
}

static inline float corner_f32(float x, float y, float z, uint32_t h)
{
    float t = fmaf(-x, x, NZ_R2);
    t = fmaf(-y, y, t);
    t = fmaf(-z, z, t);
    t = fmaxf(t, 0.0f);
    t = t * t;
    return (t * t) * grad_dot_f32(h, x, y, z);
}

float or_snoise3_f32(float x, float y, float z, uint32_t seed)
{
    const float F3 = 1.0f / 3.0f;
    const float G3 = 1.0f / 6.0f;
    const float G3x2 = 2.0f * G3;
    float s = ((x + y) + z) * F3;
    float fi = floorf(x + s), fj = floorf(y + s), fk = floorf(z + s);
    int32_t i = (int32_t)fi, j = (int32_t)fj, k = (int32_t)fk;
    float t = (float)(i + j + k) * G3;
    float x0 = x - (fi - t), y0 = y - (fj - t), z0 = z - (fk - t);
    int gxy = x0 >= y0, gyz = y0 >= z0, gxz = x0 >= z0;
    int i1 = gxy & gxz, j1 = (!gxy) & gyz, k1 = (!gxz) & (!gyz);
    int i2 = gxy | gxz, j2 = (!gxy) | gyz, k2 = (!gxz) | (!gyz);
    float x1 = (x0 - (float)i1) + G3, y1 = (y0 - (float)j1) + G3, z1 = (z0 - (float)k1) + G3;
    float x2 = (x0 - (float)i2) + G3x2, y2 = (y0 - (float)j2) + G3x2, z2 = (z0 - (float)k2) + G3x2;
    float x3 = (x0 - 1.0f) + 0.5f, y3 = (y0 - 1.0f) + 0.5f, z3 = (z0 - 1.0f) + 0.5f;
    float n0 = corner_f32(x0, y0, z0, or_grad_index(i, j, k, seed));
    float n1 = corner_f32(x1, y1, z1, or_grad_index(i + i1, j + j1, k + k1, seed));
    float n2 = corner_f32(x2, y2, z2, or_grad_index(i + i2, j + j2, k + k2, seed));
    float n3 = corner_f32(x3, y3, z3, or_grad_index(i + 1, j + 1, k + 1, seed));
    float n = ((n0 + n1) + n2) + n3;
    float r = NZ_SCALE * n;
    return fminf(fmaxf(r, -1.0f), 1.0f);
}

static inline double grad_dot_f64(uint32_t h, double x, double y, double z)
{
    double u = h < 8u ? x : y;
    double v = h < 4u ? y : ((h == 12u || h == 14u) ? x : z);
    return ((h & 1u) ? -u : u) + ((h & 2u) ? -v : v);
}

static inline double corner_f64(double x, double y, double z, uint32_t h)
{
    double t = 0.5 - x * x - y * y - z * z;
    if (t < 0.0) t = 0.0;
    t = t * t;
    return t * t * grad_dot_f64(h, x, y, z);
}

double or_snoise3_f64(double x, double y, double z, uint32_t seed)
{
    const double F3 = 1.0 / 3.0, G3 = 1.0 / 6.0;
    double s = (x + y + z) * F3;
    double fi = floor(x + s), fj = floor(y + s), fk = floor(z + s);
    int32_t i = (int32_t)fi, j = (int32_t)fj, k = (int32_t)fk;
    double t = (double)(i + j + k) * G3;
    double x0 = x - (fi - t), y0 = y - (fj - t), z0 = z - (fk - t);
    int gxy = x0 >= y0, gyz = y0 >= z0, gxz = x0 >= z0;
    int i1 = gxy & gxz, j1 = (!gxy) & gyz, k1 = (!gxz) & (!gyz);
    int i2 = gxy | gxz, j2 = (!gxy) | gyz, k2 = (!gxz) | (!gyz);
    double n = corner_f64(x0, y0, z0, or_grad_index(i, j, k, seed));
    n += corner_f64(x0 - i1 + G3, y0 - j1 + G3, z0 - k1 + G3, or_grad_index(i + i1, j + j1, k + k1, seed));
    n += corner_f64(x0 - i2 + 2 * G3, y0 - j2 + 2 * G3, z0 - k2 + 2 * G3, or_grad_index(i + i2, j + j2, k + k2, seed));
    n += corner_f64(x0 - 1 + 3 * G3, y0 - 1 + 3 * G3, z0 - 1 + 3 * G3, or_grad_index(i + 1, j + 1, k + 1, seed));
    double r = 76.8 * n;
    return r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
}

void or_fbm_params(const or_sampler_t* s, float* fo, float* ao, float* inv_norm)
{
    float f = s->freq, a = 1.0f, norm = 0.0f;
    for (int o = 0; o < s->octaves; o++) {
        fo[o] = f;
        ao[o] = a;
        norm = norm + a;
        f = f * s->lacunarity;
        a = a * s->gain;
    }
    *inv_norm = 1.0f / norm;
}

/* The FBM of the terrain, normalised and clamped to [-1, 1]. */
float or_fbm_f32(const or_sampler_t* s, const float* fo, const float* ao, float inv_norm, float x, float y, float z)
{
    float S = 0.0f;
    for (int o = 0; o < s->octaves; o++) {
        float n = or_snoise3_f32(x * fo[o], y * fo[o], z * fo[o], s->seed + (uint32_t)o);
        S = fmaf(ao[o], n, S);
    }
    float f = S * inv_norm;
    return fminf(fmaxf(f, -1.0f), 1.0f);
}

static inline float terrain_f32(const or_sampler_t* s, const float* fo, const float* ao, float inv_norm,
                                float x, float y, float z)
{
    const float f = or_fbm_f32(s, fo, ao, inv_norm, x, y, z);
    float t1 = s->amp * f;
    float t2 = y - s->y0;
    return t1 - t2;
}

float or_terrain_value_f32(const or_sampler_t* s, float x, float y, float z)
{
    float fo[32], ao[32], inv;
    or_fbm_params(s, fo, ao, &inv);
    return terrain_f32(s, fo, ao, inv, x, y, z);
}

double or_terrain_value_f64(const or_sampler_t* s, double x, double y, double z)
{
    double f = s->freq, a = 1.0, norm = 0.0, S = 0.0;
    for (int o = 0; o < s->octaves; o++) {
        S += a * or_snoise3_f64(x * f, y * f, z * f, s->seed + (uint32_t)o);
        norm += a;
        f *= s->lacunarity;
        a *= s->gain;
    }
    double v = S / norm;
    v = v < -1.0 ? -1.0 : (v > 1.0 ? 1.0 : v);
    return s->amp * v - (y - s->y0);
}

/* ------------------------------------------------------------------------ */
/* Implicit samplers (ImplicitSampler.cpp:27-65), positive inside.           */
/* ------------------------------------------------------------------------ */
static inline float implicit_f32(int kind, float res, float x, float y, float z)
{
    switch (kind) {
    case OR_SAMPLER_SPHERE: { /* -(length(p) - res*0.25) */
        float r = res * 0.25f;
        float l = sqrtf((x * x + y * y) + z * z);
        return -(l - r);
    }
    case OR_SAMPLER_CUBOID: { /* -min(max(d), length(d)), d = |p| - res/8 */
        float r = res / 8.0f;
        float dx = fabsf(x) - r, dy = fabsf(y) - r, dz = fabsf(z) - r;
        float m = fmaxf(dx, fmaxf(dy, dz));
        float l = sqrtf((dx * dx + dy * dy) + dz * dz);
        return -fminf(m, l);
    }
    case OR_SAMPLER_TORUS_Z: {
        float r1 = res / 4.0f, r2 = res / 10.0f;
        float qx = fabsf(sqrtf(x * x + y * y)) - r1;
        float len = sqrtf(qx * qx + z * z);
        return -(len - r2);
    }
    case OR_SAMPLER_PLANE_Y:
        return -y;
    }
    return 0.0f;
}

static inline double implicit_f64(int kind, double res, double x, double y, double z)
{
    switch (kind) {
    case OR_SAMPLER_SPHERE: return -(sqrt(x * x + y * y + z * z) - res * 0.25);
    case OR_SAMPLER_CUBOID: {
        double r = res / 8.0;
        double dx = fabs(x) - r, dy = fabs(y) - r, dz = fabs(z) - r;
        double m = fmax(dx, fmax(dy, dz));
        return -fmin(m, sqrt(dx * dx + dy * dy + dz * dz));
    }
    case OR_SAMPLER_TORUS_Z: {
        double qx = fabs(sqrt(x * x + y * y)) - res / 4.0;
        return -(sqrt(qx * qx + z * z) - res / 10.0);
    }
    case OR_SAMPLER_PLANE_Y: return -y;
    }
    return 0.0;
}

/* CSG: union (max) of spheres r - |p - c| and exact boxes -sdf, floored at
 * -16 (the floor keeps the field identical under per-chunk primitive
 * culling).  Gradient of the winning primitive (first index on ties),
 * zero on the floor. */
#define CSG_FLOOR (-16.0f)
float or_csg_value_f32(const or_sampler_t* s, float x, float y, float z, float grad[3])
{
    float best = CSG_FLOOR;
    float g[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < s->n_spheres; i++) {
        const float* sp = s->spheres + 4 * i;
        float dx = x - sp[0], dy = y - sp[1], dz = z - sp[2];
        float l = sqrtf((dx * dx + dy * dy) + dz * dz);
        float v = sp[3] - l;
        if (v > best) {
            best = v;
            if (grad) {
                float inv = l > 0.0f ? 1.0f / l : 0.0f;
                g[0] = -dx * inv; g[1] = -dy * inv; g[2] = -dz * inv;
            }
        }
    }
    for (int i = 0; i < s->n_boxes; i++) {
        const float* b = s->boxes + 6 * i;
        float px = x - b[0], py = y - b[1], pz = z - b[2];
        float qx = fabsf(px) - b[3], qy = fabsf(py) - b[4], qz = fabsf(pz) - b[5];
        float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
        float outside = sqrtf((mx * mx + my * my) + mz * mz);
        float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
        float v = -(outside + inside);
        if (v > best) {
            best = v;
            if (grad) {
                float sx = px < 0.0f ? -1.0f : 1.0f, sy = py < 0.0f ? -1.0f : 1.0f, sz = pz < 0.0f ? -1.0f : 1.0f;
                if (outside > 0.0f) {
                    float inv = 1.0f / outside;
                    g[0] = -mx * sx * inv; g[1] = -my * sy * inv; g[2] = -mz * sz * inv;
                } else {
                    g[0] = g[1] = g[2] = 0.0f;
                    if (qx >= qy && qx >= qz) g[0] = -sx;
                    else if (qy >= qz) g[1] = -sy;
                    else g[2] = -sz;
                }
            }
        }
    }
    if (grad) { grad[0] = g[0]; grad[1] = g[1]; grad[2] = g[2]; }
    return best;
}

double or_csg_value_f64(const or_sampler_t* s, double x, double y, double z, double grad[3])
{
    double best = (double)CSG_FLOOR;
    double g[3] = {0, 0, 0};
    for (int i = 0; i < s->n_spheres; i++) {
        const float* sp = s->spheres + 4 * i;
        double dx = x - sp[0], dy = y - sp[1], dz = z - sp[2];
        double l = sqrt(dx * dx + dy * dy + dz * dz);
        double v = sp[3] - l;
        if (v > best) {
            best = v;
            double inv = l > 0 ? 1.0 / l : 0.0;
            g[0] = -dx * inv; g[1] = -dy * inv; g[2] = -dz * inv;
        }
    }
    for (int i = 0; i < s->n_boxes; i++) {
        const float* b = s->boxes + 6 * i;
        double px = x - b[0], py = y - b[1], pz = z - b[2];
        double qx = fabs(px) - b[3], qy = fabs(py) - b[4], qz = fabs(pz) - b[5];
        double mx = fmax(qx, 0), my = fmax(qy, 0), mz = fmax(qz, 0);
        double outside = sqrt(mx * mx + my * my + mz * mz);
        double inside = fmin(fmax(qx, fmax(qy, qz)), 0);
        double v = -(outside + inside);
        if (v > best) {
            best = v;
            double sx = px < 0 ? -1 : 1, sy = py < 0 ? -1 : 1, sz = pz < 0 ? -1 : 1;
            if (outside > 0) {
                g[0] = -mx * sx / outside; g[1] = -my * sy / outside; g[2] = -mz * sz / outside;
            } else {
                g[0] = g[1] = g[2] = 0;
                if (qx >= qy && qx >= qz) g[0] = -sx;
                else if (qy >= qz) g[1] = -sy;
                else g[2] = -sz;
            }
        }
    }
    if (grad) { grad[0] = g[0]; grad[1] = g[1]; grad[2] = g[2]; }
    return best;
}

/* Sample block: NoiseSampler.cpp:19-31 / ImplicitSampler.cpp:12-19.
 * Coordinate = (float)i * delta + origin (two roundings, no FMA). */
void or_sample_block_f32(const or_sampler_t* s, int dim, const float opos[3], float delta, float* out)
{
    if (s->kind == OR_SAMPLER_NOISE) {
        or_noise_block_f32(s, dim, opos, delta, out);
        return;
    }
    float fo[32], ao[32], inv = 1.0f;
    if (s->kind == OR_SAMPLER_TERRAIN) or_fbm_params(s, fo, ao, &inv);
    for (int ix = 0; ix < dim; ix++) {
        float x = (float)ix * delta + opos[0];
        for (int iy = 0; iy < dim; iy++) {
            float y = (float)iy * delta + opos[1];
            float* row = out + ((int64_t)ix * dim + iy) * dim;
            for (int iz = 0; iz < dim; iz++) {
                float z = (float)iz * delta + opos[2];
                float v;
                if (s->kind == OR_SAMPLER_TERRAIN) v = terrain_f32(s, fo, ao, inv, x, y, z);
                else if (s->kind == OR_SAMPLER_CSG) v = or_csg_value_f32(s, x, y, z, 0);
                else v = implicit_f32(s->kind, s->world_size, x, y, z);
                row[iz] = v;
            }
        }
    }
}

static double sample_f64(const or_sampler_t* s, double x, double y, double z)
{
    if (s->kind == OR_SAMPLER_TERRAIN) return or_terrain_value_f64(s, x, y, z);
    if (s->kind == OR_SAMPLER_CSG) return or_csg_value_f64(s, x, y, z, 0);
    if (s->kind == OR_SAMPLER_NOISE) return or_noise_field_f64(s, x, y, z);
    return implicit_f64(s->kind, s->world_size, x, y, z);
}

void or_sample_block_f64(const or_sampler_t* s, int dim, const double opos[3], double delta, double* out)
{
    for (int ix = 0; ix < dim; ix++)
        for (int iy = 0; iy < dim; iy++)
            for (int iz = 0; iz < dim; iz++)
                out[((int64_t)ix * dim + iy) * dim + iz] =
                    sample_f64(s, ix * delta + opos[0], iy * delta + opos[1], iz * delta + opos[2]);
}

/* ------------------------------------------------------------------------ */
/* Labels: DMCChunk::label_grid, DMCChunk.cpp:113-162.                       */
/* Bit z%32 of word [x][y][z/32] is set when the sample is < 0 (air).        */
/* ------------------------------------------------------------------------ */
int or_label_grid(const float* dens, int dim, uint32_t* words)
{
    const int zc = (dim + 31) / 32;
    int mesh = 0, positive = 0, negative = 0;
    for (int x = 0; x < dim; x++)
        for (int y = 0; y < dim; y++)
            for (int zb = 0; zb < zc; zb++) {
                const float* s = dens + ((int64_t)x * dim + y) * dim + zb * 32;
                int zmax = dim - zb * 32;
                if (zmax > 32) zmax = 32;
                uint32_t m = 0;
                for (int z = 0; z < zmax; z++)
                    if (s[z] < 0.0f) m |= 1u << z;
                words[((int64_t)x * dim + y) * zc + zb] = m;
                if (m != 0) {
                    if (m != 0xFFFFFFFFu) mesh = 1;
                    else if (!mesh) negative = 1;
                } else if (!mesh) {
                    positive = 1;
                }
            }
    return mesh ? 1 : (negative && positive);
}

/* ------------------------------------------------------------------------ */
/* Cell masks: DMCChunk::label_edges, DMCChunk.cpp:200-438.                  */
/* Four passes over label lines; a set label bit at z lands in cell z (bit   */
/* m) and cell z-1 (bit m+1), m = 0/2/4/6 for the (x,y), (x,y+1), (x+1,y),    */
/* (x+1,y+1) lines.  Pass 1 assigns, passes 2-4 OR in.  Masks are stored 8   */
/* cells per u64: [x][y][z/8], byte z%8.                                     */
/* ------------------------------------------------------------------------ */
static void mask_line(uint64_t* row_masks, uint32_t line, int zb, int m, int nwords_row)
{
    for (int p = 0; p < 32; p++) {
        if (!(line & (1u << p))) continue;
        int z = zb * 32 + p;
        int w = z >> 3, b = z & 7;
        if (w < nwords_row) row_masks[w] |= 1ull << (b * 8 + m);
        if (z > 0) {
            int w1 = (z - 1) >> 3, b1 = (z - 1) & 7;
            if (w1 < nwords_row) row_masks[w1] |= 1ull << (b1 * 8 + m + 1);
        }
    }
}

void or_cell_masks(const uint32_t* words, int dim, uint64_t* masks)
{
    const int zc = (dim + 31) / 32, z8 = (dim + 7) / 8;
    /* pass 1 assigns the whole row (the reference assigns every word it
     * visits and writes zeros past the row, which later rows overwrite). */
    for (int x = 0; x < dim; x++)
        for (int y = 0; y < dim; y++) {
            uint64_t* rm = masks + ((int64_t)x * dim + y) * z8;
            for (int w = 0; w < z8; w++) rm[w] = 0;
            for (int zb = 0; zb < zc; zb++) {
                uint32_t line = words[((int64_t)x * dim + y) * zc + zb];
                if (line) mask_line(rm, line, zb, 0, z8);
            }
        }
    for (int pass = 1; pass < 4; pass++) {
        int dx = pass >> 1, dy = pass & 1; /* pass 1: (0,1) m=2; 2: (1,0) m=4; 3: (1,1) m=6 */
        int m = 4 * dx + 2 * dy;
        for (int x = 0; x < dim - dx; x++)
            for (int y = 0; y < dim - dy; y++) {
                uint64_t* rm = masks + ((int64_t)x * dim + y) * z8;
                for (int zb = 0; zb < zc; zb++) {
                    uint32_t line = words[((int64_t)(x + dx) * dim + (y + dy)) * zc + zb];
                    if (line) mask_line(rm, line, zb, m, z8);
                }
            }
    }
}

/* ------------------------------------------------------------------------ */
/* Vertices + triangles: DMCChunk::label_edges cell loop (:440-498),         */
/* calculate_cell (:593-655), calculate_isovertex (:664-674),                */
/* polygonize/polygonize_cell (:514-576).                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t x, y, z;
    uint8_t mask;
    int64_t v[3]; /* vertex id of the owned x/y/z edge, or -1 */
} cell_rec;

int or_mesh_from_masks(const float* dens, const uint64_t* masks, int dim,
                       float* pos, uint8_t* boundary, uint32_t* edge, uint8_t* valence, int64_t capV,
                       uint32_t* idx, int64_t capI, int64_t* V_out, int64_t* nI_out)
{
    const int z8 = (dim + 7) / 8;
    const int64_t n3 = (int64_t)dim * dim * dim;
    int64_t* inds = (int64_t*)malloc(sizeof(int64_t) * n3);
    int64_t ccap = 1024, ncells = 0;
    cell_rec* cells = (cell_rec*)malloc(sizeof(cell_rec) * ccap);
    int64_t V = 0, nI = 0;
    int overflow = 0;
    for (int64_t i = 0; i < n3; i++) inds[i] = -1;

    for (int x = 0; x < dim; x++)
        for (int y = 0; y < dim; y++)
            for (int z = 0; z < dim; z += 8) {
                uint64_t mask = masks[((int64_t)x * dim + y) * z8 + z / 8];
                if (mask == 0 || mask == ~0ull) continue;
                for (int sz = 0; sz < 8 && z + sz < dim; sz++, mask >>= 8) {
                    uint8_t cm = (uint8_t)(mask & 0xFF);
                    if (cm == 0 || cm == 255) continue;
                    int cz = z + sz;
                    cell_rec c;
                    c.x = x; c.y = y; c.z = cz; c.mask = cm;
                    /* owned edges in x, y, z order (calculate_cell) */
                    const int ax_bit[3] = {4, 2, 1};
                    const int lim[3] = {x + 1 < dim, y + 1 < dim, cz + 1 < dim};
                    for (int a = 0; a < 3; a++) {
                        c.v[a] = -1;
                        if (((cm & 1) != ((cm >> ax_bit[a]) & 1)) && lim[a]) {
                            int x1 = x + (a == 0), y1 = y + (a == 1), z1 = cz + (a == 2);
                            float s0 = dens[((int64_t)x * dim + y) * dim + cz];
                            float s1 = dens[((int64_t)x1 * dim + y1) * dim + z1];
                            float mu = (0.0f - s0) / (s1 - s0);
                            if (V < capV) {
                                float p0[3] = {(float)x, (float)y, (float)cz};
                                float p1[3] = {(float)x1, (float)y1, (float)z1};
                                for (int k = 0; k < 3; k++) pos[3 * V + k] = (p1[k] - p0[k]) * mu + p0[k];
                                boundary[V] = x == 0 || y == 0 || cz == 0 || x == dim - 1 || y == dim - 1 ||
                                              cz == dim - 1 || x1 == dim - 1 || y1 == dim - 1 || z1 == dim - 1;
                                if (edge) edge[V] = (uint32_t)((((int64_t)x * dim + y) * dim + cz) * 3 + a);
                                valence[V] = 0;
                            } else {
                                overflow = 1;
                            }
                            c.v[a] = V++;
                        }
                    }
                    inds[((int64_t)x * dim + y) * dim + cz] = ncells;
                    if (ncells == ccap) {
                        ccap *= 2;
                        cells = (cell_rec*)realloc(cells, sizeof(cell_rec) * ccap);
                    }
                    cells[ncells++] = c;
                }
            }

    /* polygonize: interior cells in push order (DMCChunk.cpp:524-534) */
    for (int64_t ci = 0; ci < ncells; ci++) {
        const cell_rec* c = &cells[ci];
        int x = c->x, y = c->y, z = c->z;
        if (x >= dim - 1 || y >= dim - 1 || z >= dim - 1) continue;
        /* local edge -> owning cell offset and axis (polygonize_cell :543-565) */
        static const int eoff[12][4] = {
            {0, 0, 0, 0}, {0, 0, 1, 0}, {0, 1, 0, 0}, {0, 1, 1, 0},
            {0, 0, 0, 1}, {0, 0, 1, 1}, {1, 0, 0, 1}, {1, 0, 1, 1},
            {0, 0, 0, 2}, {0, 1, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}};
        uint64_t row = bmf_tri_table[c->mask];
        int n = (int)(row >> 60);
        for (int s = 0; s < n; s++) {
            int e = (int)((row >> (4 * s)) & 15);
            int64_t owner = inds[((int64_t)(x + eoff[e][0]) * dim + (y + eoff[e][1])) * dim + (z + eoff[e][2])];
            int64_t vid = owner < 0 ? -1 : cells[owner].v[eoff[e][3]];
            if (vid < 0) { overflow = 2; continue; } /* cannot happen for a consistent table */
            if (vid < capV) valence[vid]++;
            if (nI < capI) idx[nI] = (uint32_t)vid;
            else overflow = 1;
            nI++;
        }
    }
    free(inds);
    free(cells);
    *V_out = V;
    *nI_out = nI;
    return overflow == 0 ? 0 : (overflow == 1 ? -1 : -2);
}

/* ------------------------------------------------------------------------ */
/* Smoothing: MeshProcessor<3>::init/init_primitives (MeshProcessor.cpp:25-  */
/* 128), optimize_dual_grid (:130-236), optimize_primal_grid (:238-306), as  */
/* driven by ChunkGenerator.cpp:112-124: K dual steps, a primal step after   */
/* each of the first K-1, then one more primal step.  Adjacency lists hold   */
/* triangle ids in ascending order; sums start from +0 and run in list       */
/* order; boundary vertices stay put unless boundary_processing.             */
/* ------------------------------------------------------------------------ */
#define SMOOTH_BODY(T)                                                                       \
    if (iters <= 0 || V == 0 || nI < 3) return 0;                                            \
    int64_t nT = nI / 3;                                                                     \
    uint32_t* val = (uint32_t*)calloc((size_t)V, sizeof(uint32_t));                          \
    uint32_t* off = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(V + 1));                   \
    uint32_t* nxt = (uint32_t*)calloc((size_t)V, sizeof(uint32_t));                          \
    uint32_t* adj = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(3 * nT));                  \
    T* cen = (T*)malloc(sizeof(T) * (size_t)(3 * nT));                                       \
    for (int64_t i = 0; i < 3 * nT; i++) val[idx[i]]++;                                      \
    off[0] = 0;                                                                              \
    for (int64_t v = 0; v < V; v++) off[v + 1] = off[v] + val[v];                            \
    for (int64_t t = 0; t < nT; t++)                                                         \
        for (int k = 0; k < 3; k++) {                                                        \
            uint32_t v = idx[3 * t + k];                                                     \
            adj[off[v] + nxt[v]++] = (uint32_t)t;                                            \
        }                                                                                    \
    for (int m = 0; m < iters; m++) {                                                        \
        for (int64_t t = 0; t < nT; t++)                                                     \
            for (int k = 0; k < 3; k++) {                                                    \
                T acc = (T)0;                                                                \
                acc = acc + pos[3 * idx[3 * t + 0] + k];                                     \
                acc = acc + pos[3 * idx[3 * t + 1] + k];                                     \
                acc = acc + pos[3 * idx[3 * t + 2] + k];                                     \
                cen[3 * t + k] = acc / (T)3;                                                 \
            }                                                                                \
        /* primal: inside the loop for m < K-1, then once more after it */                   \
        for (int64_t v = 0; v < V; v++) {                                                    \
            if (nxt[v] == 0 || (!boundary_processing && boundary[v])) continue;              \
            T a[3] = {(T)0, (T)0, (T)0};                                                     \
            for (uint32_t q = off[v]; q < off[v] + nxt[v]; q++)                              \
                for (int k = 0; k < 3; k++) a[k] = a[k] + cen[3 * adj[q] + k];               \
            for (int k = 0; k < 3; k++) pos[3 * v + k] = a[k] / (T)nxt[v];                   \
        }                                                                                    \
    }                                                                                        \
    free(val); free(off); free(nxt); free(adj); free(cen);                                   \
    return 0;

int or_smooth_f32(float* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI,
                  int iters, int boundary_processing)
{
    SMOOTH_BODY(float)
}

int or_smooth_f64(double* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI,
                  int iters, int boundary_processing)
{
    SMOOTH_BODY(double)
}

/* ------------------------------------------------------------------------ */
/* One chunk through ChunkGenerator::extract_chunk (ChunkGenerator.cpp:80-   */
/* 147): label_grid -> label_edges -> polygonize -> MeshProcessor<3> when    */
/* K > 0 and the chunk has a mesh.                                           */
/* ------------------------------------------------------------------------ */
int or_extract_chunk(const or_sampler_t* s, const or_chunk_t* c, or_mesh_t* out, float* dens_out)
{
    const int dim = c->dim;
    const int64_t n3 = (int64_t)dim * dim * dim;
    const int zc = (dim + 31) / 32, z8 = (dim + 7) / 8;
    memset(out, 0, sizeof(*out));
    or_chunk_geometry(c, &out->delta, out->opos);
    float* dens = dens_out ? dens_out : (float*)malloc(sizeof(float) * n3);
    or_sample_block_f32(s, dim, out->opos, out->delta, dens);
    uint32_t* words = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)dim * dim * zc);
    out->contains_mesh = or_label_grid(dens, dim, words);
    if (out->contains_mesh) {
        uint64_t* masks = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)dim * dim * z8 + 4));
        or_cell_masks(words, dim, masks);
        int64_t capV = 4096, capI = 8192, V = 0, nI = 0;
        for (;;) {
            out->pos = (float*)malloc(sizeof(float) * 3 * capV);
            out->boundary = (uint8_t*)malloc((size_t)capV);
            out->edge = (uint32_t*)malloc(sizeof(uint32_t) * capV);
            out->idx = (uint32_t*)malloc(sizeof(uint32_t) * capI);
            uint8_t* valence = (uint8_t*)malloc((size_t)capV);
            int rc = or_mesh_from_masks(dens, masks, dim, out->pos, out->boundary, out->edge, valence, capV,
                                        out->idx, capI, &V, &nI);
            free(valence);
            if (rc == 0) break;
            free(out->pos); free(out->boundary); free(out->edge); free(out->idx);
            if (rc == -2) { free(masks); free(words); if (!dens_out) free(dens); return -2; }
            capV = V + 1;
            capI = nI + 1;
        }
        out->V = V;
        out->nI = nI;
        free(masks);
        if (c->iters > 0 && V > 0 && nI > 0)
            or_smooth_f32(out->pos, out->boundary, V, out->idx, nI, c->iters, c->boundary_processing);
    }
    free(words);
    if (!dens_out) free(dens);
    return 0;
}

void or_free_mesh(or_mesh_t* m)
{
    free(m->pos); free(m->boundary); free(m->edge); free(m->idx);
    m->pos = 0; m->boundary = 0; m->edge = 0; m->idx = 0;
}

int or_positions_f64(const or_sampler_t* s, const or_chunk_t* c, const or_mesh_t* topo, double* pos)
{
    const int dim = c->dim;
    double opos[3] = {topo->opos[0], topo->opos[1], topo->opos[2]};
    double delta = topo->delta;
    for (int64_t v = 0; v < topo->V; v++) {
        uint32_t e = topo->edge[v];
        int a = (int)(e % 3);
        int64_t p = e / 3;
        int z = (int)(p % dim), y = (int)((p / dim) % dim), x = (int)(p / ((int64_t)dim * dim));
        int x1 = x + (a == 0), y1 = y + (a == 1), z1 = z + (a == 2);
        double s0 = sample_f64(s, x * delta + opos[0], y * delta + opos[1], z * delta + opos[2]);
        double s1 = sample_f64(s, x1 * delta + opos[0], y1 * delta + opos[1], z1 * delta + opos[2]);
        double mu = -s0 / (s1 - s0);
        pos[3 * v + 0] = x + (a == 0 ? mu : 0.0);
        pos[3 * v + 1] = y + (a == 1 ? mu : 0.0);
        pos[3 * v + 2] = z + (a == 2 ? mu : 0.0);
    }
    if (c->iters > 0 && topo->V > 0 && topo->nI > 0)
        or_smooth_f64(pos, topo->boundary, topo->V, topo->idx, topo->nI, c->iters, c->boundary_processing);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Batch: #pragma omp parallel for over chunks (ChunkGenerator.cpp:91).      */
/* ------------------------------------------------------------------------ */
/* Order-sensitive checksum of a mesh: sum over the u32 stream (position
 * bit patterns, then indices) of word_k * (2k + 1), times an odd constant,
 * modulo 2^64. */
static uint64_t mesh_checksum(const float* pos, int64_t V, const uint32_t* idx, int64_t nI)
{
    uint64_t h = 0, k = 0;
    for (int64_t i = 0; i < 3 * V; i++, k++) {
        uint32_t b;
        memcpy(&b, &pos[i], 4);
        h += (uint64_t)b * (2 * k + 1);
    }
    for (int64_t i = 0; i < nI; i++, k++) h += (uint64_t)idx[i] * (2 * k + 1);
    return h * 0x9E3779B97F4A7C15ull;
}

int or_extract_batch(const or_sampler_t* s, const or_chunk_t* chunks, int n, int threads,
                     int64_t* V, int64_t* nI, int32_t* contains, uint64_t* checksum)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
    omp_set_max_active_levels(1);
#endif
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) {
        or_mesh_t m;
        or_extract_chunk(s, &chunks[i], &m, 0);
        uint64_t h = mesh_checksum(m.pos, m.V, m.idx, m.nI);
        V[i] = m.V;
        nI[i] = m.nI;
        contains[i] = m.contains_mesh;
        checksum[i] = h;
        or_free_mesh(&m);
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* QEF: qef_simd.h:392-579 restated per scalar lane.  The reference uses the */
/* approximate _mm_rsqrt_ps for the Givens cosine (qef_simd.h:196); here it  */
/* is an exact 1/sqrt.  The truncation band keeps 1/sigma when              */
/* min(|sigma|, |1/sigma|) >= tol (qef_simd.h:346-357); the error uses the   */
/* uncentred ATb (qef_simd.h:434-440, 457).                                  */
/* ------------------------------------------------------------------------ */
#define QEF_BODY(T, SQRT)                                                                     \
    if (count < 2 || count > 12) { out[0] = out[1] = out[2] = 0; return 0; }                  \
    T ata[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, atb[3] = {0, 0, 0}, pa[4] = {0, 0, 0, 0}; \
    for (int i = 0; i < count; i++) {                                                         \
        const T px = p[3 * i], py = p[3 * i + 1], pz = p[3 * i + 2];                          \
        const T nn[3] = {n[3 * i], n[3 * i + 1], n[3 * i + 2]};                               \
        for (int r = 0; r < 3; r++)                                                           \
            for (int c = 0; c < 3; c++) ata[r][c] = ata[r][c] + nn[r] * nn[c];                \
        T d = (px * nn[0] + py * nn[1]) + pz * nn[2];                                         \
        for (int r = 0; r < 3; r++) atb[r] = atb[r] + d * nn[r];                              \
        pa[0] = pa[0] + px; pa[1] = pa[1] + py; pa[2] = pa[2] + pz; pa[3] = pa[3] + (T)1;     \
    }                                                                                         \
    T mp[3] = {pa[0] / pa[3], pa[1] / pa[3], pa[2] / pa[3]};                                  \
    T rhs[3];                                                                                 \
    for (int c = 0; c < 3; c++)                                                               \
        rhs[c] = atb[c] - ((mp[0] * ata[0][c] + mp[1] * ata[1][c]) + mp[2] * ata[2][c]);      \
    /* Jacobi sweeps on the upper triangle (svd_solve_sym :304-342) */                        \
    T a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};                                   \
    memcpy(a, ata, sizeof(a));                                                                \
    static const int PA[3] = {0, 0, 1}, PB[3] = {1, 2, 2};                                    \
    for (int sw = 0; sw < 5; sw++)                                                            \
        for (int q = 0; q < 3; q++) {                                                         \
            int i0 = PA[q], i1 = PB[q];                                                       \
            if (a[i0][i1] == (T)0) continue;                                                  \
            T pp = a[i0][i0], pq = a[i0][i1], qq = a[i1][i1];                                 \
            T tau = (qq - pp) / (pq * (T)2);                                                  \
            T stt = SQRT(tau * tau + (T)1);                                                   \
            T tn = (T)1 / (tau >= (T)0 ? tau + stt : tau - stt);                              \
            T cs = (T)1 / SQRT((T)1 + tn * tn);                                               \
            T sn = tn * cs;                                                                   \
            /* rotateq_xy :215-260 */                                                         \
            T cc = cs * cs, ss = sn * sn, mx = (((T)2 * cs) * sn) * pq;                       \
            a[i0][i0] = (cc * pp - mx) + ss * qq;                                             \
            a[i1][i1] = (ss * pp + mx) + cc * qq;                                             \
            /* rotate_xy :264-300: columns i0,i1 of V and the remaining                       \
             * off-diagonal pair of the upper triangle */                                     \
            int ua_r = 0, ua_c = 3 - i1, va_r = 1 - i0, va_c = 2;                             \
            T u3 = a[ua_r][ua_c], v3 = a[va_r][va_c];                                         \
            for (int r = 0; r < 3; r++) {                                                     \
                T vu = v[r][i0], vv = v[r][i1];                                               \
                v[r][i0] = cs * vu - sn * vv;                                                 \
                v[r][i1] = sn * vu + cs * vv;                                                 \
            }                                                                                 \
            a[ua_r][ua_c] = cs * u3 - sn * v3;                                                \
            a[va_r][va_c] = sn * u3 + cs * v3;                                                \
            a[i0][i1] = (T)0;                                                                 \
        }                                                                                     \
    T sig[3] = {a[0][0], a[1][1], a[2][2]}, inv[3];                                           \
    for (int k = 0; k < 3; k++) {                                                             \
        T r1 = (T)1 / sig[k];                                                                 \
        T ax = sig[k] < 0 ? -sig[k] : sig[k], ar = r1 < 0 ? -r1 : r1;                         \
        inv[k] = ((ax < ar ? ax : ar) >= tol) ? r1 : (T)0;                                    \
    }                                                                                         \
    T pinv[3][3];                                                                             \
    for (int r = 0; r < 3; r++)                                                               \
        for (int c = 0; c < 3; c++)                                                           \
            pinv[r][c] = (v[c][0] * inv[0] * v[r][0] + v[c][1] * inv[1] * v[r][1]) +          \
                         v[c][2] * inv[2] * v[r][2];                                          \
    T x[3];                                                                                   \
    for (int c = 0; c < 3; c++) x[c] = (rhs[0] * pinv[0][c] + rhs[1] * pinv[1][c]) + rhs[2] * pinv[2][c]; \
    T e[3];                                                                                   \
    for (int c = 0; c < 3; c++)                                                               \
        e[c] = atb[c] - ((x[0] * ata[0][c] + x[1] * ata[1][c]) + x[2] * ata[2][c]);           \
    T err = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];                                        \
    out[0] = x[0] + mp[0]; out[1] = x[1] + mp[1]; out[2] = x[2] + mp[2];                      \
    return err;

float or_qef_solve_f32(const float* p, const float* n, int count, float tol, float out[3])
{
    QEF_BODY(float, sqrtf)
}

double or_qef_solve_f64(const double* p, const double* n, int count, double tol, double out[3])
{
    QEF_BODY(double, sqrt)
}

/* The sampler's FP32 value at one point (every octave of the terrain). */
float or_sample_value_f32(const or_sampler_t* s, float x, float y, float z)
{
    if (s->kind == OR_SAMPLER_TERRAIN) return or_terrain_value_f32(s, x, y, z);
    if (s->kind == OR_SAMPLER_CSG) return or_csg_value_f32(s, x, y, z, NULL);
    if (s->kind == OR_SAMPLER_NOISE) return or_noise_field_f32(s, x, y, z);
    return implicit_f32(s->kind, s->world_size, x, y, z);
}
