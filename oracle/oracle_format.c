/*
 * oracle_format.c -- CPU restatement of the output formatting of a batch
 * (test infrastructure only; never linked into the product).
 *
 * WorldOctreeNode::format (WorldOctreeNode.cpp:68-86) -> GLChunk::format_data
 * with unwind_verts = false (GLChunk.cpp:278-297) copies each meshed chunk's
 * vertex positions, normals and colours into the VBO arrays, and the vertex
 * shader places a vertex at p * scale + overlap_pos (shaders/main_vs.glsl:34,
 * chunk_pos = (overlap_pos, scale)); DMCChunk::copy_verts_and_inds
 * (DMCChunk.cpp:578-591) concatenates the chunks with contains_mesh in batch
 * order and adds each chunk's first vertex to its indices.
 *
 * Vertex colours: ColorMapper::generate_colors (ColorMapper.cpp:15-62) takes
 * a 4-octave simplex FBM at the vertex position (scale 1) and maps
 * n = 4 * noise to hsl_to_rgb((n + 1) * 0.5 * 360, 0.72, 1.0) (:60-119, an HSV
 * conversion).  The reference's noise is FastNoiseSIMD's SimplexFractal,
 * which is not in the reference tree; here it is the project's own FBM of 3D
 * simplex noise (oracle.c) with FastNoiseSIMD's default frequency 0.01,
 * lacunarity 2, gain 0.5 and seed 1337 and ColorMapper's 4 octaves.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

/* ColorMapper::hsl_to_rgb (ColorMapper.cpp:60-119) in FP32, s > 0. */
void or_hsv_to_rgb_f32(float h, float s, float v, float out[3])
{
    float hh = fmodf(fabsf(h), 360.0f);
    hh = hh / 60.0f;
    const int i = (int)hh;
    const float ff = hh - (float)i;
    const float p = v * (1.0f - s);
    const float q = v * (1.0f - (s * ff));
    const float t = v * (1.0f - (s * (1.0f - ff)));
    float r, g, b;
    switch (i) {
        case 0: r = v; g = t; b = p; break;
        case 1: r = q; g = v; b = p; break;
        case 2: r = p; g = v; b = t; break;
        case 3: r = p; g = q; b = v; break;
        case 4: r = t; g = p; b = v; break;
        default: r = v; g = p; b = q; break;
    }
    out[0] = r;
    out[1] = g;
    out[2] = b;
}

/* The colour noise's FBM: oracle.c's terrain FBM with these parameters. */
static void color_sampler(or_sampler_t* s)
{
    memset(s, 0, sizeof(*s));
    s->kind = OR_SAMPLER_TERRAIN;
    s->freq = 0.01f;
    s->lacunarity = 2.0f;
    s->gain = 0.5f;
    s->octaves = 4;
    s->seed = 1337u;
    s->amp = 1.0f;
}

float or_color_noise_f32(float x, float y, float z)
{
    or_sampler_t s;
    color_sampler(&s);
    float fo[32], ao[32], inv;
    or_fbm_params(&s, fo, ao, &inv);
    return or_fbm_f32(&s, fo, ao, inv, x, y, z);
}

void or_vertex_color_f32(const float p[3], float rgb[3])
{
    const float n = or_color_noise_f32(p[0], p[1], p[2]) * 4.0f;
    or_hsv_to_rgb_f32(((n + 1.0f) * 0.5f) * 360.0f, 0.72f, 1.0f, rgb);
}

/* One batch, formatted.  Inputs per chunk j (n chunks): its vertex and index
 * counts, contains_mesh, delta and overlap_pos, and its vertices / indices
 * at offsets voff[j] / ioff[j] of pos [.][3], nrm [.][3] (may be NULL) and
 * idx.  space: 0 grid units, 1 world.  colors: 0 none, 1 white (DualVertex
 * default, DMCChunk.cpp:681), 2 noise.  global: add the chunk's first
 * formatted vertex to its indices.  Writes the formatted arrays and returns
 * the totals. */
void or_format_batch(int n, const int64_t* V, const int64_t* I, const int32_t* contains_mesh, const float* delta,
                     const float* opos, const int64_t* voff, const int64_t* ioff, const float* pos,
                     const float* nrm, const uint32_t* idx, int space, int colors, int global, float* out_pos,
                     float* out_col, float* out_nrm, uint32_t* out_idx, int64_t* total_v, int64_t* total_i)
{
    int64_t tv = 0, ti = 0;
    for (int j = 0; j < n; j++) {
        if (!contains_mesh[j]) continue;
        for (int64_t v = 0; v < V[j]; v++) {
            const float* p = pos + 3 * (voff[j] + v);
            float* o = out_pos + 3 * (tv + v);
            for (int k = 0; k < 3; k++) {
                if (space == 1) {
                    const float m = p[k] * delta[j];
                    o[k] = m + opos[3 * j + k];
                } else {
                    o[k] = p[k];
                }
            }
            if (colors == 1) {
                for (int k = 0; k < 3; k++) out_col[3 * (tv + v) + k] = 1.0f;
            } else if (colors == 2) {
                or_vertex_color_f32(p, out_col + 3 * (tv + v));
            }
            if (nrm && out_nrm)
                for (int k = 0; k < 3; k++) out_nrm[3 * (tv + v) + k] = nrm[3 * (voff[j] + v) + k];
        }
        for (int64_t i = 0; i < I[j]; i++) out_idx[ti + i] = idx[ioff[j] + i] + (global ? (uint32_t)tv : 0u);
        tv += V[j];
        ti += I[j];
    }
    *total_v = tv;
    *total_i = ti;
}
