/*
 * Noise-field samplers of the oracle -- TEST INFRASTRUCTURE ONLY (tests/,
 * the smoke check and bench.py's CPU baseline load it; the product never does).
 *
 * The shapes are the reference's NoiseSamplers block functions
 * (NoiseSampler.cpp:113-318): terrain2d_block and terrain2d_pert_block
 * evaluate a noise over (x, z) only (a NOISE_BLOCK one sample high) and
 * broadcast it along y, terrain3d_block, terrain3d_pert_block and
 * windy3d_block over the volume, and every one returns
 * -dy - n * nm with dy = ((float)iy * scale + p.y) * g_scale (* ym).  Their
 * noise comes from FastNoiseSIMD, which is not in the reference tree, so
 * the noise here is the project's own, pinned to its definition
 * (tests/test_oracle_noise_field.py):
 *
 *   sample position p (two roundings, as every sampler); c = g_scale * p;
 *   perturbation of c (none; GRADIENT: per octave k, frequency pf_k =
 *     perturb_freq * 2^k and amplitude pa_k = perturb_amp * perturb_gain^k,
 *     c += pa_k * (s(c pf_k, seed + 1000 + 3k + a))_a for the axes a the
 *     field varies in, s = the 3D simplex noise of oracle.c (at y = 0 for a
 *     2D field); OFFSET_XZ (3D only): w = the normalised FBM of s at
 *     p * perturb_freq 2^k with gains perturb_gain^k and seed + 2000 + k,
 *     then c.x += perturb_amp w and c.z += perturb_amp w -- one field for
 *     both axes, as windy3d_block fills its x and y offset sets with the same
 *     settings);
 *   basis b: VALUE = lattice values in [-1, 1) (the simplex lattice hash
 *     mixed to 24 bits, k / 2^23 - 1) blended with the quintic fade
 *     6t^5 - 15t^4 + 10t^3, x then y then z; SIMPLEX = oracle.c's simplex
 *     (at y = 0 for a 2D field);
 *   fractal: FBM = the normalised, clamped sum of a_o b(c f_o, seed + o) as
 *     the terrain's FBM; RIGID = t_0 - sum_{o >= 1} a_o t_o with
 *     t_o = 1 - |b(c f_o, seed + o)| (f_o = freq L^o, a_o = gain^o);
 *   d = -((p.y * g_scale) * y_scale) - n * height.
 *
 * FP32 with the device's operation order (explicit fmaf, no contraction);
 * the FP64 twins evaluate the same definition in double.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

#define NZ_P1 0x5BD1E9u
#define NZ_P2 0x27D4EBu
#define NZ_P3 0x165667u
#define NZ_M1 0x9E3779u
#define NZ_M2 0x85EBCBu
#define OR_MAX_PERTURB 24

static inline uint32_t mul24(uint32_t a, uint32_t b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }

static inline uint32_t value_mix(uint32_t h)
{
    h ^= h >> 13;
    h = mul24(h ^ (h >> 24), NZ_M1);
    h ^= h >> 11;
    h = mul24(h ^ (h >> 24), NZ_M2);
    h ^= h >> 16;
    return h & 0xFFFFFFu;
}

/* Lattice value of (i, j, k): k / 2^23 - 1 for the 24-bit mixed hash k. */
float or_lattice_value(int32_t i, int32_t j, int32_t k, uint32_t seed)
{
    const uint32_t h = mul24((uint32_t)i, NZ_P1) ^ mul24((uint32_t)j, NZ_P2) ^ mul24((uint32_t)k, NZ_P3) ^ seed;
    return (float)value_mix(h) * 0x1p-23f - 1.0f;
}

static inline float fade5_f32(float t)
{
    const float s = fmaf(fmaf(t, 6.0f, -15.0f), t, 10.0f);
    return ((t * t) * t) * s;
}

static inline float lerp_f32(float a, float b, float u) { return fmaf(u, b - a, a); }

float or_value3_f32(float x, float y, float z, uint32_t seed)
{
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int32_t i = (int32_t)fx, j = (int32_t)fy, k = (int32_t)fz;
    const float ux = fade5_f32(x - fx), uy = fade5_f32(y - fy), uz = fade5_f32(z - fz);
    const float v000 = or_lattice_value(i, j, k, seed), v100 = or_lattice_value(i + 1, j, k, seed);
    const float v010 = or_lattice_value(i, j + 1, k, seed), v110 = or_lattice_value(i + 1, j + 1, k, seed);
    const float v001 = or_lattice_value(i, j, k + 1, seed), v101 = or_lattice_value(i + 1, j, k + 1, seed);
    const float v011 = or_lattice_value(i, j + 1, k + 1, seed), v111 = or_lattice_value(i + 1, j + 1, k + 1, seed);
    const float a00 = lerp_f32(v000, v100, ux), a10 = lerp_f32(v010, v110, ux);
    const float a01 = lerp_f32(v001, v101, ux), a11 = lerp_f32(v011, v111, ux);
    const float b0 = lerp_f32(a00, a10, uy), b1 = lerp_f32(a01, a11, uy);
    return lerp_f32(b0, b1, uz);
}

float or_value2_f32(float x, float z, uint32_t seed)
{
    const float fx = floorf(x), fz = floorf(z);
    const int32_t i = (int32_t)fx, k = (int32_t)fz;
    const float ux = fade5_f32(x - fx), uz = fade5_f32(z - fz);
    const float v00 = or_lattice_value(i, 0, k, seed), v10 = or_lattice_value(i + 1, 0, k, seed);
    const float v01 = or_lattice_value(i, 0, k + 1, seed), v11 = or_lattice_value(i + 1, 0, k + 1, seed);
    const float a0 = lerp_f32(v00, v10, ux), a1 = lerp_f32(v01, v11, ux);
    return lerp_f32(a0, a1, uz);
}

static inline double fade5_f64(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
static inline double lerp_f64(double a, double b, double u) { return a + u * (b - a); }

double or_value3_f64(double x, double y, double z, uint32_t seed)
{
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const int32_t i = (int32_t)fx, j = (int32_t)fy, k = (int32_t)fz;
    const double ux = fade5_f64(x - fx), uy = fade5_f64(y - fy), uz = fade5_f64(z - fz);
    double v[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) v[a][b][c] = (double)or_lattice_value(i + a, j + b, k + c, seed);
    const double b0 = lerp_f64(lerp_f64(v[0][0][0], v[1][0][0], ux), lerp_f64(v[0][1][0], v[1][1][0], ux), uy);
    const double b1 = lerp_f64(lerp_f64(v[0][0][1], v[1][0][1], ux), lerp_f64(v[0][1][1], v[1][1][1], ux), uy);
    return lerp_f64(b0, b1, uz);
}

double or_value2_f64(double x, double z, uint32_t seed)
{
    const double fx = floor(x), fz = floor(z);
    const int32_t i = (int32_t)fx, k = (int32_t)fz;
    const double ux = fade5_f64(x - fx), uz = fade5_f64(z - fz);
    const double a0 = lerp_f64(or_lattice_value(i, 0, k, seed), or_lattice_value(i + 1, 0, k, seed), ux);
    const double a1 = lerp_f64(or_lattice_value(i, 0, k + 1, seed), or_lattice_value(i + 1, 0, k + 1, seed), ux);
    return lerp_f64(a0, a1, uz);
}

/* Per-sampler constants: FP32 product chains as the library builds them. */
typedef struct {
    float fo[32], ao[32], inv_norm;
    float pf[OR_MAX_PERTURB], pa[OR_MAX_PERTURB], pinv;
} noise_consts;

static void noise_params(const or_sampler_t* s, noise_consts* k)
{
    or_fbm_params(s, k->fo, k->ao, &k->inv_norm);
    float f = s->perturb_freq, a = s->perturb == OR_PERTURB_GRADIENT ? s->perturb_amp : 1.0f, norm = 0.0f;
    for (int q = 0; q < s->perturb_octaves && q < OR_MAX_PERTURB; q++) {
        k->pf[q] = f;
        k->pa[q] = a;
        norm = norm + a;
        f = f * 2.0f;
        a = a * s->perturb_gain;
    }
    k->pinv = norm > 0.0f ? 1.0f / norm : 0.0f;
}

static inline float basis_f32(const or_sampler_t* s, float x, float y, float z, uint32_t seed)
{
    if (s->noise_dims == 2)
        return s->noise_basis == OR_NOISE_VALUE ? or_value2_f32(x, z, seed) : or_snoise3_f32(x, 0.0f, z, seed);
    return s->noise_basis == OR_NOISE_VALUE ? or_value3_f32(x, y, z, seed) : or_snoise3_f32(x, y, z, seed);
}

static float fractal_f32(const or_sampler_t* s, const noise_consts* k, float cx, float cy, float cz)
{
    if (s->noise_fractal == OR_FRACTAL_RIGID) {
        float r = 0.0f;
        for (int o = 0; o < s->octaves; o++) {
            const float f = k->fo[o];
            const float t = 1.0f - fabsf(basis_f32(s, cx * f, cy * f, cz * f, s->seed + (uint32_t)o));
            r = o == 0 ? t : fmaf(-t, k->ao[o], r);
        }
        return r;
    }
    float acc = 0.0f;
    for (int o = 0; o < s->octaves; o++) {
        const float f = k->fo[o];
        acc = fmaf(k->ao[o], basis_f32(s, cx * f, cy * f, cz * f, s->seed + (uint32_t)o), acc);
    }
    return fminf(fmaxf(acc * k->inv_norm, -1.0f), 1.0f);
}

static void warp_f32(const or_sampler_t* s, const noise_consts* k, float px, float py, float pz, float c[3])
{
    c[0] = px * s->g_scale;
    c[1] = py * s->g_scale;
    c[2] = pz * s->g_scale;
    if (s->perturb == OR_PERTURB_GRADIENT) {
        for (int q = 0; q < s->perturb_octaves; q++) {
            const float f = k->pf[q], a = k->pa[q];
            const uint32_t sq = s->seed + 1000u + 3u * (uint32_t)q;
            const float qx = c[0] * f, qy = c[1] * f, qz = c[2] * f;
            if (s->noise_dims == 2) {
                const float wx = or_snoise3_f32(qx, 0.0f, qz, sq), wz = or_snoise3_f32(qx, 0.0f, qz, sq + 2u);
                c[0] = fmaf(a, wx, c[0]);
                c[2] = fmaf(a, wz, c[2]);
            } else {
                const float wx = or_snoise3_f32(qx, qy, qz, sq), wy = or_snoise3_f32(qx, qy, qz, sq + 1u);
                const float wz = or_snoise3_f32(qx, qy, qz, sq + 2u);
                c[0] = fmaf(a, wx, c[0]);
                c[1] = fmaf(a, wy, c[1]);
                c[2] = fmaf(a, wz, c[2]);
            }
        }
    } else if (s->perturb == OR_PERTURB_OFFSET_XZ) {
        float acc = 0.0f;
        for (int q = 0; q < s->perturb_octaves; q++) {
            const float f = k->pf[q];
            acc = fmaf(k->pa[q], or_snoise3_f32(px * f, py * f, pz * f, s->seed + 2000u + (uint32_t)q), acc);
        }
        const float w = fminf(fmaxf(acc * k->pinv, -1.0f), 1.0f);
        c[0] = fmaf(s->perturb_amp, w, c[0]);
        c[2] = fmaf(s->perturb_amp, w, c[2]);
    }
}

static float noise_n_f32(const or_sampler_t* s, const noise_consts* k, float x, float y, float z)
{
    float c[3];
    warp_f32(s, k, x, y, z, c);
    return fractal_f32(s, k, c[0], c[1], c[2]);
}

static inline float density_from_f32(const or_sampler_t* s, float n, float y)
{
    const float dy = (y * s->g_scale) * s->y_scale;
    return (-dy) - n * s->height;
}

float or_noise_fractal_f32(const or_sampler_t* s, float x, float y, float z)
{
    noise_consts k;
    noise_params(s, &k);
    return noise_n_f32(s, &k, x, y, z);
}

float or_noise_field_f32(const or_sampler_t* s, float x, float y, float z)
{
    return density_from_f32(s, or_noise_fractal_f32(s, x, y, z), y);
}

/* Block of a noise field: a 2D field evaluates its noise once per (x, z)
 * column (the same FP32 value at every y of the column). */
void or_noise_block_f32(const or_sampler_t* s, int dim, const float opos[3], float delta, float* out)
{
    noise_consts k;
    noise_params(s, &k);
    for (int ix = 0; ix < dim; ix++) {
        const float x = (float)ix * delta + opos[0];
        for (int iz = 0; iz < dim; iz++) {
            const float z = (float)iz * delta + opos[2];
            const float ncol = s->noise_dims == 2 ? noise_n_f32(s, &k, x, 0.0f, z) : 0.0f;
            for (int iy = 0; iy < dim; iy++) {
                const float y = (float)iy * delta + opos[1];
                const float n = s->noise_dims == 2 ? ncol : noise_n_f32(s, &k, x, y, z);
                out[((int64_t)ix * dim + iy) * dim + iz] = density_from_f32(s, n, y);
            }
        }
    }
}

/* ---- FP64 twin of the definition ---- */
static inline double basis_f64(const or_sampler_t* s, double x, double y, double z, uint32_t seed)
{
    if (s->noise_dims == 2)
        return s->noise_basis == OR_NOISE_VALUE ? or_value2_f64(x, z, seed) : or_snoise3_f64(x, 0.0, z, seed);
    return s->noise_basis == OR_NOISE_VALUE ? or_value3_f64(x, y, z, seed) : or_snoise3_f64(x, y, z, seed);
}

double or_noise_field_f64(const or_sampler_t* s, double x, double y, double z)
{
    double c[3] = {x * s->g_scale, y * s->g_scale, z * s->g_scale};
    if (s->perturb == OR_PERTURB_GRADIENT) {
        double f = s->perturb_freq, a = s->perturb_amp;
        for (int q = 0; q < s->perturb_octaves; q++) {
            const uint32_t sq = s->seed + 1000u + 3u * (uint32_t)q;
            const double qx = c[0] * f, qy = c[1] * f, qz = c[2] * f;
            if (s->noise_dims == 2) {
                const double wx = or_snoise3_f64(qx, 0.0, qz, sq), wz = or_snoise3_f64(qx, 0.0, qz, sq + 2u);
                c[0] += a * wx;
                c[2] += a * wz;
            } else {
                const double wx = or_snoise3_f64(qx, qy, qz, sq), wy = or_snoise3_f64(qx, qy, qz, sq + 1u);
                const double wz = or_snoise3_f64(qx, qy, qz, sq + 2u);
                c[0] += a * wx;
                c[1] += a * wy;
                c[2] += a * wz;
            }
            f *= 2.0;
            a *= s->perturb_gain;
        }
    } else if (s->perturb == OR_PERTURB_OFFSET_XZ) {
        double f = s->perturb_freq, a = 1.0, acc = 0.0, norm = 0.0;
        for (int q = 0; q < s->perturb_octaves; q++) {
            acc += a * or_snoise3_f64(x * f, y * f, z * f, s->seed + 2000u + (uint32_t)q);
            norm += a;
            f *= 2.0;
            a *= s->perturb_gain;
        }
        double w = norm > 0.0 ? acc / norm : 0.0;
        w = w < -1.0 ? -1.0 : (w > 1.0 ? 1.0 : w);
        c[0] += s->perturb_amp * w;
        c[2] += s->perturb_amp * w;
    }
    double f = s->freq, a = 1.0, n;
    if (s->noise_fractal == OR_FRACTAL_RIGID) {
        n = 0.0;
        for (int o = 0; o < s->octaves; o++) {
            const double t = 1.0 - fabs(basis_f64(s, c[0] * f, c[1] * f, c[2] * f, s->seed + (uint32_t)o));
            n = o == 0 ? t : n - a * t;
            f *= s->lacunarity;
            a *= s->gain;
        }
    } else {
        double acc = 0.0, norm = 0.0;
        for (int o = 0; o < s->octaves; o++) {
            acc += a * basis_f64(s, c[0] * f, c[1] * f, c[2] * f, s->seed + (uint32_t)o);
            norm += a;
            f *= s->lacunarity;
            a *= s->gain;
        }
        n = acc / norm;
        n = n < -1.0 ? -1.0 : (n > 1.0 ? 1.0 : n);
    }
    return -(y * s->g_scale * s->y_scale) - n * s->height;
}
