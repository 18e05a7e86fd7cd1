// Device density samplers.  FP32 with an exact, fixed operation sequence:
// the library is compiled with -ffp-contract=off and every fused multiply-add
// below is an explicit __builtin_fmaf, so the values are reproducible on any
// IEEE FP32 implementation that follows the same sequence.
//
// Terrain: d = amp * fbm(p) - (p.y - y0), the sampler family of
// NoiseSamplers::terrain3d_block (NoiseSampler.cpp:196-227) with the noise
// being the project's own FBM of 3D simplex noise (the reference calls
// FastNoiseSIMD there, which is not part of the reference tree).
//
// Simplex noise after S. Gustavson, "Simplex noise demystified" (2005):
// skew 1/3, unskew 1/6, kernel (0.5 - |d|^2)^4, gradient set and selection of
// K. Perlin, "Improving Noise" (2002), a lattice hash of 24-bit multiplies
// and xor-shifts (v_mul_u32_u24 runs at full rate), output x76.8 clamped to
// [-1, 1].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bmf_internal.h"

namespace bmf {

constexpr uint32_t kP1 = 0x5BD1E9u, kP2 = 0x27D4EBu, kP3 = 0x165667u;
constexpr uint32_t kM1 = 0x9E3779u, kM2 = 0x85EBCBu;
constexpr float kNoiseScale = 76.8f;
constexpr float kNoiseR2 = 0.5f;
constexpr float kCsgFloor = -16.0f;

__device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) { return __umul24(a, b); }

// mul24 by a constant as the full-rate v_mul_u32_u24 itself: where only the
// low bits of the product are used, the compiler otherwise drops the operand
// mask and emits the quarter-rate 32-bit v_mul_lo_u32 (same low 24 bits).
template <uint32_t M>
__device__ __forceinline__ uint32_t mul24_const(uint32_t a) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(M), "v"(a));
    return r;
}

// a ^ b ^ c as one three-input bitwise operation (truth table 0x96)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Gradient index of a lattice point from its combined coordinate hash
// h = mul24(i, P1) ^ mul24(j, P2) ^ mul24(k, P3) ^ seed.
__device__ __forceinline__ uint32_t grad_mix(uint32_t h) {
    h ^= h >> 13;
    h = mul24_const<kM1>(h ^ (h >> 24));
    h ^= h >> 11;
    h = mul24_const<kM2>(h ^ (h >> 24));
    h ^= h >> 16;
    return (h >> 4) & 15u;
}

__device__ __forceinline__ float grad_dot(uint32_t h, float x, float y, float z) {
    float u = h < 8u ? x : y;
    float v = h < 4u ? y : ((h == 12u || h == 14u) ? x : z);
    float su = (h & 1u) ? -u : u;
    float sv = (h & 2u) ? -v : v;
    return su + sv;
}

__device__ __forceinline__ float corner(float x, float y, float z, uint32_t h) {
    float t = __builtin_fmaf(-x, x, kNoiseR2);
    t = __builtin_fmaf(-y, y, t);
    t = __builtin_fmaf(-z, z, t);
    t = fmaxf(t, 0.0f);
    t = t * t;
    return (t * t) * grad_dot(h, x, y, z);
}

__device__ __forceinline__ float snoise3(float x, float y, float z, uint32_t seed) {
    const float F3 = 1.0f / 3.0f;
    const float G3 = 1.0f / 6.0f;
    const float G3x2 = 2.0f * G3;
    float s = ((x + y) + z) * F3;
    float fi = floorf(x + s), fj = floorf(y + s), fk = floorf(z + s);
    int32_t i = (int32_t)fi, j = (int32_t)fj, k = (int32_t)fk;
    float t = (float)(i + j + k) * G3;
    float x0 = x - (fi - t), y0 = y - (fj - t), z0 = z - (fk - t);
    // the rank flags stay lane masks (v_cmp results combined by scalar ops)
    const bool gxy = x0 >= y0, gyz = y0 >= z0, gxz = x0 >= z0;
    const bool i1 = gxy && gxz, j1 = !gxy && gyz, k1 = !gxz && !gyz;
    const bool i2 = gxy || gxz, j2 = !gxy || gyz, k2 = !gxz || !gyz;
    // x0 - (float)i1 is x0 - 1 or exactly x0 (i1 is 0 or 1), so the middle
    // corners select between x0 and the last corner's x0 - 1
    const float xm = x0 - 1.0f, ym = y0 - 1.0f, zm = z0 - 1.0f;
    float x1 = (i1 ? xm : x0) + G3, y1 = (j1 ? ym : y0) + G3, z1 = (k1 ? zm : z0) + G3;
    float x2 = (i2 ? xm : x0) + G3x2, y2 = (j2 ? ym : y0) + G3x2, z2 = (k2 ? zm : z0) + G3x2;
    float x3 = xm + 0.5f, y3 = ym + 0.5f, z3 = zm + 0.5f;
    // the corners' lattice coordinates are i or i + 1 per axis: six products
    // (the seed folded into the z terms) serve the four corner hashes
    const uint32_t a0 = mul24((uint32_t)i, kP1), a1 = mul24((uint32_t)(i + 1), kP1);
    const uint32_t b0 = mul24((uint32_t)j, kP2), b1 = mul24((uint32_t)(j + 1), kP2);
    const uint32_t c0 = mul24((uint32_t)k, kP3) ^ seed, c1 = mul24((uint32_t)(k + 1), kP3) ^ seed;
    float n0 = corner(x0, y0, z0, grad_mix(xor3(a0, b0, c0)));
    float n1 = corner(x1, y1, z1, grad_mix(xor3(i1 ? a1 : a0, j1 ? b1 : b0, k1 ? c1 : c0)));
    float n2 = corner(x2, y2, z2, grad_mix(xor3(i2 ? a1 : a0, j2 ? b1 : b0, k2 ? c1 : c0)));
    float n3 = corner(x3, y3, z3, grad_mix(xor3(a1, b1, c1)));
    float n = ((n0 + n1) + n2) + n3;
    float r = kNoiseScale * n;
    return fminf(fmaxf(r, -1.0f), 1.0f);
}

// FP32 operation count of one simplex evaluation as defined (the direct
// sequence; snoise3 shares a few of these between corners, with identical
// results), one per add/sub/mul/min/max/floor/convert/compare/select on
// float data, two per fma; the integer hash work is excluded (reported
// separately as kHashIntOps):
//   skew: 2 add + 1 mul = 3;  x+s: 3 add + 3 floor + 3 cvt = 9;
//   t: 1 cvt + 1 mul = 2;  x0: 6 sub;  rank compares: 3;
//   corners 1 and 2: 2 x (3 cvt + 3 sub + 3 add) = 18;  corner 3: 6;
//   per corner: 3 fma (6) + max + 2 mul + final mul + 4 selects + add = 16,
//   x 4 = 64;  sum 3 + scale 1 + clamp 2 = 6.   Total 117.
constexpr int kSimplexFlops = 117;
// Integer ops of the lattice hash per simplex evaluation as defined: 4
// corners x (3 mul24 + 3 xor for the combine, 2 mul24 + 6 xor/shift for the
// mix, 2 shift/and for the index) plus 9 adds for the corner lattice points
// (snoise3 shares the per-axis products between corners).
constexpr int kHashIntOps = 73;
// Per octave outside snoise3: 3 coordinate multiplies + the fma accumulate.
constexpr int kOctaveFlops = 5;
// Per terrain sample outside the octave loop: 3 x (cvt + mul + add) for the
// coordinate, normalise, 2 clamp, amp multiply, y - y0, final subtract.
constexpr int kTerrainFlops = 15;

// FBM of 3D simplex noise: sum of a_o * snoise3(p * f_o) (fused
// accumulate), normalised by the sum of the amplitudes, clamped to [-1, 1].
__device__ __forceinline__ float fbm_value(const float* fo, const float* ao, float inv_norm, int octaves,
                                           uint32_t seed, float x, float y, float z) {
    float acc = 0.0f;
    for (int o = 0; o < octaves; o++) {
        const float f = fo[o];
        float n = snoise3(x * f, y * f, z * f, seed + (uint32_t)o);
        acc = __builtin_fmaf(ao[o], n, acc);
    }
    float fb = acc * inv_norm;
    return fminf(fmaxf(fb, -1.0f), 1.0f);
}

__device__ __forceinline__ float terrain_density(const SamplerDev& S, float x, float y, float z) {
    const float fb = fbm_value(S.fo, S.ao, S.inv_norm, S.octaves, S.seed, x, y, z);
    float t1 = S.amp * fb;
    float t2 = y - S.y0;
    return t1 - t2;
}

// ImplicitFunctions (ImplicitSampler.cpp:27-65), positive inside.
__device__ __forceinline__ float implicit_density(int kind, float res, float x, float y, float z) {
    switch (kind) {
        case BMF_SAMPLER_SPHERE: {
            float r = res * 0.25f;
            float l = sqrtf((x * x + y * y) + z * z);
            return -(l - r);
        }
        case BMF_SAMPLER_CUBOID: {
            float r = res / 8.0f;
            float dx = fabsf(x) - r, dy = fabsf(y) - r, dz = fabsf(z) - r;
            float m = fmaxf(dx, fmaxf(dy, dz));
            float l = sqrtf((dx * dx + dy * dy) + dz * dz);
            return -fminf(m, l);
        }
        case BMF_SAMPLER_TORUS_Z: {
            float r1 = res / 4.0f, r2 = res / 10.0f;
            float qx = fabsf(sqrtf(x * x + y * y)) - r1;
            float len = sqrtf(qx * qx + z * z);
            return -(len - r2);
        }
        case BMF_SAMPLER_PLANE_Y:
            return -y;
    }
    return 0.0f;
}

// FP32 operations per primitive of the CSG union below (add, sub, mul, sqrt,
// abs, min/max each 1): sphere 3 sub + 3 mul + 2 add + sqrt + sub + max;
// box 3 sub + 3 abs + 3 sub + 3 max + 3 mul + 2 add + sqrt + 3 min/max + add
// + neg + max.
constexpr int kCsgSphereFlops = 11;
constexpr int kCsgBoxFlops = 24;

// CSG union of spheres r - |p - c| and exact boxes -sdf, floored.
__device__ __forceinline__ float csg_density(const SamplerDev& S, float x, float y, float z) {
    float best = kCsgFloor;
    for (int i = 0; i < S.n_spheres; i++) {
        const float* sp = S.spheres + 4 * i;
        float dx = x - sp[0], dy = y - sp[1], dz = z - sp[2];
        float l = sqrtf((dx * dx + dy * dy) + dz * dz);
        best = fmaxf(best, sp[3] - l);
    }
    for (int i = 0; i < S.n_boxes; i++) {
        const float* b = S.boxes + 6 * i;
        float px = x - b[0], py = y - b[1], pz = z - b[2];
        float qx = fabsf(px) - b[3], qy = fabsf(py) - b[4], qz = fabsf(pz) - b[5];
        float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
        float outside = sqrtf((mx * mx + my * my) + mz * mz);
        float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
        best = fmaxf(best, -(outside + inside));
    }
    return best;
}

// The same union over a chunk's culled primitive list: sphere ids, then box
// ids, in their original order.  Primitives left out cannot exceed the floor
// anywhere in the chunk, so the value is identical.
__device__ __forceinline__ float csg_density_list(const SamplerDev& S, const uint16_t* __restrict__ list, int ns,
                                                  int nb, float x, float y, float z) {
    float best = kCsgFloor;
    for (int k = 0; k < ns; k++) {
        const float* sp = S.spheres + 4 * list[k];
        float dx = x - sp[0], dy = y - sp[1], dz = z - sp[2];
        float l = sqrtf((dx * dx + dy * dy) + dz * dz);
        best = fmaxf(best, sp[3] - l);
    }
    for (int k = 0; k < nb; k++) {
        const float* b = S.boxes + 6 * list[ns + k];
        float px = x - b[0], py = y - b[1], pz = z - b[2];
        float qx = fabsf(px) - b[3], qy = fabsf(py) - b[4], qz = fabsf(pz) - b[5];
        float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
        float outside = sqrtf((mx * mx + my * my) + mz * mz);
        float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
        best = fmaxf(best, -(outside + inside));
    }
    return best;
}

// Primitives of a chunk's culled list staged in LDS: two float4 per
// primitive, (cx, cy, cz, r) and unused for a sphere, (cx, cy, cz, hx) and
// (hy, hz, -, -) for a box; spheres first.  Same union, same operation order.
constexpr int kCsgStage = 256;  // primitives per chunk staged in LDS (larger lists read global memory)

__device__ __forceinline__ void stage_csg(const SamplerDev& S, const uint16_t* __restrict__ list, int ns, int nb,
                                          float4* s_csg) {
    for (int k = threadIdx.x; k < ns + nb; k += blockDim.x) {
        if (k < ns) {
            const float* sp = S.spheres + 4 * list[k];
            s_csg[2 * k] = make_float4(sp[0], sp[1], sp[2], sp[3]);
            s_csg[2 * k + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            const float* b = S.boxes + 6 * list[k];
            s_csg[2 * k] = make_float4(b[0], b[1], b[2], b[3]);
            s_csg[2 * k + 1] = make_float4(b[4], b[5], 0.0f, 0.0f);
        }
    }
}

__device__ __forceinline__ float csg_density_lds(const float4* s_csg, int ns, int nb, float x, float y, float z) {
    float best = kCsgFloor;
    for (int k = 0; k < ns; k++) {
        const float4 sp = s_csg[2 * k];
        float dx = x - sp.x, dy = y - sp.y, dz = z - sp.z;
        float l = sqrtf((dx * dx + dy * dy) + dz * dz);
        best = fmaxf(best, sp.w - l);
    }
    for (int k = ns; k < ns + nb; k++) {
        const float4 b0 = s_csg[2 * k], b1 = s_csg[2 * k + 1];
        float px = x - b0.x, py = y - b0.y, pz = z - b0.z;
        float qx = fabsf(px) - b0.w, qy = fabsf(py) - b1.x, qz = fabsf(pz) - b1.y;
        float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
        float outside = sqrtf((mx * mx + my * my) + mz * mz);
        float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
        best = fmaxf(best, -(outside + inside));
    }
    return best;
}

// ---- noise fields: the NoiseSamplers block shapes (NoiseSampler.cpp:113-318)
// with the project's own noise.  At sample position p: c = g_scale p,
// perturbed; n = the fractal of the basis at c; d = -((p.y g_scale) y_scale)
// - height n.  A 2D field varies over (x, z) only.

// Lattice value in [-1, 1) of the corner whose per-axis hash products are
// a, b, c (c with the seed folded in): the simplex lattice hash mixed to 24
// bits k, then k / 2^23 - 1 (exact).
__device__ __forceinline__ float lattice_value(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t h = xor3(a, b, c);
    h ^= h >> 13;
    h = mul24_const<kM1>(h ^ (h >> 24));
    h ^= h >> 11;
    h = mul24_const<kM2>(h ^ (h >> 24));
    h ^= h >> 16;
    return (float)(h & 0xFFFFFFu) * 0x1p-23f - 1.0f;
}

// Quintic fade 6t^5 - 15t^4 + 10t^3 and the blend a + u (b - a).
__device__ __forceinline__ float fade5(float t) {
    const float s = __builtin_fmaf(__builtin_fmaf(t, 6.0f, -15.0f), t, 10.0f);
    return ((t * t) * t) * s;
}
__device__ __forceinline__ float blend(float a, float b, float u) { return __builtin_fmaf(u, b - a, a); }

// Value noise: the lattice values blended along x, then y, then z.
__device__ __forceinline__ float value3(float x, float y, float z, uint32_t seed) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int32_t i = (int32_t)fx, j = (int32_t)fy, k = (int32_t)fz;
    const float ux = fade5(x - fx), uy = fade5(y - fy), uz = fade5(z - fz);
    const uint32_t a0 = mul24((uint32_t)i, kP1), a1 = mul24((uint32_t)(i + 1), kP1);
    const uint32_t b0 = mul24((uint32_t)j, kP2), b1 = mul24((uint32_t)(j + 1), kP2);
    const uint32_t c0 = mul24((uint32_t)k, kP3) ^ seed, c1 = mul24((uint32_t)(k + 1), kP3) ^ seed;
    const float x00 = blend(lattice_value(a0, b0, c0), lattice_value(a1, b0, c0), ux);
    const float x10 = blend(lattice_value(a0, b1, c0), lattice_value(a1, b1, c0), ux);
    const float x01 = blend(lattice_value(a0, b0, c1), lattice_value(a1, b0, c1), ux);
    const float x11 = blend(lattice_value(a0, b1, c1), lattice_value(a1, b1, c1), ux);
    return blend(blend(x00, x10, uy), blend(x01, x11, uy), uz);
}
__device__ __forceinline__ float value2(float x, float z, uint32_t seed) {
    const float fx = floorf(x), fz = floorf(z);
    const int32_t i = (int32_t)fx, k = (int32_t)fz;
    const float ux = fade5(x - fx), uz = fade5(z - fz);
    const uint32_t a0 = mul24((uint32_t)i, kP1), a1 = mul24((uint32_t)(i + 1), kP1);
    const uint32_t c0 = mul24((uint32_t)k, kP3) ^ seed, c1 = mul24((uint32_t)(k + 1), kP3) ^ seed;
    const float x0 = blend(lattice_value(a0, 0u, c0), lattice_value(a1, 0u, c0), ux);
    const float x1 = blend(lattice_value(a0, 0u, c1), lattice_value(a1, 0u, c1), ux);
    return blend(x0, x1, uz);
}

__device__ __forceinline__ float noise_basis(const SamplerDev& S, float x, float y, float z, uint32_t seed) {
    if (S.noise_dims == 2) return S.noise_basis == BMF_NOISE_VALUE ? value2(x, z, seed) : snoise3(x, 0.0f, z, seed);
    return S.noise_basis == BMF_NOISE_VALUE ? value3(x, y, z, seed) : snoise3(x, y, z, seed);
}

// The fractal at c: FBM (normalised, clamped) or rigid t_0 - sum_{o>=1} a_o t_o,
// t_o = 1 - |b(c f_o)|.
__device__ __forceinline__ float noise_fractal(const SamplerDev& S, float cx, float cy, float cz) {
    if (S.noise_fractal == BMF_FRACTAL_RIGID) {
        float r = 0.0f;
        for (int o = 0; o < S.octaves; o++) {
            const float f = S.fo[o];
            const float t = 1.0f - fabsf(noise_basis(S, cx * f, cy * f, cz * f, S.seed + (uint32_t)o));
            r = o == 0 ? t : __builtin_fmaf(-t, S.ao[o], r);
        }
        return r;
    }
    float acc = 0.0f;
    for (int o = 0; o < S.octaves; o++) {
        const float f = S.fo[o];
        acc = __builtin_fmaf(S.ao[o], noise_basis(S, cx * f, cy * f, cz * f, S.seed + (uint32_t)o), acc);
    }
    return fminf(fmaxf(acc * S.inv_norm, -1.0f), 1.0f);
}

// c = g_scale p and its perturbation: GRADIENT adds, per octave k,
// pa_k (s(c pf_k, seed + 1000 + 3k + a))_a on the field's axes a;
// OFFSET_XZ adds perturb_amp w to x and z, w = the normalised FBM of s at
// p pf_k (one field for both axes, as windy3d_block's two offset sets).
__device__ __forceinline__ void noise_warp(const SamplerDev& S, float px, float py, float pz, float& cx, float& cy,
                                           float& cz) {
    cx = px * S.g_scale;
    cy = py * S.g_scale;
    cz = pz * S.g_scale;
    if (S.perturb == BMF_PERTURB_GRADIENT) {
        for (int k = 0; k < S.perturb_octaves; k++) {
            const float f = S.pf[k], a = S.pa[k];
            const uint32_t sk = S.seed + 1000u + 3u * (uint32_t)k;
            const float qx = cx * f, qy = cy * f, qz = cz * f;
            if (S.noise_dims == 2) {
                const float wx = snoise3(qx, 0.0f, qz, sk), wz = snoise3(qx, 0.0f, qz, sk + 2u);
                cx = __builtin_fmaf(a, wx, cx);
                cz = __builtin_fmaf(a, wz, cz);
            } else {
                const float wx = snoise3(qx, qy, qz, sk), wy = snoise3(qx, qy, qz, sk + 1u);
                const float wz = snoise3(qx, qy, qz, sk + 2u);
                cx = __builtin_fmaf(a, wx, cx);
                cy = __builtin_fmaf(a, wy, cy);
                cz = __builtin_fmaf(a, wz, cz);
            }
        }
    } else if (S.perturb == BMF_PERTURB_OFFSET_XZ) {
        float acc = 0.0f;
        for (int k = 0; k < S.perturb_octaves; k++) {
            const float f = S.pf[k];
            acc = __builtin_fmaf(S.pa[k], snoise3(px * f, py * f, pz * f, S.seed + 2000u + (uint32_t)k), acc);
        }
        const float w = fminf(fmaxf(acc * S.pinv, -1.0f), 1.0f);
        cx = __builtin_fmaf(S.perturb_amp, w, cx);
        cz = __builtin_fmaf(S.perturb_amp, w, cz);
    }
}

// n at a sample (a 2D field's n depends on x and z only).
__device__ __forceinline__ float noise_value(const SamplerDev& S, float px, float py, float pz) {
    float cx, cy, cz;
    noise_warp(S, px, py, pz, cx, cy, cz);
    return noise_fractal(S, cx, cy, cz);
}

__device__ __forceinline__ float noise_density_from(const SamplerDev& S, float n, float py) {
    const float dy = (py * S.g_scale) * S.y_scale;
    return (-dy) - n * S.height;
}

__device__ __forceinline__ float sample_density(const SamplerDev& S, float x, float y, float z) {
    if (S.kind == BMF_SAMPLER_TERRAIN) return terrain_density(S, x, y, z);
    if (S.kind == BMF_SAMPLER_CSG) return csg_density(S, x, y, z);
    if (S.kind == BMF_SAMPLER_NOISE) return noise_density_from(S, noise_value(S, x, y, z), y);
    return implicit_density(S.kind, S.world_size, x, y, z);
}

}  // namespace bmf
