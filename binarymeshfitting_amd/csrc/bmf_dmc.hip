// Quad mode of the extraction path (BMF_MESH_QUADS): dual marching cubes
// over the reference's polygon table (Tables.hpp:96-354) with mass-point or
// QEF vertex placement (qef_simd.h:392-579), quads around the interior
// sign-changing edges, and MeshProcessor<4>/<3> smoothing over per-vertex
// rings.
//
// Per batch (after sample, count and refine, which in this mode count per
// label row the polygons of the row's cells and the quads of the row's
// interior sign-changing edges, and leave every edge endpoint's density
// exact):
//   (k_scan)       row bases, arena offsets
//   k_dmc_vertices per row: vertex positions; per surface cell its first
//                  vertex id and mask (cell_tb / cell_mask arenas); for the
//                  chunks that smooth, each vertex's ring (the primitives
//                  around it in ascending id, the reference's order), from
//                  the quad ids of its polygon's edges
//   k_dmc_quads    per row: quad (or two-triangle) indices
//   K x (k_dual_n, k_primal_ring)
#include <hip/hip_runtime.h>

#include "bmf_device_tables.h"
#include "bmf_internal.h"
#include "bmf_rows.h"
#include "samplers.h"

namespace bmf {

namespace {

// Per block: polygon count of each mask, and for each mask and local edge
// the polygon holding it (nibbles, 0xF = none).
struct DmcTables {
    uint8_t np[256];
    uint64_t pe[256];
};

__device__ void load_dmc_tables(DmcTables& t) {
    for (int m = threadIdx.x; m < 256; m += blockDim.x) {
        const uint64_t row = c_dmc_table[m];
        const int n = (int)(row >> 60);
        uint64_t pe = ~0ull;
        int p = 0, any = 0;
        for (int s = 0; s < n; s++) {
            const int e = (int)((row >> (4 * s)) & 15ull);
            if (e == 15) {
                p++;
                any = 0;
                continue;
            }
            pe = (pe & ~(15ull << (4 * e))) | ((uint64_t)p << (4 * e));
            any = 1;
        }
        t.np[m] = (uint8_t)(p + any);
        t.pe[m] = pe;
    }
}

// Sum over the wave of a per-lane count c < 8 (bit-sliced ballots), and the
// exclusive prefix of the lanes below.
__device__ __forceinline__ uint32_t wave_sum3(uint32_t c, uint64_t lt, uint32_t& pre) {
    const uint64_t b0 = __ballot(c & 1u), b1 = __ballot(c & 2u), b2 = __ballot(c & 4u);
    pre = (uint32_t)__popcll(b0 & lt) + 2u * (uint32_t)__popcll(b1 & lt) + 4u * (uint32_t)__popcll(b2 & lt);
    return (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
}

// ------------------------------------------------------------------------
// CSG gradient and QEF (FP32, the CPU oracle's operation sequence)
// ------------------------------------------------------------------------
// Gradient of the winning primitive of the max-union (first index on ties),
// zero on the floor.
// The primitives are the chunk's culled list (sphere ids, then box ids, in
// order) when there is one: the others never exceed the floor in the chunk.
__device__ void csg_gradient(const SamplerDev& S, const uint16_t* list, int ns, int nb, float x, float y, float z,
                             float g[3]) {
    float best = kCsgFloor;
    g[0] = g[1] = g[2] = 0.0f;
    if (!list) {
        ns = S.n_spheres;
        nb = S.n_boxes;
    }
    for (int k = 0; k < ns; k++) {
        const float* sp = S.spheres + 4 * (list ? list[k] : k);
        const float dx = x - sp[0], dy = y - sp[1], dz = z - sp[2];
        const float l = sqrtf((dx * dx + dy * dy) + dz * dz);
        const float v = sp[3] - l;
        if (v > best) {
            best = v;
            const float inv = l > 0.0f ? 1.0f / l : 0.0f;
            g[0] = -dx * inv;
            g[1] = -dy * inv;
            g[2] = -dz * inv;
        }
    }
    for (int k = 0; k < nb; k++) {
        const float* b = S.boxes + 6 * (list ? list[ns + k] : k);
        const float px = x - b[0], py = y - b[1], pz = z - b[2];
        const float qx = fabsf(px) - b[3], qy = fabsf(py) - b[4], qz = fabsf(pz) - b[5];
        const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
        const float outside = sqrtf((mx * mx + my * my) + mz * mz);
        const float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
        const float v = -(outside + inside);
        if (v > best) {
            best = v;
            const float sx = px < 0.0f ? -1.0f : 1.0f, sy = py < 0.0f ? -1.0f : 1.0f, sz = pz < 0.0f ? -1.0f : 1.0f;
            if (outside > 0.0f) {
                const float inv = 1.0f / outside;
                g[0] = -mx * sx * inv;
                g[1] = -my * sy * inv;
                g[2] = -mz * sz * inv;
            } else {
                g[0] = g[1] = g[2] = 0.0f;
                if (qx >= qy && qx >= qz) g[0] = -sx;
                else if (qy >= qz) g[1] = -sy;
                else g[2] = -sz;
            }
        }
    }
}

// The same with the chunk's primitives staged in LDS (stage_csg layout):
// the loop finds the winning primitive (first index on ties), whose
// gradient is then formed once, with the same operations as above.
__device__ void csg_gradient_lds(const float4* s_csg, int ns, int nb, float x, float y, float z, float g[3]) {
    float best = kCsgFloor;
    int win = -1;
    for (int k = 0; k < ns; k++) {
        const float4 sp = s_csg[2 * k];
        const float dx = x - sp.x, dy = y - sp.y, dz = z - sp.z;
        const float l = sqrtf((dx * dx + dy * dy) + dz * dz);
        const float v = sp.w - l;
        if (v > best) {
            best = v;
            win = k;
        }
    }
    for (int k = ns; k < ns + nb; k++) {
        const float4 b0 = s_csg[2 * k], b1 = s_csg[2 * k + 1];
        const float px = x - b0.x, py = y - b0.y, pz = z - b0.z;
        const float qx = fabsf(px) - b0.w, qy = fabsf(py) - b1.x, qz = fabsf(pz) - b1.y;
        const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
        const float outside = sqrtf((mx * mx + my * my) + mz * mz);
        const float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
        const float v = -(outside + inside);
        if (v > best) {
            best = v;
            win = k;
        }
    }
    g[0] = g[1] = g[2] = 0.0f;
    if (win < 0) return;  // the floor
    if (win < ns) {
        const float4 sp = s_csg[2 * win];
        const float dx = x - sp.x, dy = y - sp.y, dz = z - sp.z;
        const float l = sqrtf((dx * dx + dy * dy) + dz * dz);
        const float inv = l > 0.0f ? 1.0f / l : 0.0f;
        g[0] = -dx * inv;
        g[1] = -dy * inv;
        g[2] = -dz * inv;
        return;
    }
    const float4 b0 = s_csg[2 * win], b1 = s_csg[2 * win + 1];
    const float px = x - b0.x, py = y - b0.y, pz = z - b0.z;
    const float qx = fabsf(px) - b0.w, qy = fabsf(py) - b1.x, qz = fabsf(pz) - b1.y;
    const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
    const float outside = sqrtf((mx * mx + my * my) + mz * mz);
    const float sx = px < 0.0f ? -1.0f : 1.0f, sy = py < 0.0f ? -1.0f : 1.0f, sz = pz < 0.0f ? -1.0f : 1.0f;
    if (outside > 0.0f) {
        const float inv = 1.0f / outside;
        g[0] = -mx * sx * inv;
        g[1] = -my * sy * inv;
        g[2] = -mz * sz * inv;
    } else if (qx >= qy && qx >= qz) {
        g[0] = -sx;
    } else if (qy >= qz) {
        g[1] = -sy;
    } else {
        g[2] = -sz;
    }
}

// QEF of `count` (point, unit normal) pairs (qef_simd.h:392-579): normal
// equations about the mass point, 5 Jacobi sweeps over the pairs (0,1),
// (0,2), (1,2) with an exact 1/sqrt Givens cosine, 1/sigma kept when
// min(|sigma|, |1/sigma|) >= tol, residual with the uncentred ATb.
// The sums are accumulated one pair at a time in pair order (qef_add), then
// solved (qef_finish); no pair is kept.
struct QefSums {
    float ata[3][3];
    float atb[3];
    float pa[4];
};

__device__ __forceinline__ void qef_add(QefSums& q, float px, float py, float pz, const float nn[3]) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) q.ata[r][c] = q.ata[r][c] + nn[r] * nn[c];
    const float d = (px * nn[0] + py * nn[1]) + pz * nn[2];
#pragma unroll
    for (int r = 0; r < 3; r++) q.atb[r] = q.atb[r] + d * nn[r];
    q.pa[0] = q.pa[0] + px;
    q.pa[1] = q.pa[1] + py;
    q.pa[2] = q.pa[2] + pz;
    q.pa[3] = q.pa[3] + 1.0f;
}

__device__ float qef_finish(const QefSums& q, int count, float tol, float out[3]) {
    if (count < 2 || count > 12) {
        out[0] = out[1] = out[2] = 0.0f;
        return 0.0f;
    }
    const float(&ata)[3][3] = q.ata;
    const float(&atb)[3] = q.atb;
    const float(&pa)[4] = q.pa;
    const float mp[3] = {pa[0] / pa[3], pa[1] / pa[3], pa[2] / pa[3]};
    float rhs[3];
#pragma unroll
    for (int c = 0; c < 3; c++) rhs[c] = atb[c] - ((mp[0] * ata[0][c] + mp[1] * ata[1][c]) + mp[2] * ata[2][c]);
    float a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) a[r][c] = ata[r][c];
    for (int sw = 0; sw < 5; sw++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int i0 = q == 2 ? 1 : 0, i1 = q == 0 ? 1 : 2;
            if (a[i0][i1] == 0.0f) continue;
            const float pp = a[i0][i0], pq = a[i0][i1], qq = a[i1][i1];
            const float tau = (qq - pp) / (pq * 2.0f);
            const float stt = sqrtf(tau * tau + 1.0f);
            const float tn = 1.0f / (tau >= 0.0f ? tau + stt : tau - stt);
            const float cs = 1.0f / sqrtf(1.0f + tn * tn);
            const float sn = tn * cs;
            const float cc = cs * cs, ss = sn * sn, mx = ((2.0f * cs) * sn) * pq;
            a[i0][i0] = (cc * pp - mx) + ss * qq;
            a[i1][i1] = (ss * pp + mx) + cc * qq;
            const int ua_r = 0, ua_c = 3 - i1, va_r = 1 - i0, va_c = 2;
            const float u3 = a[ua_r][ua_c], v3 = a[va_r][va_c];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const float vu = v[r][i0], vv = v[r][i1];
                v[r][i0] = cs * vu - sn * vv;
                v[r][i1] = sn * vu + cs * vv;
            }
            a[ua_r][ua_c] = cs * u3 - sn * v3;
            a[va_r][va_c] = sn * u3 + cs * v3;
            a[i0][i1] = 0.0f;
        }
    const float sig[3] = {a[0][0], a[1][1], a[2][2]};
    float inv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float r1 = 1.0f / sig[k];
        const float ax = sig[k] < 0.0f ? -sig[k] : sig[k], ar = r1 < 0.0f ? -r1 : r1;
        inv[k] = ((ax < ar ? ax : ar) >= tol) ? r1 : 0.0f;
    }
    float pinv[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            pinv[r][c] = (v[c][0] * inv[0] * v[r][0] + v[c][1] * inv[1] * v[r][1]) + v[c][2] * inv[2] * v[r][2];
    float x[3];
#pragma unroll
    for (int c = 0; c < 3; c++) x[c] = (rhs[0] * pinv[0][c] + rhs[1] * pinv[1][c]) + rhs[2] * pinv[2][c];
    float e[3];
#pragma unroll
    for (int c = 0; c < 3; c++) e[c] = atb[c] - ((x[0] * ata[0][c] + x[1] * ata[1][c]) + x[2] * ata[2][c]);
    const float err = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
    out[0] = x[0] + mp[0];
    out[1] = x[1] + mp[1];
    out[2] = x[2] + mp[2];
    return err;
}

// Local edge -> offset of its origin sample from the cell origin and axis
// (DMCChunk.cpp:543-565).
__constant__ int8_t c_le[12][4] = {{0, 0, 0, 0}, {0, 0, 1, 0}, {0, 1, 0, 0}, {0, 1, 1, 0},
                                   {0, 0, 0, 1}, {0, 0, 1, 1}, {1, 0, 0, 1}, {1, 0, 1, 1},
                                   {0, 0, 0, 2}, {0, 1, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}};

}  // namespace

// ------------------------------------------------------------------------
// Rings.  Quad ids follow the (x, y, z, axis) order of their lattice edges:
// a row's base (row_t after the scan, per quad) plus the row's interior
// sign-changing edges before the edge's sample (all axes) and before its axis
// at that sample.  A vertex (a polygon of a cell) lies on the quads of its
// polygon's interior edges, so visiting the cell's local edges in the order of
// their (dx, dy, dz, axis) offsets lists its quads in ascending id.  With
// quads written as triangles (a, b, c), (c, d, a) (MeshProcessor.cpp:74-91) a
// quad's first and last ring cells (corners a and c) are on both triangles,
// the others on one, chosen by the quad's winding.
// ------------------------------------------------------------------------
// For label row d = 2 dx + dy of cell (x, y, z), i.e. row (x + dx, y + dy):
// the row's first quad id plus its interior sign-changing edges before
// sample z (all axes), and flag byte d: the interior edge bits of axes 0-2 at
// z (bits 0-2) and at z + 1 (bits 3-5), the label at z and z + 1 (6, 7).
// The rows come from the staged planes x, x+1, x+2.
template <int W>
__device__ __forceinline__ void quad_row_at(const uint64_t (*s_planes)[plane_words(W)], const uint32_t* rtc,
                                            uint32_t div, int dim, int x, int y, int z, int d, uint32_t& pre,
                                            uint32_t& fl) {
    const int dx = d >> 1, dy = d & 1, xx = x + dx, yy = y + dy;
    const RowBits<W> ra = lds_row<W>(s_planes[dx], dim, yy, true);
    const RowBits<W> rb = lds_row<W>(s_planes[dx], dim, yy + 1, yy + 1 < dim);
    const RowBits<W> rc = lds_row<W>(s_planes[dx + 1], dim, yy, xx + 1 < dim);
    RowBits<W> e0, e1, e2;
    interior_edges<W>(ra, rb, rc, dim, xx, yy, e0, e1, e2);
    pre = rtc[(int64_t)xx * dim + yy] / div + prefix_pop<W>(e0, z) + prefix_pop<W>(e1, z) + prefix_pop<W>(e2, z);
    const uint32_t f = bit_at<W>(e0, z) | (bit_at<W>(e1, z) << 1) | (bit_at<W>(e2, z) << 2) |
                       (bit_at<W>(e0, z + 1) << 3) | (bit_at<W>(e1, z + 1) << 4) | (bit_at<W>(e2, z + 1) << 5) |
                       (bit_at<W>(ra, z) << 6) | (bit_at<W>(ra, z + 1) << 7);
    fl |= f << (8 * d);
}

// Writes vertex g's ring (polygon p of a cell with polygon-of-edge nibbles
// pe) from the four rows' quad prefixes and flags; returns its size.
__device__ __forceinline__ uint32_t dmc_ring(const Arenas& A, const uint32_t (&pre)[4], uint32_t fl, uint64_t pe,
                                             uint32_t p, int64_t g, bool tris) {
    uint32_t n = 0;
    bool over = false;
    const int64_t cap = A.cap_vertices;
    auto put = [&](uint32_t pid) {
        if (n < (uint32_t)kRingSlots) A.ring[(int64_t)n * cap + g] = pid;
        else over = true;
        n++;
    };
#pragma unroll
    for (int i = 0; i < 12; i++) {
        // local edges in (dx, dy, dz, axis) order: 0, 4, 8, 1, 5, 2, 9, 3, 6, 10, 7, 11
        constexpr int kOrder[12] = {0, 4, 8, 1, 5, 2, 9, 3, 6, 10, 7, 11};
        const int le = kOrder[i];
        if ((uint32_t)((pe >> (4 * le)) & 15ull) != p) continue;
        // c_le[le]: the origin offset and axis (a = le >> 2, the other two
        // offsets bits 1 and 0 of le)
        const int a = le >> 2, b1 = (le >> 1) & 1, b0 = le & 1;
        const int dx = a == 0 ? 0 : b1, dy = a == 0 ? b1 : (a == 2 ? b0 : 0), dz = a == 2 ? 0 : b0;
        const int d = 2 * dx + dy;
        const uint32_t f = fl >> (8 * d);
        const uint32_t cb = dz ? f >> 3 : f;  // interior edge bits of the edge's sample
        if (!((cb >> a) & 1u)) continue;      // on the chunk's boundary: no quad
        const uint32_t qid = pre[d] + (dz ? (f & 1u) + ((f >> 1) & 1u) + ((f >> 2) & 1u) : 0u) +
                             (a >= 1 ? cb & 1u : 0u) + (a == 2 ? (cb >> 1) & 1u : 0u);
        if (!tris) {
            put(qid);
            continue;
        }
        // ring cell r = 3 - (le & 3) of the edge (c_ring order); the quad is
        // (v0, v1, v3, v2), with v1 and v2 swapped when the solid end of the
        // edge is not on the side its axis winds toward
        const int r = 3 - (le & 3);
        const uint32_t prim = 2u * qid;
        if (r == 0 || r == 3) {
            put(prim);
            put(prim + 1u);
        } else {
            const bool solid_origin = ((f >> (6 + dz)) & 1u) == 0u;
            const bool flip = solid_origin != (a != 1);
            put(prim + ((r == 1) == flip ? 1u : 0u));
        }
    }
    if (over) A.ctr->error = 10;
    return n;
}

// ------------------------------------------------------------------------
// vertices: one per polygon of each surface cell.  Row pass (wave per row,
// lane = z): per-cell first vertex id and mask, per vertex its cell and
// polygon code and boundary flag.  The plane's vertices are a contiguous id
// range, so a dense pass of the whole block then places them (mass point or
// QEF), free of the row pass's idle lanes.
// ------------------------------------------------------------------------
__device__ void place_vertex(const Arenas& A, const ChunkMeta& M, const SamplerDev& S, const uint16_t* csg,
                             const float4* s_csg, int x, uint32_t code, uint32_t mask, bool qef, float out[3],
                             float& err) {
    const int dim = M.dim;
    const int y = (int)((code >> 8) & 255u), z = (int)((code >> 16) & 255u), p = (int)(code >> 24);
    const float* __restrict__ D = A.dens + M.dens_off;
    const uint64_t row = c_dmc_table[mask];
    const int nn = (int)(row >> 60);
    // skip to polygon p
    int s = 0;
    for (int q = 0; q < p && s < nn; s++)
        if (((row >> (4 * s)) & 15ull) == 15ull) q++;
    QefSums qs = {{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, {0, 0, 0}, {0, 0, 0, 0}};
    // the polygon's edges (at most 7; an eighth is an error), their endpoint
    // densities loaded together before any is used; local edge e runs along
    // axis e >> 2 from the corner whose other two offsets are bits 1 and 0
    // of e (the rows of c_le)
    int ee[7], k = 0;
    float d0[7], d1[7];
    const int64_t own = ((int64_t)x * dim + y) * dim + z;
#pragma unroll
    for (int j = 0; j < 7; j++) {
        const int e = s + j < nn ? (int)((row >> (4 * (s + j))) & 15ull) : 15;
        const bool valid = e != 15 && k == j;  // the polygon ends at its first separator
        ee[j] = valid ? e : 0;
        const int a = ee[j] >> 2, b1 = (ee[j] >> 1) & 1, b0 = ee[j] & 1;
        const int ox = a == 0 ? 0 : b1, oy = a == 0 ? b1 : (a == 2 ? b0 : 0), oz = a == 2 ? 0 : b0;
        const int64_t i0 = (((int64_t)(x + ox) * dim) + (y + oy)) * dim + (z + oz);
        const int64_t i1 = i0 + (a == 0 ? (int64_t)dim * dim : (a == 1 ? (int64_t)dim : 1));
        d0[j] = D[valid ? i0 : own];
        d1[j] = D[valid ? i1 : own];
        k += valid ? 1 : 0;
    }
    {
        const int e8 = s + 7 < nn ? (int)((row >> (4 * (s + 7))) & 15ull) : 15;
        if (k == 7 && e8 != 15) k = 8;  // more than 7 edges: reported below
    }
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
    for (int j = 0; j < 7; j++) {
        if (j >= k) break;
        const int e = ee[j];
        const int x0 = x + c_le[e][0], y0 = y + c_le[e][1], z0 = z + c_le[e][2], a = c_le[e][3];
        const int x1 = x0 + (a == 0), y1 = y0 + (a == 1), z1 = z0 + (a == 2);
        const float s0 = d0[j], s1 = d1[j];
        // DMCChunk::calculate_isovertex (DMCChunk.cpp:657-674)
        const float mu = (0.0f - s0) / (s1 - s0);
        const float dx = (float)x1 - (float)x0, dy = (float)y1 - (float)y0, dz = (float)z1 - (float)z0;
        const float qx = dx * mu + (float)x0, qy = dy * mu + (float)y0, qz = dz * mu + (float)z0;
        ax = ax + qx;
        ay = ay + qy;
        az = az + qz;
        if (qef) {
            // unit normal of the CSG field at the point (world coordinates)
            const float wx = qx * M.delta + M.opos[0], wy = qy * M.delta + M.opos[1], wz = qz * M.delta + M.opos[2];
            float g[3];
            if (s_csg) csg_gradient_lds(s_csg, M.csg_ns, M.csg_nb, wx, wy, wz, g);
            else csg_gradient(S, csg, M.csg_ns, M.csg_nb, wx, wy, wz, g);
            const float len = sqrtf((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
            float nrm[3];
            for (int c = 0; c < 3; c++) nrm[c] = len > 0.0f ? g[c] / len : 0.0f;
            qef_add(qs, qx, qy, qz, nrm);
        }
    }
    err = 0.0f;
    if (k == 0 || k > 7) {
        A.ctr->error = 8;
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    if (qef) {
        err = qef_finish(qs, k, A.qef_tol, out);
    } else {
        const float fk = (float)k;
        out[0] = ax / fk;
        out[1] = ay / fk;
        out[2] = az / fk;
    }
}

template <int W>
__global__ void __launch_bounds__(kRowThreads) k_dmc_vertices(Arenas A, int n) {
    __shared__ DmcTables s_t;
    __shared__ uint64_t s_planes[3][plane_words(W)];  // labels of planes x, x+1, x+2
    __shared__ float4 s_csg[2 * kCsgStage];  // the chunk's culled CSG primitives (QEF normals)
    load_dmc_tables(s_t);
    const SamplerDev* __restrict__ S = A.sampler;
    const int wave = uniform_int(threadIdx.x >> 6);
    const int lane = lane_id();
    const uint64_t lt = lanemask_lt();
    const int64_t base = A.ctr->class_base[W - 1], total = A.ctr->class_planes[W - 1];
    for (int64_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int2 pm = A.plane_map[base + item];
        const int j = uniform_int(pm.x);
        const ChunkMeta M = read_chunk(A.meta, j);
        const int dim = M.dim;
        if (M.status != BMF_OK || M.V == 0) continue;
        const int x = uniform_int(pm.y);
        if (x >= dim - 1) continue;
        const bool qef = A.placement == BMF_PLACE_QEF && S->kind == BMF_SAMPLER_CSG;
        const uint16_t* csg = A.csg_list ? A.csg_list + (int64_t)j * A.csg_stride : nullptr;
        const bool staged = qef && csg && M.csg_ns + M.csg_nb <= kCsgStage;  // uniform across the block
        __syncthreads();
        {
            uint64_t* const dst[3] = {s_planes[0], s_planes[1], s_planes[2]};
            const uint64_t* const src[3] = {A.labels + M.lab_off, A.labels + M.lab_off, A.labels + M.lab_off};
            const int xp[3] = {x, x + 1, x + 2};
            stage_plane_group<3>(dst, src, M.lab_words, dim, xp);
        }
        if (staged) stage_csg(*S, csg, M.csg_ns, M.csg_nb, s_csg);
        __syncthreads();
        const uint32_t* rv = A.row_v + M.row_off + (int64_t)x * dim;
        const int64_t vo = M.v_off;
        const uint32_t Vc = (uint32_t)M.V;
        const bool smooth = (M.iters > 0 || A.normals) && M.T > 0;  // the chunk runs the smoothing kernels
        const bool tris = A.quad_tris != 0;
        const uint32_t* rtc = A.row_t + M.row_off;
        // ---- row pass ----
        for (int y = wave; y < dim - 1; y += kRowThreads / 64) {
            if (rv[y + 1] == rv[y]) continue;  // the row's cells have no polygon
            const RowBits<W> r00 = lds_row<W>(s_planes[0], dim, y, true);
            const RowBits<W> r01 = lds_row<W>(s_planes[0], dim, y + 1, true);
            const RowBits<W> r10 = lds_row<W>(s_planes[1], dim, y, true);
            const RowBits<W> r11 = lds_row<W>(s_planes[1], dim, y + 1, true);
            if (quad_uniform<W>(r00, r01, r10, r11, dim)) continue;
            uint32_t carry = rv[y];
            for (int z0 = 0; z0 < dim - 1; z0 += 64) {
                const int z = z0 + lane;
                const uint32_t mask = z < dim - 1 ? cell_mask<W>(r00, r01, r10, r11, z) : 0u;
                const uint32_t np = s_t.np[mask];
                uint32_t pre;
                const uint32_t tot = wave_sum3(np, lt, pre);
                const uint32_t vb = carry + pre;
                carry += tot;
                if (!np) continue;
                const int64_t ci = M.dens_off + ((int64_t)x * dim + y) * dim + z;
                A.cell_tb[ci] = vb;
                A.cell_mask[ci] = (uint8_t)mask;
                for (uint32_t p = 0; p < np; p++) {
                    const uint32_t v = vb + p;
                    if (v >= Vc) {
                        A.ctr->error = 8;
                        break;
                    }
                    A.boundary[vo + v] =
                        (uint8_t)(x == 0 || y == 0 || z == 0 || x + 1 == dim - 1 || y + 1 == dim - 1 || z + 1 == dim - 1);
                    A.vedge[vo + v] = (uint32_t)x | ((uint32_t)y << 8) | ((uint32_t)z << 16) | (p << 24);
                }
            }
        }
        __syncthreads();  // the plane's vertex codes are written
        // ---- dense placement of the plane's vertices [rv[0], next plane's first) ----
        const uint32_t v0 = rv[0], v1 = min(x + 1 < dim ? rv[dim] : Vc, Vc);
        for (uint32_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
            const uint32_t code = A.vedge[vo + v];
            const int y = (int)((code >> 8) & 255u), z = (int)((code >> 16) & 255u);
            const uint32_t mask = A.cell_mask[M.dens_off + ((int64_t)x * dim + y) * dim + z];
            float o[3], err;
            place_vertex(A, M, *S, csg, staged ? s_csg : nullptr, x, code, mask, qef, o, err);
            float* P = A.pos + (vo + v) * 3;
            P[0] = o[0];
            P[1] = o[1];
            P[2] = o[2];
            if (smooth) A.pos4[vo + v] = make_float4(o[0], o[1], o[2], 0.0f);
            if (A.qef_err) A.qef_err[vo + v] = err;
            if (A.normals) {  // a vertex with no polygon keeps the zero normal
                A.nrm4[vo + v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                float* O = A.nrm + (vo + v) * 3;
                O[0] = 0.0f;
                O[1] = 0.0f;
                O[2] = 0.0f;
            }
        }
        if (!smooth) continue;
        // ---- dense pass: the rings of the plane's vertices ----
        for (uint32_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
            const uint32_t code = A.vedge[vo + v];
            const int y = (int)((code >> 8) & 255u), z = (int)((code >> 16) & 255u);
            const uint32_t mask = A.cell_mask[M.dens_off + ((int64_t)x * dim + y) * dim + z];
            uint32_t pre[4] = {0u, 0u, 0u, 0u}, fl = 0u;
            // one row at a time (few live registers), its prefix kept by selects
#pragma unroll 1
            for (int d = 0; d < 4; d++) {
                uint32_t pd;
                quad_row_at<W>(s_planes, rtc, tris ? 2u : 1u, dim, x, y, z, d, pd, fl);
#pragma unroll
                for (int q = 0; q < 4; q++) pre[q] = q == d ? pd : pre[q];
            }
            A.ring_n[vo + v] = dmc_ring(A, pre, fl, s_t.pe[mask], code >> 24, vo + v, tris);
        }
    }
}

// ------------------------------------------------------------------------
// quads around the interior sign-changing edges owned by the row's samples
// ------------------------------------------------------------------------
template <int W>
__global__ void __launch_bounds__(kRowThreads) k_dmc_quads(Arenas A, int n) {
    __shared__ DmcTables s_t;
    __shared__ uint64_t s_planes[2][plane_words(W)];
    load_dmc_tables(s_t);
    const int wave = uniform_int(threadIdx.x >> 6);
    const int lane = lane_id();
    const int64_t base = A.ctr->class_base[W - 1], total = A.ctr->class_planes[W - 1];
    const bool tris = A.quad_tris != 0;
    for (int64_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int2 pm = A.plane_map[base + item];
        const ChunkMeta M = read_chunk(A.meta, pm.x);
        const int dim = M.dim;
        if (M.status != BMF_OK || M.T == 0) continue;
        const int x = uniform_int(pm.y);
        __syncthreads();
        {
            uint64_t* const dst[2] = {s_planes[0], s_planes[1]};
            const uint64_t* const src[2] = {A.labels + M.lab_off, A.labels + M.lab_off};
            const int xp[2] = {x, x + 1};
            stage_plane_group<2>(dst, src, M.lab_words, dim, xp);
        }
        __syncthreads();
        const uint32_t* rt = A.row_t + M.row_off + (int64_t)x * dim;
        const uint32_t Vc = (uint32_t)M.V, Tc = (uint32_t)M.T;
        uint32_t* __restrict__ I = A.idx + M.i_off;
        for (int y = wave; y < dim; y += kRowThreads / 64) {
            const bool x1 = x + 1 < dim, y1 = y + 1 < dim;
            // rows whose samples own no quad (the next row's base, or the next
            // plane's, or the chunk's total after the last row) skip the label work
            const uint32_t next = y1 ? rt[y + 1] : (x1 ? rt[dim] : (uint32_t)M.T);
            if (next == rt[y]) continue;
            const RowBits<W> r00 = lds_row<W>(s_planes[0], dim, y, true);
            const RowBits<W> r01 = lds_row<W>(s_planes[0], dim, y + 1, y1);
            const RowBits<W> r10 = lds_row<W>(s_planes[1], dim, y, x1);
            RowBits<W> ex, ey, ez;
            interior_edges<W>(r00, r01, r10, dim, x, y, ex, ey, ez);
            if (total_pop<W>(ex) + total_pop<W>(ey) + total_pop<W>(ez) == 0) continue;
            const uint32_t qbase = rt[y] / (tris ? 2u : 1u);
            for (int z0 = 0; z0 < dim; z0 += 64) {
                const int z = z0 + lane;
                if (z >= dim) continue;
                const uint32_t f[3] = {bit_at<W>(ex, z), bit_at<W>(ey, z), bit_at<W>(ez, z)};
                if (!(f[0] | f[1] | f[2])) continue;
                uint32_t qid = qbase + prefix_pop<W>(ex, z) + prefix_pop<W>(ey, z) + prefix_pop<W>(ez, z);
                const bool solid_origin = bit_at<W>(r00, z) == 0u;
                // the rings of the sample's three edges are cells of the 2 x 2 x 2
                // block (x-1 .. x, y-1 .. y, z-1 .. z): its eight first-vertex ids
                // and masks load together (a cell outside the chunk reads cell 0;
                // only the rings of interior edges, all inside, use them)
                uint32_t tb8[8], mk8[8];
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const int cx = x - 1 + (c >> 2), cy = y - 1 + ((c >> 1) & 1), cz = z - 1 + (c & 1);
                    const bool in = cx >= 0 && cy >= 0 && cz >= 0 && cx <= dim - 2 && cy <= dim - 2 && cz <= dim - 2;
                    const int64_t ci = M.dens_off + (in ? ((int64_t)cx * dim + cy) * dim + cz : 0);
                    tb8[c] = A.cell_tb[ci];
                    mk8[c] = A.cell_mask[ci];
                }
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    if (!f[a]) continue;
                    uint32_t v[4];
                    bool ok = true;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        // ring cell r of axis a: offsets (c_ring) -1 / 0 per axis -> block cell
                        const int c = ((c_ring[a][r][0] + 1) << 2) | ((c_ring[a][r][1] + 1) << 1) | (c_ring[a][r][2] + 1);
                        const uint32_t le = (uint32_t)c_ring[a][r][3];
                        const uint32_t pi = (uint32_t)((s_t.pe[mk8[c]] >> (4 * le)) & 15ull);
                        v[r] = tb8[c] + pi;
                        if (pi == 15u || v[r] >= Vc) ok = false;
                    }
                    const uint32_t prim = tris ? 2u * qid : qid;
                    if (!ok || prim + (tris ? 1u : 0u) >= Tc) {
                        A.ctr->error = 9;
                        qid++;
                        continue;
                    }
                    // cyclic c0 c1 c3 c2: right-hand normal along -x, +y, -z;
                    // flipped so that it points to the solid end of the edge
                    const bool toward_origin = a != 1;
                    uint32_t o[4] = {v[0], v[1], v[3], v[2]};
                    if (solid_origin != toward_origin) {
                        o[1] = v[2];
                        o[3] = v[1];
                    }
                    if (tris) {  // MeshProcessor::flush_to_tris (MeshProcessor.cpp:74-91)
                        uint32_t* t = I + 3ll * prim;
                        t[0] = o[0];
                        t[1] = o[1];
                        t[2] = o[2];
                        t[3] = o[2];
                        t[4] = o[3];
                        t[5] = o[0];
                    } else {
                        uint32_t* t = I + 4ll * prim;
                        t[0] = o[0];
                        t[1] = o[1];
                        t[2] = o[2];
                        t[3] = o[3];
                    }
                    qid++;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// smoothing over the rings (chunks of the smoothing list, XCD-grouped tiles
// as in the triangle path)
// ------------------------------------------------------------------------
// Dual step (MeshProcessor.cpp:130-236) for N-gons (N = 3 or 4, one per
// chunk): ((0 + p0) + p1 ...) / N, the corner ids and then the corners
// loaded together.  Tiles come from the per-frame tile map.
__global__ void __launch_bounds__(kSmoothThreads) k_dual_n(Arenas A, int m) {
    const int xcd = blockIdx.x % kXcds;
    const int64_t end = A.smooth_xcd_t[xcd + 1];
    for (int64_t tile = A.smooth_xcd_t[xcd] + blockIdx.x / kXcds; tile < end; tile += gridDim.x / kXcds) {
        const int2 tm = A.smooth_tmap[tile];  // (chunk, tile within the chunk)
        const ChunkMeta& M = A.meta[uniform_int(tm.x)];
        const int iters = uniform_int(M.iters), N = uniform_int(M.ipp), T = uniform_int(M.T);
        const uint32_t V = (uint32_t)uniform_int(M.V);
        const int64_t i_off = M.i_off, v_off = M.v_off;
        if (iters <= m) continue;
        const int64_t t = (int64_t)uniform_int(tm.y) * kSmoothThreads + threadIdx.x;
        if (t >= T) continue;
        const uint32_t* __restrict__ I = A.idx + i_off + (int64_t)N * t;
        const float4* __restrict__ P = A.pos4 + v_off;
        const uint32_t u0 = I[0], u1 = I[1], u2 = I[2], u3 = N == 4 ? I[3] : u0;
        if (u0 >= V || u1 >= V || u2 >= V || u3 >= V) {
            A.ctr->error = 11;
            continue;
        }
        const float4 a = P[u0], b = P[u1], c = P[u2], d = P[u3];
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        cx = cx + a.x;
        cy = cy + a.y;
        cz = cz + a.z;
        cx = cx + b.x;
        cy = cy + b.y;
        cz = cz + b.z;
        cx = cx + c.x;
        cy = cy + c.y;
        cz = cz + c.z;
        if (N == 4) {
            cx = cx + d.x;
            cy = cy + d.y;
            cz = cz + d.z;
        }
        const float fn = (float)N;
        A.cen4[i_off / N + t] = make_float4(cx / fn, cy / fn, cz / fn, 0.0f);
    }
}

// Primal step (MeshProcessor.cpp:238-306) over the ring (ascending primitive
// id): the centroids summed from +0 in ring order, eight ring slots and then
// their centroids loaded per round.
constexpr int kRingRoundN = 8;
__global__ void __launch_bounds__(kSmoothThreads) k_primal_ring(Arenas A, int m) {
    const int xcd = blockIdx.x % kXcds;
    const int64_t end = A.smooth_xcd_v[xcd + 1];
    const int64_t cap = A.cap_vertices;
    for (int64_t tile = A.smooth_xcd_v[xcd] + blockIdx.x / kXcds; tile < end; tile += gridDim.x / kXcds) {
        const int2 tm = A.smooth_vmap[tile];  // (chunk, tile within the chunk)
        const ChunkMeta& M = A.meta[uniform_int(tm.x)];
        const int K = uniform_int(M.iters), N = uniform_int(M.ipp), V = uniform_int(M.V), bp = uniform_int(M.bp);
        const int64_t i_off = M.i_off, v_off = M.v_off;
        if (K <= m) continue;
        const int64_t v = (int64_t)uniform_int(tm.y) * kSmoothThreads + threadIdx.x;
        if (v >= V) continue;
        const int64_t g = v_off + v;
        const uint8_t bnd = A.boundary[g];
        const int n = (int)min(A.ring_n[g], (uint32_t)kRingSlots);
        if ((!bp && bnd) || n == 0) continue;
        const float4* __restrict__ C = A.cen4 + i_off / N;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (int i0 = 0; i0 < n; i0 += kRingRoundN) {
            uint32_t id[kRingRoundN];
#pragma unroll
            for (int k = 0; k < kRingRoundN; k++) id[k] = A.ring[(int64_t)(i0 + k < n ? i0 + k : i0) * cap + g];
            float4 q[kRingRoundN];
#pragma unroll
            for (int k = 0; k < kRingRoundN; k++) q[k] = C[id[k]];
#pragma unroll
            for (int k = 0; k < kRingRoundN; k++) {
                const bool ok = i0 + k < n;
                sx = ok ? sx + q[k].x : sx;
                sy = ok ? sy + q[k].y : sy;
                sz = ok ? sz + q[k].z : sz;
            }
        }
        const float fc = (float)n;
        const float px = sx / fc, py = sy / fc, pz = sz / fc;
        A.pos4[g] = make_float4(px, py, pz, 0.0f);
        if (m == K - 1) {
            float* __restrict__ P = A.pos + g * 3;
            P[0] = px;
            P[1] = py;
            P[2] = pz;
        }
    }
}

// With vertex normals (the smooth_normals branch of MeshProcessor<N>,
// MeshProcessor.cpp:159-216, 271-287): every chunk of the list runs max(K, 1)
// steps.  Step m's face normal while m == 0, m < max_norms (min(K/2 - 3, 10))
// or m < 3 is, for a triangle, -cross(normalize(p0 - p1), normalize(p0 - p2));
// for a quad, n1 = cross(normalize(p0 - p1), normalize(p0 - p2)) and
// n2 = cross(normalize(p2 - p3), normalize(p2 - p0)) with the reference's NaN
// fallbacks (a NaN n1.x takes n2, or (0, 1, 0) when n2 is NaN too; a NaN
// n2.x then takes n1), and -normalize((n1 + n2) * 0.5) (:187-205); later
// steps take the sum of the corners' normals.  normalize(v) is
// v / sqrt((x*x + y*y) + z*z).
__device__ __forceinline__ float3 unit_cross(float ux, float uy, float uz, float wx, float wy, float wz) {
    const float lu = sqrtf((ux * ux + uy * uy) + uz * uz), lw = sqrtf((wx * wx + wy * wy) + wz * wz);
    ux = ux / lu;
    uy = uy / lu;
    uz = uz / lu;
    wx = wx / lw;
    wy = wy / lw;
    wz = wz / lw;
    return make_float3(uy * wz - wy * uz, uz * wx - wz * ux, ux * wy - wx * uy);
}

__global__ void __launch_bounds__(kSmoothThreads) k_dual_n_nrm(Arenas A, int m) {
    const int nsm = A.ctr->n_smooth;
    const int xcd = blockIdx.x % kXcds;
    const int64_t end = A.smooth_xcd_t[xcd + 1];
    for (int64_t tile = A.smooth_xcd_t[xcd] + blockIdx.x / kXcds; tile < end; tile += gridDim.x / kXcds) {
        const int k = find_chunk(A.smooth_tprefix, nsm, tile);
        const ChunkMeta& M = A.meta[A.smooth_list[k]];
        const int K = M.iters;
        if (max(K, 1) <= m) continue;
        const int64_t t = (tile - A.smooth_tprefix[k]) * kSmoothThreads + threadIdx.x;
        if (t >= M.T) continue;
        const int N = M.ipp;
        const uint32_t* __restrict__ I = A.idx + M.i_off + (int64_t)N * t;
        const float4* __restrict__ P = A.pos4 + M.v_off;
        const float4* __restrict__ NV = A.nrm4 + M.v_off;
        uint32_t u[4] = {0u, 0u, 0u, 0u};
        bool ok = true;
        for (int c = 0; c < N; c++) {
            u[c] = I[c];
            if (u[c] >= (uint32_t)M.V) ok = false;
        }
        if (!ok) {
            A.ctr->error = 11;
            continue;
        }
        float4 q[4];
        for (int c = 0; c < N; c++) q[c] = P[u[c]];
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        for (int c = 0; c < N; c++) {
            cx = cx + q[c].x;
            cy = cy + q[c].y;
            cz = cz + q[c].z;
        }
        const float fN = (float)N;
        const int64_t tt = M.i_off / N + t;
        A.cen4[tt] = make_float4(cx / fN, cy / fN, cz / fN, 0.0f);
        const int max_norms = min(K / 2 - 3, 10);
        float4 fn;
        if (m == 0 || m < max_norms || m < 3) {
            if (N == 3) {
                const float3 f = unit_cross(q[0].x - q[1].x, q[0].y - q[1].y, q[0].z - q[1].z, q[0].x - q[2].x,
                                            q[0].y - q[2].y, q[0].z - q[2].z);
                fn = make_float4(-f.x, -f.y, -f.z, 0.0f);
            } else {
                float3 n1 = unit_cross(q[0].x - q[1].x, q[0].y - q[1].y, q[0].z - q[1].z, q[0].x - q[2].x,
                                       q[0].y - q[2].y, q[0].z - q[2].z);
                float3 n2 = unit_cross(q[2].x - q[3].x, q[2].y - q[3].y, q[2].z - q[3].z, q[2].x - q[0].x,
                                       q[2].y - q[0].y, q[2].z - q[0].z);
                if (__builtin_isnan(n1.x)) {
                    n1 = n2;
                    if (__builtin_isnan(n1.x)) n1 = make_float3(0.0f, 1.0f, 0.0f);
                }
                if (__builtin_isnan(n2.x)) n2 = n1;
                const float hx = (n1.x + n2.x) * 0.5f, hy = (n1.y + n2.y) * 0.5f, hz = (n1.z + n2.z) * 0.5f;
                const float lh = sqrtf((hx * hx + hy * hy) + hz * hz);
                fn = make_float4(-(hx / lh), -(hy / lh), -(hz / lh), 0.0f);
            }
        } else {
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            for (int c = 0; c < N; c++) {
                const float4 nc = NV[u[c]];
                sx = sx + nc.x;
                sy = sy + nc.y;
                sz = sz + nc.z;
            }
            fn = make_float4(sx, sy, sz, 0.0f);
        }
        A.fn4[tt] = fn;
    }
}

// Primal step with normals over the sorted ring: the ring's centroids and
// face normals summed from +0 in ascending primitive order and divided by
// the ring size; the normal is normalised when m == 3 or (m == 0 and K <= 3)
// and kept when its y is not 0, for every vertex with a ring; positions of
// boundary vertices stay unless boundary processing (K > 0 only); the last
// step writes the unit normal (0 for a zero normal).
__global__ void __launch_bounds__(kSmoothThreads) k_primal_ring_nrm(Arenas A, int m) {
    const int nsm = A.ctr->n_smooth;
    const int xcd = blockIdx.x % kXcds;
    const int64_t end = A.smooth_xcd_v[xcd + 1];
    for (int64_t tile = A.smooth_xcd_v[xcd] + blockIdx.x / kXcds; tile < end; tile += gridDim.x / kXcds) {
        const int k = find_chunk(A.smooth_prefix, nsm, tile);
        const ChunkMeta& M = A.meta[A.smooth_list[k]];
        const int K = M.iters;
        const int steps = max(K, 1);
        if (steps <= m) continue;
        const int64_t v = (tile - A.smooth_prefix[k]) * kSmoothThreads + threadIdx.x;
        if (v >= M.V) continue;
        const int64_t g = M.v_off + v;
        const int n = (int)min(A.ring_n[g], (uint32_t)kRingSlots);
        const int64_t t_off = M.i_off / M.ipp;
        const float4* __restrict__ C = A.cen4 + t_off;
        const float4* __restrict__ F = A.fn4 + t_off;
        float4 nv = A.nrm4[g];
        if (n > 0) {
            float sx = 0.0f, sy = 0.0f, sz = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f;
            for (int i = 0; i < n; i++) {
                const uint32_t pid = A.ring[(int64_t)i * A.cap_vertices + g];
                const float4 q = C[pid], f = F[pid];
                sx = sx + q.x;
                sy = sy + q.y;
                sz = sz + q.z;
                nx = nx + f.x;
                ny = ny + f.y;
                nz = nz + f.z;
            }
            const float fc = (float)n;
            float ax = nx / fc, ay = ny / fc, az = nz / fc;
            const bool unit = m != steps - 1 && (m == 3 || (m == 0 && K <= 3));
            if (unit) {
                const float l = sqrtf((ax * ax + ay * ay) + az * az);
                ax = ax / l;
                ay = ay / l;
                az = az / l;
            }
            if (ay != 0.0f) {
                nv = make_float4(ax, ay, az, 0.0f);
                A.nrm4[g] = nv;
            }
            if (K > 0 && (M.bp || !A.boundary[g])) {
                const float px = sx / fc, py = sy / fc, pz = sz / fc;
                A.pos4[g] = make_float4(px, py, pz, 0.0f);
                if (m == K - 1) {
                    float* __restrict__ P = A.pos + g * 3;
                    P[0] = px;
                    P[1] = py;
                    P[2] = pz;
                }
            }
        }
        if (m == steps - 1) {
            const float l = sqrtf((nv.x * nv.x + nv.y * nv.y) + nv.z * nv.z);
            float* __restrict__ O = A.nrm + g * 3;
            O[0] = l > 0.0f ? nv.x / l : 0.0f;
            O[1] = l > 0.0f ? nv.y / l : 0.0f;
            O[2] = l > 0.0f ? nv.z / l : 0.0f;
        }
    }
}

// ------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------

hipError_t launch_dmc_emit(const Arenas& A, int n, int blocks, int wclasses, hipStream_t s) {
    // per class: vertices, then the quads that index them
    if (wclasses & 1) hipLaunchKernelGGL(k_dmc_vertices<1>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 2) hipLaunchKernelGGL(k_dmc_vertices<2>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 4) hipLaunchKernelGGL(k_dmc_vertices<3>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 8) hipLaunchKernelGGL(k_dmc_vertices<4>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 1) hipLaunchKernelGGL(k_dmc_quads<1>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 2) hipLaunchKernelGGL(k_dmc_quads<2>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 4) hipLaunchKernelGGL(k_dmc_quads<3>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    if (wclasses & 8) hipLaunchKernelGGL(k_dmc_quads<4>, dim3(blocks), dim3(kRowThreads), 0, s, A, n);
    return hipGetLastError();
}

hipError_t launch_dmc_smooth(const Arenas& A, int m, int blocks, hipStream_t s) {
    blocks = blocks < kXcds ? kXcds : blocks - blocks % kXcds;  // whole XCD groups
    if (m == 0 && !A.normals) {
        const hipError_t e = launch_smooth_map(A, blocks, s);
        if (e != hipSuccess) return e;
    }
    if (A.normals) {
        hipLaunchKernelGGL(k_dual_n_nrm, dim3(blocks), dim3(kSmoothThreads), 0, s, A, m);
        hipLaunchKernelGGL(k_primal_ring_nrm, dim3(blocks), dim3(kSmoothThreads), 0, s, A, m);
    } else {
        hipLaunchKernelGGL(k_dual_n, dim3(blocks), dim3(kSmoothThreads), 0, s, A, m);
        hipLaunchKernelGGL(k_primal_ring, dim3(blocks), dim3(kSmoothThreads), 0, s, A, m);
    }
    return hipGetLastError();
}

}  // namespace bmf
