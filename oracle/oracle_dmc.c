/*
 * CPU oracle for the dual-marching-cubes quad mode -- TEST INFRASTRUCTURE
 * ONLY.  Only tests/ load this code; the product never does.
 *
 * The reference carries the Nielson dual-marching-cubes polygon table
 * (Tables.hpp:76-354, EdgeTable) and a quad mesh processor
 * (MeshProcessor<4>, flush_to_tris: MeshProcessor.cpp:73-91) but no live
 * caller.  The quad mode restated here is defined as follows:
 *   - cells (x,y,z), x,y,z < dim-1, in lexicographic order; a cell whose
 *     corner mask m (bit 4dx+2dy+dz = density < 0) is not 0 or 255 gets one
 *     vertex per polygon of EdgeTable[m] (up to its -2 end marker), in
 *     table order;
 *   - a vertex sits at the mass point of its polygon's edge points
 *     ((0 + q_0) + q_1 + ...) / k, the edge points being the marching-cubes
 *     vertices (DMCChunk.cpp:657-674), or at the QEF solution of those
 *     points with the unit analytic gradients of the sampler as normals
 *     (qef_simd.h:392-579 semantics, FP32, truncated at tol);
 *   - every sign-changing lattice edge whose four surrounding cells lie in
 *     the chunk gets one quad joining, in each of those cells, the vertex
 *     of the polygon that contains the edge; quads are numbered in edge
 *     order (x, y, z, axis) and wound so that the right-hand normal points
 *     to the solid end of the edge (the winding of the marching-cubes table);
 *   - boundary flag: the cell has a corner on a chunk face;
 *   - as triangles: quad (a,b,c,d) -> (a,b,c), (c,d,a) (flush_to_tris).
 */
#include "oracle.h"
#include "../include/bmf_mc_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* local edge -> (dx, dy, dz, axis) from the cell origin (DMCChunk.cpp:543-565) */
static const int LE[12][4] = {
    {0, 0, 0, 0}, {0, 0, 1, 0}, {0, 1, 0, 0}, {0, 1, 1, 0},
    {0, 0, 0, 1}, {0, 0, 1, 1}, {1, 0, 0, 1}, {1, 0, 1, 1},
    {0, 0, 0, 2}, {0, 1, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}};

/* the four cells around an edge along each axis, lexicographic, with the
 * edge's local number in each */
static const int RING[3][4][4] = {
    {{0, -1, -1, 3}, {0, -1, 0, 2}, {0, 0, -1, 1}, {0, 0, 0, 0}},
    {{-1, 0, -1, 7}, {-1, 0, 0, 6}, {0, 0, -1, 5}, {0, 0, 0, 4}},
    {{-1, -1, 0, 11}, {-1, 0, 0, 10}, {0, -1, 0, 9}, {0, 0, 0, 8}}};

int or_dmc_polygons(int mask, int* edges, int* start)
{
    const uint64_t row = bmf_dmc_poly_table[mask];
    const int n = (int)(row >> 60);
    int np = 0, ne = 0;
    start[0] = 0;
    for (int s = 0; s < n; s++) {
        const int e = (int)((row >> (4 * s)) & 15);
        if (e == 15) {
            start[++np] = ne;
        } else {
            edges[ne++] = e;
        }
    }
    if (ne > start[np]) start[++np] = ne;
    return np;
}

int or_dmc_poly_of_edge(int mask, int le)
{
    int edges[16], start[8];
    const int np = or_dmc_polygons(mask, edges, start);
    for (int p = 0; p < np; p++)
        for (int k = start[p]; k < start[p + 1]; k++)
            if (edges[k] == le) return p;
    return -1;
}

static inline int lab(const float* d, int dim, int x, int y, int z)
{
    return d[((int64_t)x * dim + y) * dim + z] < 0.0f;
}

static int cell_mask(const float* d, int dim, int x, int y, int z)
{
    int m = 0;
    for (int dx = 0; dx < 2; dx++)
        for (int dy = 0; dy < 2; dy++)
            for (int dz = 0; dz < 2; dz++) m |= lab(d, dim, x + dx, y + dy, z + dz) << (4 * dx + 2 * dy + dz);
    return m;
}

/* marching-cubes vertex of local edge le of cell (x,y,z), grid units */
static void edge_point(const float* d, int dim, int x, int y, int z, int le, float q[3])
{
    const int x0 = x + LE[le][0], y0 = y + LE[le][1], z0 = z + LE[le][2], a = LE[le][3];
    const int x1 = x0 + (a == 0), y1 = y0 + (a == 1), z1 = z0 + (a == 2);
    const float s0 = d[((int64_t)x0 * dim + y0) * dim + z0], s1 = d[((int64_t)x1 * dim + y1) * dim + z1];
    const float mu = (0.0f - s0) / (s1 - s0);
    const float p0[3] = {(float)x0, (float)y0, (float)z0}, p1[3] = {(float)x1, (float)y1, (float)z1};
    for (int k = 0; k < 3; k++) q[k] = (p1[k] - p0[k]) * mu + p0[k];
}

int or_dmc_mesh(const or_sampler_t* s, const float* dens, int dim, const float opos[3], float delta, int placement,
                float tol, int as_triangles, float* pos, uint8_t* boundary, uint32_t* code, float* qerr,
                int64_t capV, uint32_t* idx, int64_t capI, int64_t* V_out, int64_t* nI_out)
{
    const int c = dim - 1;
    int64_t* base = (int64_t*)malloc(sizeof(int64_t) * (size_t)((int64_t)c * c * c + 1));
    int64_t V = 0, nI = 0;
    int overflow = 0;
    for (int x = 0; x < c; x++)
        for (int y = 0; y < c; y++)
            for (int z = 0; z < c; z++) {
                const int64_t ci = ((int64_t)x * c + y) * c + z;
                const int m = cell_mask(dens, dim, x, y, z);
                base[ci] = -1;
                if (m == 0 || m == 255) continue;
                base[ci] = V;
                int edges[16], start[8];
                const int np = or_dmc_polygons(m, edges, start);
                for (int p = 0; p < np; p++, V++) {
                    if (V >= capV) {
                        overflow = 1;
                        continue;
                    }
                    const int k = start[p + 1] - start[p];
                    float q[12][3], nrm[12][3];
                    float acc[3] = {0.0f, 0.0f, 0.0f};
                    for (int j = 0; j < k; j++) {
                        edge_point(dens, dim, x, y, z, edges[start[p] + j], q[j]);
                        for (int a = 0; a < 3; a++) acc[a] = acc[a] + q[j][a];
                    }
                    float err = 0.0f;
                    if (placement == 1 && s) {
                        for (int j = 0; j < k; j++) {
                            float g[3];
                            const float wx = q[j][0] * delta + opos[0], wy = q[j][1] * delta + opos[1],
                                        wz = q[j][2] * delta + opos[2];
                            or_csg_value_f32(s, wx, wy, wz, g);
                            const float len = sqrtf((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
                            for (int a = 0; a < 3; a++) nrm[j][a] = len > 0.0f ? g[a] / len : 0.0f;
                        }
                        float out[3];
                        err = or_qef_solve_f32(&q[0][0], &nrm[0][0], k, tol, out);
                        for (int a = 0; a < 3; a++) pos[3 * V + a] = out[a];
                    } else {
                        for (int a = 0; a < 3; a++) pos[3 * V + a] = acc[a] / (float)k;
                    }
                    boundary[V] = x == 0 || y == 0 || z == 0 || x + 1 == c || y + 1 == c || z + 1 == c;
                    if (code) code[V] = (uint32_t)x | ((uint32_t)y << 8) | ((uint32_t)z << 16) | ((uint32_t)p << 24);
                    if (qerr) qerr[V] = err;
                }
            }
    /* quads around interior sign-changing edges, edge order */
    for (int x = 0; x < dim; x++)
        for (int y = 0; y < dim; y++)
            for (int z = 0; z < dim; z++)
                for (int a = 0; a < 3; a++) {
                    const int x1 = x + (a == 0), y1 = y + (a == 1), z1 = z + (a == 2);
                    if (x1 >= dim || y1 >= dim || z1 >= dim) continue;
                    const int l0 = lab(dens, dim, x, y, z);
                    if (l0 == lab(dens, dim, x1, y1, z1)) continue;
                    int64_t v[4];
                    int ok = 1;
                    for (int r = 0; r < 4 && ok; r++) {
                        const int cx = x + RING[a][r][0], cy = y + RING[a][r][1], cz = z + RING[a][r][2];
                        if (cx < 0 || cy < 0 || cz < 0 || cx >= c || cy >= c || cz >= c) {
                            ok = 0;
                            break;
                        }
                        const int64_t ci = ((int64_t)cx * c + cy) * c + cz;
                        const int pi = or_dmc_poly_of_edge(cell_mask(dens, dim, cx, cy, cz), RING[a][r][3]);
                        if (base[ci] < 0 || pi < 0) {
                            overflow = 2;
                            ok = 0;
                            break;
                        }
                        v[r] = base[ci] + pi;
                    }
                    if (!ok) continue;
                    /* cyclic order c0, c1, c3, c2 has its right-hand normal along
                     * -x, +y, -z for x-, y-, z-edges; flip so it points to the
                     * solid end (label 0) */
                    const int solid_origin = !l0;
                    const int toward_origin = a != 1;
                    int64_t q[4] = {v[0], v[1], v[3], v[2]};
                    if (solid_origin != toward_origin) {
                        q[1] = v[2];
                        q[3] = v[1];
                    }
                    const int nq = as_triangles ? 6 : 4;
                    if (nI + nq > capI) {
                        overflow = 1;
                        nI += nq;
                        continue;
                    }
                    if (as_triangles) {
                        const int64_t t[6] = {q[0], q[1], q[2], q[2], q[3], q[0]};
                        for (int j = 0; j < 6; j++) idx[nI++] = (uint32_t)t[j];
                    } else {
                        for (int j = 0; j < 4; j++) idx[nI++] = (uint32_t)q[j];
                    }
                }
    free(base);
    *V_out = V;
    *nI_out = nI;
    return overflow == 0 ? 0 : (overflow == 1 ? -1 : -2);
}

/* MeshProcessor<N> smoothing (MeshProcessor.cpp:130-306) for N-gons: dual
 * centroid (((0 + p0) + p1) ...) / N, primal mean over the ascending
 * adjacency, K steps; boundary vertices stay unless boundary_processing. */
#define SMOOTH_N_BODY(T)                                                                       \
    if (iters <= 0 || V == 0 || nI < N) return 0;                                              \
    const int64_t nP = nI / N;                                                                 \
    uint32_t* val = (uint32_t*)calloc((size_t)V + 1, sizeof(uint32_t));                        \
    uint32_t* off = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(V + 1));                     \
    uint32_t* nxt = (uint32_t*)calloc((size_t)V + 1, sizeof(uint32_t));                        \
    uint32_t* adj = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(N * nP + 1));                \
    T* cen = (T*)malloc(sizeof(T) * (size_t)(3 * nP + 3));                                     \
    for (int64_t i = 0; i < N * nP; i++) val[idx[i]]++;                                        \
    off[0] = 0;                                                                                \
    for (int64_t v = 0; v < V; v++) off[v + 1] = off[v] + val[v];                              \
    for (int64_t t = 0; t < nP; t++)                                                           \
        for (int k = 0; k < N; k++) {                                                          \
            const uint32_t v = idx[N * t + k];                                                 \
            adj[off[v] + nxt[v]++] = (uint32_t)t;                                              \
        }                                                                                      \
    for (int m = 0; m < iters; m++) {                                                          \
        for (int64_t t = 0; t < nP; t++)                                                       \
            for (int k = 0; k < 3; k++) {                                                      \
                T acc = (T)0;                                                                  \
                for (int j = 0; j < N; j++) acc = acc + pos[3 * idx[N * t + j] + k];           \
                cen[3 * t + k] = acc / (T)N;                                                   \
            }                                                                                  \
        for (int64_t v = 0; v < V; v++) {                                                      \
            if (nxt[v] == 0 || (!boundary_processing && boundary[v])) continue;                \
            T a[3] = {(T)0, (T)0, (T)0};                                                       \
            for (uint32_t q = off[v]; q < off[v] + nxt[v]; q++)                                \
                for (int k = 0; k < 3; k++) a[k] = a[k] + cen[3 * adj[q] + k];                 \
            for (int k = 0; k < 3; k++) pos[3 * v + k] = a[k] / (T)nxt[v];                     \
        }                                                                                      \
    }                                                                                          \
    free(val); free(off); free(nxt); free(adj); free(cen);                                     \
    return 0;

int or_smooth_n_f32(float* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int N,
                    int iters, int boundary_processing)
{
    SMOOTH_N_BODY(float)
}

int or_smooth_n_f64(double* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int N,
                    int iters, int boundary_processing)
{
    SMOOTH_N_BODY(double)
}

/* FP64 twin of the vertex placement of a quad mesh (same polygons, given by
 * code[]): FP64 densities and, for the QEF, FP64 gradients and solver. */
void or_dmc_positions_f64(const or_sampler_t* s, const double* dens, int dim, const double opos[3], double delta,
                          int placement, double tol, const uint32_t* code, int64_t V, double* pos, double* qerr)
{
    for (int64_t v = 0; v < V; v++) {
        const int x = (int)(code[v] & 255u), y = (int)((code[v] >> 8) & 255u), z = (int)((code[v] >> 16) & 255u);
        const int p = (int)(code[v] >> 24);
        int m = 0;
        for (int dx = 0; dx < 2; dx++)
            for (int dy = 0; dy < 2; dy++)
                for (int dz = 0; dz < 2; dz++)
                    m |= (dens[((int64_t)(x + dx) * dim + (y + dy)) * dim + (z + dz)] < 0.0) << (4 * dx + 2 * dy + dz);
        int edges[16], start[8];
        or_dmc_polygons(m, edges, start);
        const int k = start[p + 1] - start[p];
        double q[12][3], nrm[12][3], acc[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < k; j++) {
            const int le = edges[start[p] + j];
            const int x0 = x + LE[le][0], y0 = y + LE[le][1], z0 = z + LE[le][2], a = LE[le][3];
            const int x1 = x0 + (a == 0), y1 = y0 + (a == 1), z1 = z0 + (a == 2);
            const double s0 = dens[((int64_t)x0 * dim + y0) * dim + z0], s1 = dens[((int64_t)x1 * dim + y1) * dim + z1];
            const double mu = -s0 / (s1 - s0);
            q[j][0] = x0 + (a == 0 ? mu : 0.0);
            q[j][1] = y0 + (a == 1 ? mu : 0.0);
            q[j][2] = z0 + (a == 2 ? mu : 0.0);
            for (int c2 = 0; c2 < 3; c2++) acc[c2] += q[j][c2];
        }
        if (placement == 1 && s) {
            for (int j = 0; j < k; j++) {
                double g[3];
                or_csg_value_f64(s, q[j][0] * delta + opos[0], q[j][1] * delta + opos[1], q[j][2] * delta + opos[2], g);
                const double len = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                for (int c2 = 0; c2 < 3; c2++) nrm[j][c2] = len > 0.0 ? g[c2] / len : 0.0;
            }
            double out[3];
            const double e = or_qef_solve_f64(&q[0][0], &nrm[0][0], k, tol, out);
            for (int c2 = 0; c2 < 3; c2++) pos[3 * v + c2] = out[c2];
            if (qerr) qerr[v] = e;
        } else {
            for (int c2 = 0; c2 < 3; c2++) pos[3 * v + c2] = acc[c2] / k;
            if (qerr) qerr[v] = 0.0;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Vertex normals of the triangle path.  The reference computes them only    */
/* with smooth_normals (MeshProcessor.cpp:159-216, 271-287), which its live  */
/* path leaves off (ChunkGenerator.cpp:116); they are defined here from that */
/* branch for N = 3:                                                          */
/*   dual step m: face normal -cross(normalize(p0-p1), normalize(p0-p2)) of   */
/*     the positions at step m while m == 0 || m < max_norms || m < 3, with  */
/*     max_norms = min(K/2 - 3, 10); later steps take the sum of the three    */
/*     vertex normals;                                                        */
/*   primal steps: n = (sum of the ring's face normals, ascending) / count,   */
/*     normalised when m == 3 or (m == 0 and K <= 3), kept only if n.y != 0;  */
/*     every vertex with a ring gets a normal (positions of boundary          */
/*     vertices stay pinned as before);                                       */
/*   K = 0: one face-normal step and one primal step, positions unchanged;    */
/*   output: the final normal divided by its length (0 when zero).           */
/* normalize(v) = v / sqrt((x*x + y*y) + z*z); cross(a, b) =                  */
/* (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y).                  */
/* ------------------------------------------------------------------------ */
#define NORMALS_BODY(T, SQRT)                                                                   \
    if (V == 0 || nI < 3) return 0;                                                             \
    const int64_t nT = nI / 3;                                                                  \
    const int K = iters;                                                                        \
    const int steps = K > 0 ? K : 1;                                                            \
    const int max_norms = (K / 2 - 3 < 10) ? K / 2 - 3 : 10;                                    \
    uint32_t* val = (uint32_t*)calloc((size_t)V + 1, sizeof(uint32_t));                         \
    uint32_t* off = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(V + 1));                      \
    uint32_t* nxt = (uint32_t*)calloc((size_t)V + 1, sizeof(uint32_t));                         \
    uint32_t* adj = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(3 * nT + 1));                 \
    T* cen = (T*)malloc(sizeof(T) * (size_t)(3 * nT + 3));                                      \
    T* fn = (T*)malloc(sizeof(T) * (size_t)(3 * nT + 3));                                       \
    T* vn = (T*)calloc((size_t)(3 * V + 3), sizeof(T));                                         \
    for (int64_t i = 0; i < 3 * nT; i++) val[idx[i]]++;                                         \
    off[0] = 0;                                                                                 \
    for (int64_t v = 0; v < V; v++) off[v + 1] = off[v] + val[v];                               \
    for (int64_t t = 0; t < nT; t++)                                                            \
        for (int k = 0; k < 3; k++) {                                                           \
            const uint32_t v = idx[3 * t + k];                                                  \
            adj[off[v] + nxt[v]++] = (uint32_t)t;                                               \
        }                                                                                       \
    for (int m = 0; m < steps; m++) {                                                           \
        for (int64_t t = 0; t < nT; t++) {                                                      \
            const T* a = pos + 3 * idx[3 * t];                                                  \
            const T* b = pos + 3 * idx[3 * t + 1];                                              \
            const T* c = pos + 3 * idx[3 * t + 2];                                              \
            for (int k = 0; k < 3; k++) {                                                       \
                T acc = (T)0;                                                                   \
                acc = acc + a[k];                                                               \
                acc = acc + b[k];                                                               \
                acc = acc + c[k];                                                               \
                cen[3 * t + k] = acc / (T)3;                                                    \
            }                                                                                   \
            if (m == 0 || m < max_norms || m < 3) {                                             \
                T u[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};                               \
                T w[3] = {a[0] - c[0], a[1] - c[1], a[2] - c[2]};                               \
                const T lu = SQRT((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);                   \
                const T lw = SQRT((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);                   \
                for (int k = 0; k < 3; k++) {                                                   \
                    u[k] = u[k] / lu;                                                           \
                    w[k] = w[k] / lw;                                                           \
                }                                                                               \
                fn[3 * t + 0] = -(u[1] * w[2] - w[1] * u[2]);                                   \
                fn[3 * t + 1] = -(u[2] * w[0] - w[2] * u[0]);                                   \
                fn[3 * t + 2] = -(u[0] * w[1] - w[0] * u[1]);                                   \
            } else {                                                                            \
                for (int k = 0; k < 3; k++) {                                                   \
                    T s = (T)0;                                                                 \
                    s = s + vn[3 * idx[3 * t] + k];                                             \
                    s = s + vn[3 * idx[3 * t + 1] + k];                                         \
                    s = s + vn[3 * idx[3 * t + 2] + k];                                         \
                    fn[3 * t + k] = s;                                                          \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
        const int last = m == steps - 1;                                                        \
        const int unit = !last && (m == 3 || (m == 0 && K <= 3));                               \
        for (int64_t v = 0; v < V; v++) {                                                       \
            if (nxt[v] == 0) continue;                                                          \
            T ap[3] = {(T)0, (T)0, (T)0}, an[3] = {(T)0, (T)0, (T)0};                           \
            for (uint32_t q = off[v]; q < off[v] + nxt[v]; q++)                                 \
                for (int k = 0; k < 3; k++) {                                                   \
                    ap[k] = ap[k] + cen[3 * adj[q] + k];                                        \
                    an[k] = an[k] + fn[3 * adj[q] + k];                                         \
                }                                                                               \
            for (int k = 0; k < 3; k++) an[k] = an[k] / (T)nxt[v];                              \
            if (unit) {                                                                         \
                const T l = SQRT((an[0] * an[0] + an[1] * an[1]) + an[2] * an[2]);              \
                for (int k = 0; k < 3; k++) an[k] = an[k] / l;                                  \
            }                                                                                   \
            if (an[1] != (T)0)                                                                  \
                for (int k = 0; k < 3; k++) vn[3 * v + k] = an[k];                              \
            if (K > 0 && (boundary_processing || !boundary[v]))                                 \
                for (int k = 0; k < 3; k++) pos[3 * v + k] = ap[k] / (T)nxt[v];                 \
        }                                                                                       \
    }                                                                                           \
    for (int64_t v = 0; v < V; v++) {                                                           \
        const T* n = vn + 3 * v;                                                                \
        const T l = SQRT((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);                            \
        for (int k = 0; k < 3; k++) normals[3 * v + k] = l > (T)0 ? n[k] / l : (T)0;            \
    }                                                                                           \
    free(val); free(off); free(nxt); free(adj); free(cen); free(fn); free(vn);                  \
    return 0;

int or_smooth_normals_f32(float* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int iters,
                          int boundary_processing, float* normals)
{
    NORMALS_BODY(float, sqrtf)
}

int or_smooth_normals_f64(double* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int iters,
                          int boundary_processing, double* normals)
{
    NORMALS_BODY(double, sqrt)
}

/* ------------------------------------------------------------------------ */
/* Vertex normals of an N-gon mesh (N = 3 or 4): the smooth_normals branch  */
/* of MeshProcessor<N> (MeshProcessor.cpp:159-216, 271-287).  Same schedule */
/* as the triangle path above; the face normal of step m while              */
/* m == 0 || m < max_norms || m < 3 is                                       */
/*   N = 3: -cross(normalize(p0-p1), normalize(p0-p2));                     */
/*   N = 4: n1 = cross(normalize(p0-p1), normalize(p0-p2)),                 */
/*          n2 = cross(normalize(p2-p3), normalize(p2-p0)); a NaN n1.x takes */
/*          n2 (and (0,1,0) when that is NaN too), then a NaN n2.x takes n1; */
/*          the face normal is -normalize((n1 + n2) * 0.5)                  */
/* (MeshProcessor.cpp:187-205); later steps take the sum of the N corner    */
/* normals in corner order.  Centroids ((0 + p0) + ...)/N, rings in        */
/* ascending primitive order.                                               */
/* ------------------------------------------------------------------------ */
#define CROSS_NORM(T, SQRT, out, u0, u1, u2, w0, w1, w2)                                        \
    do {                                                                                        \
        const T lu_ = SQRT((u0 * u0 + u1 * u1) + u2 * u2);                                      \
        const T lw_ = SQRT((w0 * w0 + w1 * w1) + w2 * w2);                                      \
        const T a0_ = u0 / lu_, a1_ = u1 / lu_, a2_ = u2 / lu_;                                  \
        const T b0_ = w0 / lw_, b1_ = w1 / lw_, b2_ = w2 / lw_;                                  \
        out[0] = a1_ * b2_ - b1_ * a2_;                                                         \
        out[1] = a2_ * b0_ - b2_ * a0_;                                                         \
        out[2] = a0_ * b1_ - b0_ * a1_;                                                         \
    } while (0)

#define NORMALS_N_BODY(T, SQRT)                                                                 \
    if (V == 0 || nI < N || (N != 3 && N != 4)) return 0;                                       \
    const int64_t nP = nI / N;                                                                  \
    const int K = iters;                                                                        \
    const int steps = K > 0 ? K : 1;                                                            \
    const int max_norms = (K / 2 - 3 < 10) ? K / 2 - 3 : 10;                                    \
    uint32_t* val = (uint32_t*)calloc((size_t)V + 1, sizeof(uint32_t));                         \
    uint32_t* off = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(V + 1));                      \
    uint32_t* nxt = (uint32_t*)calloc((size_t)V + 1, sizeof(uint32_t));                         \
    uint32_t* adj = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(N * nP + 1));                 \
    T* cen = (T*)malloc(sizeof(T) * (size_t)(3 * nP + 3));                                      \
    T* fn = (T*)malloc(sizeof(T) * (size_t)(3 * nP + 3));                                       \
    T* vn = (T*)calloc((size_t)(3 * V + 3), sizeof(T));                                         \
    for (int64_t i = 0; i < N * nP; i++) val[idx[i]]++;                                         \
    off[0] = 0;                                                                                 \
    for (int64_t v = 0; v < V; v++) off[v + 1] = off[v] + val[v];                               \
    for (int64_t t = 0; t < nP; t++)                                                            \
        for (int k = 0; k < N; k++) {                                                           \
            const uint32_t v = idx[N * t + k];                                                  \
            adj[off[v] + nxt[v]++] = (uint32_t)t;                                               \
        }                                                                                       \
    for (int m = 0; m < steps; m++) {                                                           \
        for (int64_t t = 0; t < nP; t++) {                                                      \
            const T* p[4];                                                                      \
            for (int c = 0; c < N; c++) p[c] = pos + 3 * idx[N * t + c];                        \
            for (int k = 0; k < 3; k++) {                                                       \
                T acc = (T)0;                                                                   \
                for (int c = 0; c < N; c++) acc = acc + p[c][k];                                \
                cen[3 * t + k] = acc / (T)N;                                                    \
            }                                                                                   \
            if (m == 0 || m < max_norms || m < 3) {                                             \
                if (N == 3) {                                                                   \
                    T f[3];                                                                     \
                    CROSS_NORM(T, SQRT, f, (p[0][0] - p[1][0]), (p[0][1] - p[1][1]), (p[0][2] - p[1][2]), \
                               (p[0][0] - p[2][0]), (p[0][1] - p[2][1]), (p[0][2] - p[2][2]));  \
                    for (int k = 0; k < 3; k++) fn[3 * t + k] = -f[k];                          \
                } else {                                                                        \
                    T n1[3], n2[3];                                                             \
                    CROSS_NORM(T, SQRT, n1, (p[0][0] - p[1][0]), (p[0][1] - p[1][1]), (p[0][2] - p[1][2]), \
                               (p[0][0] - p[2][0]), (p[0][1] - p[2][1]), (p[0][2] - p[2][2]));  \
                    CROSS_NORM(T, SQRT, n2, (p[2][0] - p[3][0]), (p[2][1] - p[3][1]), (p[2][2] - p[3][2]), \
                               (p[2][0] - p[0][0]), (p[2][1] - p[0][1]), (p[2][2] - p[0][2]));  \
                    if (isnan(n1[0])) {                                                         \
                        for (int k = 0; k < 3; k++) n1[k] = n2[k];                              \
                        if (isnan(n1[0])) {                                                     \
                            n1[0] = (T)0;                                                       \
                            n1[1] = (T)1;                                                       \
                            n1[2] = (T)0;                                                       \
                        }                                                                       \
                    }                                                                           \
                    if (isnan(n2[0]))                                                           \
                        for (int k = 0; k < 3; k++) n2[k] = n1[k];                              \
                    T h[3];                                                                     \
                    for (int k = 0; k < 3; k++) h[k] = (n1[k] + n2[k]) * (T)0.5;               \
                    const T lh = SQRT((h[0] * h[0] + h[1] * h[1]) + h[2] * h[2]);               \
                    for (int k = 0; k < 3; k++) fn[3 * t + k] = -(h[k] / lh);                   \
                }                                                                               \
            } else {                                                                            \
                for (int k = 0; k < 3; k++) {                                                   \
                    T s = (T)0;                                                                 \
                    for (int c = 0; c < N; c++) s = s + vn[3 * idx[N * t + c] + k];             \
                    fn[3 * t + k] = s;                                                          \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
        const int last = m == steps - 1;                                                        \
        const int unit = !last && (m == 3 || (m == 0 && K <= 3));                               \
        for (int64_t v = 0; v < V; v++) {                                                       \
            if (nxt[v] == 0) continue;                                                          \
            T ap[3] = {(T)0, (T)0, (T)0}, an[3] = {(T)0, (T)0, (T)0};                           \
            for (uint32_t q = off[v]; q < off[v] + nxt[v]; q++)                                 \
                for (int k = 0; k < 3; k++) {                                                   \
                    ap[k] = ap[k] + cen[3 * adj[q] + k];                                        \
                    an[k] = an[k] + fn[3 * adj[q] + k];                                         \
                }                                                                               \
            for (int k = 0; k < 3; k++) an[k] = an[k] / (T)nxt[v];                              \
            if (unit) {                                                                         \
                const T l = SQRT((an[0] * an[0] + an[1] * an[1]) + an[2] * an[2]);              \
                for (int k = 0; k < 3; k++) an[k] = an[k] / l;                                  \
            }                                                                                   \
            if (an[1] != (T)0)                                                                  \
                for (int k = 0; k < 3; k++) vn[3 * v + k] = an[k];                              \
            if (K > 0 && (boundary_processing || !boundary[v]))                                 \
                for (int k = 0; k < 3; k++) pos[3 * v + k] = ap[k] / (T)nxt[v];                 \
        }                                                                                       \
    }                                                                                           \
    for (int64_t v = 0; v < V; v++) {                                                           \
        const T* n = vn + 3 * v;                                                                \
        const T l = SQRT((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);                            \
        for (int k = 0; k < 3; k++) normals[3 * v + k] = l > (T)0 ? n[k] / l : (T)0;            \
    }                                                                                           \
    free(val); free(off); free(nxt); free(adj); free(cen); free(fn); free(vn);                  \
    return 0;

int or_smooth_normals_n_f32(float* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int N,
                            int iters, int boundary_processing, float* normals)
{
    NORMALS_N_BODY(float, sqrtf)
}

int or_smooth_normals_n_f64(double* pos, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int N,
                            int iters, int boundary_processing, double* normals)
{
    NORMALS_N_BODY(double, sqrt)
}
