/*
 * CPU oracle for the differentiable extraction -- TEST INFRASTRUCTURE ONLY.
 * Only tests/ load this code; the product never does.
 *
 * The reference has no adjoint.  What is restated here is the forward it
 * does have, as a function of the densities with the topology frozen:
 *   vertex placement  p = x_e + axis * mu, mu = (0 - s0) / (s1 - s0)
 *                     (DMCChunk.cpp:657-674, DMCChunk::_get_intersection)
 *   smoothing         K dual/primal steps (MeshProcessor.cpp:130-306 as
 *                     driven by ChunkGenerator.cpp:112-124): linear in the
 *                     positions for a fixed topology
 * plus the nearest-target loss, its gradient, the adjoint (transpose) of
 * the two maps, a forward-mode product (JVP) and central finite
 * differences, all in FP64.
 *
 * Pins (tests/test_oracle_diff.py): the adjoint against the JVP by the
 * dot-product identity <J v, g> = <v, J^T g> and against central finite
 * differences of the FP64 forward; the kd-tree nearest target against numpy
 * brute force; the loss gradient against its closed form.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void edge_decode(uint32_t e, int dim, int64_t* i0, int64_t* i1, int* axis, int c[3])
{
    const int a = (int)(e % 3);
    const int64_t p = e / 3;
    c[2] = (int)(p % dim);
    c[1] = (int)((p / dim) % dim);
    c[0] = (int)(p / ((int64_t)dim * dim));
    *axis = a;
    *i0 = p;
    const int64_t st[3] = {(int64_t)dim * dim, dim, 1};
    *i1 = p + st[a];
}

void or_place_f64(const double* dens, int dim, const uint32_t* edge, int64_t V, double* pos)
{
    for (int64_t v = 0; v < V; v++) {
        int64_t i0, i1;
        int a, c[3];
        edge_decode(edge[v], dim, &i0, &i1, &a, c);
        const double s0 = dens[i0], s1 = dens[i1];
        const double mu = (0.0 - s0) / (s1 - s0);
        for (int k = 0; k < 3; k++) pos[3 * v + k] = (double)c[k] + (k == a ? mu : 0.0);
    }
}

/* ------------------------------------------------------------------------ */
/* Adjacency in MeshProcessor::set_adj order (triangle ids ascending).       */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t* off;  /* V + 1 */
    uint32_t* adj;  /* 3T */
    uint8_t* moved; /* V */
} ring_t;

static void ring_build(ring_t* r, const uint8_t* boundary, int64_t V, const uint32_t* idx, int64_t nI, int bp)
{
    const int64_t nT = nI / 3;
    uint32_t* cnt = (uint32_t*)calloc((size_t)V + 1, sizeof(uint32_t));
    r->off = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(V + 1));
    r->adj = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(3 * nT + 1));
    r->moved = (uint8_t*)malloc((size_t)V + 1);
    for (int64_t i = 0; i < 3 * nT; i++) cnt[idx[i]]++;
    r->off[0] = 0;
    for (int64_t v = 0; v < V; v++) r->off[v + 1] = r->off[v] + cnt[v];
    memset(cnt, 0, sizeof(uint32_t) * (size_t)V);
    for (int64_t t = 0; t < nT; t++)
        for (int k = 0; k < 3; k++) {
            const uint32_t v = idx[3 * t + k];
            r->adj[r->off[v] + cnt[v]++] = (uint32_t)t;
        }
    for (int64_t v = 0; v < V; v++) r->moved[v] = (uint8_t)((bp || !boundary[v]) && cnt[v] > 0);
    free(cnt);
}

static void ring_free(ring_t* r)
{
    free(r->off);
    free(r->adj);
    free(r->moved);
}

/* K dual/primal steps on pos (3V) with a prebuilt ring; cen is 3T scratch. */
static void smooth_ring(double* pos, int64_t V, const uint32_t* idx, int64_t nT, const ring_t* r, int K, double* cen)
{
    for (int m = 0; m < K; m++) {
        for (int64_t t = 0; t < nT; t++)
            for (int k = 0; k < 3; k++) {
                double a = 0.0;
                a = a + pos[3 * idx[3 * t] + k];
                a = a + pos[3 * idx[3 * t + 1] + k];
                a = a + pos[3 * idx[3 * t + 2] + k];
                cen[3 * t + k] = a / 3.0;
            }
        for (int64_t v = 0; v < V; v++) {
            if (!r->moved[v]) continue;
            double a[3] = {0.0, 0.0, 0.0};
            for (uint32_t q = r->off[v]; q < r->off[v + 1]; q++)
                for (int k = 0; k < 3; k++) a[k] = a[k] + cen[3 * r->adj[q] + k];
            const double n = (double)(r->off[v + 1] - r->off[v]);
            for (int k = 0; k < 3; k++) pos[3 * v + k] = a[k] / n;
        }
    }
}

void or_forward_f64(const double* dens, int dim, const uint32_t* edge, const uint8_t* boundary, int64_t V,
                    const uint32_t* idx, int64_t nI, int K, int bp, double* pos)
{
    or_place_f64(dens, dim, edge, V, pos);
    if (K <= 0 || V == 0 || nI < 3) return;
    ring_t r;
    ring_build(&r, boundary, V, idx, nI, bp);
    double* cen = (double*)malloc(sizeof(double) * (size_t)nI);
    smooth_ring(pos, V, idx, nI / 3, &r, K, cen);
    free(cen);
    ring_free(&r);
}

/* J v: the derivative of the final positions along a density direction.
 * Smoothing has no constant term, so it maps displacements like positions. */
void or_jvp_f64(const double* dens, int dim, const uint32_t* edge, const uint8_t* boundary, int64_t V,
                const uint32_t* idx, int64_t nI, int K, int bp, const double* ddens, double* dpos)
{
    for (int64_t v = 0; v < V; v++) {
        int64_t i0, i1;
        int a, c[3];
        edge_decode(edge[v], dim, &i0, &i1, &a, c);
        const double s0 = dens[i0], s1 = dens[i1], ds = s1 - s0;
        const double dmu = (-s1 / (ds * ds)) * ddens[i0] + (s0 / (ds * ds)) * ddens[i1];
        for (int k = 0; k < 3; k++) dpos[3 * v + k] = k == a ? dmu : 0.0;
    }
    if (K <= 0 || V == 0 || nI < 3) return;
    ring_t r;
    ring_build(&r, boundary, V, idx, nI, bp);
    double* cen = (double*)malloc(sizeof(double) * (size_t)nI);
    smooth_ring(dpos, V, idx, nI / 3, &r, K, cen);
    free(cen);
    ring_free(&r);
}

/* J^T g: transposed smoothing steps, then the placement derivative
 * scattered to both endpoint densities of each vertex's edge. */
void or_adjoint_f64(const double* dens, int dim, const uint32_t* edge, const uint8_t* boundary, int64_t V,
                    const uint32_t* idx, int64_t nI, int K, int bp, const double* dldp, double* dldd)
{
    const int64_t n3 = (int64_t)dim * dim * dim;
    const int64_t nT = nI / 3;
    double* g = (double*)malloc(sizeof(double) * (size_t)(3 * V + 1));
    memcpy(g, dldp, sizeof(double) * (size_t)(3 * V));
    if (K > 0 && V > 0 && nI >= 3) {
        ring_t r;
        ring_build(&r, boundary, V, idx, nI, bp);
        double* gc = (double*)malloc(sizeof(double) * (size_t)(3 * nT));
        for (int m = K - 1; m >= 0; m--) {
            for (int64_t t = 0; t < nT; t++)
                for (int k = 0; k < 3; k++) {
                    double a = 0.0;
                    for (int j = 0; j < 3; j++) {
                        const uint32_t u = idx[3 * t + j];
                        if (r.moved[u]) a += g[3 * u + k] / (double)(r.off[u + 1] - r.off[u]);
                    }
                    gc[3 * t + k] = a;
                }
            for (int64_t v = 0; v < V; v++) {
                double s[3] = {0.0, 0.0, 0.0};
                for (uint32_t q = r.off[v]; q < r.off[v + 1]; q++)
                    for (int k = 0; k < 3; k++) s[k] += gc[3 * r.adj[q] + k];
                for (int k = 0; k < 3; k++) g[3 * v + k] = (r.moved[v] ? 0.0 : g[3 * v + k]) + s[k] / 3.0;
            }
        }
        free(gc);
        ring_free(&r);
    }
    memset(dldd, 0, sizeof(double) * (size_t)n3);
    for (int64_t v = 0; v < V; v++) {
        int64_t i0, i1;
        int a, c[3];
        edge_decode(edge[v], dim, &i0, &i1, &a, c);
        const double s0 = dens[i0], s1 = dens[i1], ds = s1 - s0;
        dldd[i0] += g[3 * v + a] * (-s1 / (ds * ds));
        dldd[i1] += g[3 * v + a] * (s0 / (ds * ds));
    }
    free(g);
}

/* ------------------------------------------------------------------------ */
/* Exact nearest target: kd-tree with per-node boxes; ties go to the lower   */
/* target index, so a node is pruned only when its box is strictly farther   */
/* than the best distance.                                                    */
/* ------------------------------------------------------------------------ */
#define KD_LEAF 8

struct or_kdtree {
    int64_t n;
    double* pts;   /* n x 3, permuted */
    int64_t* id;   /* original index of each permuted point */
    int64_t nnodes;
    int64_t* lo;   /* node range [lo, hi) */
    int64_t* hi;
    int64_t* left; /* child node ids, -1 for leaves */
    int64_t* right;
    double* box;   /* node x 6: min xyz, max xyz */
};

static int64_t kd_build_rec(or_kdtree* t, int64_t lo, int64_t hi)
{
    const int64_t node = t->nnodes++;
    double* b = t->box + 6 * node;
    for (int k = 0; k < 3; k++) {
        b[k] = DBL_MAX;
        b[3 + k] = -DBL_MAX;
    }
    for (int64_t i = lo; i < hi; i++)
        for (int k = 0; k < 3; k++) {
            const double x = t->pts[3 * i + k];
            if (x < b[k]) b[k] = x;
            if (x > b[3 + k]) b[3 + k] = x;
        }
    t->lo[node] = lo;
    t->hi[node] = hi;
    t->left[node] = t->right[node] = -1;
    if (hi - lo <= KD_LEAF) return node;
    int ax = 0;
    for (int k = 1; k < 3; k++)
        if (b[3 + k] - b[k] > b[3 + ax] - b[ax]) ax = k;
    /* quickselect the median along ax */
    const int64_t mid = (lo + hi) / 2;
    int64_t l = lo, r = hi - 1;
    while (l < r) {
        const double pv = t->pts[3 * ((l + r) / 2) + ax];
        int64_t i = l, j = r;
        while (i <= j) {
            while (t->pts[3 * i + ax] < pv) i++;
            while (t->pts[3 * j + ax] > pv) j--;
            if (i <= j) {
                for (int k = 0; k < 3; k++) {
                    const double tmp = t->pts[3 * i + k];
                    t->pts[3 * i + k] = t->pts[3 * j + k];
                    t->pts[3 * j + k] = tmp;
                }
                const int64_t ti = t->id[i];
                t->id[i] = t->id[j];
                t->id[j] = ti;
                i++;
                j--;
            }
        }
        if (mid <= j) r = j;
        else if (mid >= i) l = i;
        else break;
    }
    const int64_t a = kd_build_rec(t, lo, mid);
    const int64_t c = kd_build_rec(t, mid, hi);
    t->left[node] = a;
    t->right[node] = c;
    return node;
}

or_kdtree* or_kd_build(const double* pts, int64_t n)
{
    or_kdtree* t = (or_kdtree*)calloc(1, sizeof(or_kdtree));
    t->n = n;
    t->pts = (double*)malloc(sizeof(double) * (size_t)(3 * n + 3));
    t->id = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    memcpy(t->pts, pts, sizeof(double) * (size_t)(3 * n));
    for (int64_t i = 0; i < n; i++) t->id[i] = i;
    const int64_t cap = 2 * (n / KD_LEAF + 1) * 2 + 8;
    t->lo = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    t->hi = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    t->left = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    t->right = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    t->box = (double*)malloc(sizeof(double) * 6 * (size_t)cap);
    if (n > 0) kd_build_rec(t, 0, n);
    return t;
}

void or_kd_free(or_kdtree* t)
{
    if (!t) return;
    free(t->pts); free(t->id); free(t->lo); free(t->hi); free(t->left); free(t->right); free(t->box);
    free(t);
}

static double box_lb2(const double* b, const double* p)
{
    double s = 0.0;
    for (int k = 0; k < 3; k++) {
        double d = 0.0;
        if (p[k] < b[k]) d = b[k] - p[k];
        else if (p[k] > b[3 + k]) d = p[k] - b[3 + k];
        s += d * d;
    }
    return s;
}

static void kd_query_one(const or_kdtree* t, const double* p, int64_t* best_id, double* best_d2)
{
    int64_t stack[256];
    int sp = 0;
    double best = DBL_MAX;
    int64_t bi = INT64_MAX;
    stack[sp++] = 0;
    while (sp > 0) {
        const int64_t node = stack[--sp];
        if (box_lb2(t->box + 6 * node, p) > best) continue;
        if (t->left[node] < 0) {
            for (int64_t i = t->lo[node]; i < t->hi[node]; i++) {
                const double dx = p[0] - t->pts[3 * i], dy = p[1] - t->pts[3 * i + 1], dz = p[2] - t->pts[3 * i + 2];
                const double d = dx * dx + dy * dy + dz * dz;
                if (d < best || (d == best && t->id[i] < bi)) {
                    best = d;
                    bi = t->id[i];
                }
            }
            continue;
        }
        const int64_t a = t->left[node], b = t->right[node];
        const double la = box_lb2(t->box + 6 * a, p), lb = box_lb2(t->box + 6 * b, p);
        /* push the farther child first so the nearer one is visited next */
        if (la <= lb) {
            stack[sp++] = b;
            stack[sp++] = a;
        } else {
            stack[sp++] = a;
            stack[sp++] = b;
        }
    }
    *best_id = bi;
    *best_d2 = best;
}

void or_kd_nearest(const or_kdtree* t, const double* q, int64_t m, int64_t* out, double* d2, int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < m; i++) {
        double dd;
        kd_query_one(t, q + 3 * i, &out[i], &dd);
        if (d2) d2[i] = dd;
    }
}

/* Nearest target in FP32 arithmetic: squared distance ((dx*dx + dy*dy) +
 * dz*dz) with every operation rounded to FP32, ties to the lower target
 * index.  The tree holds the FP32 coordinates exactly (as doubles); a box is
 * pruned only when its FP64 distance, shrunk by 4e-6 relative (more than the
 * four FP32 roundings of a distance can take off), is still above the best
 * FP32 distance, so the result is that of a brute-force FP32 scan.          */
static void kd_query_one_f32(const or_kdtree* t, const float* p, int64_t* best_id, float* best_d2)
{
    int64_t stack[256];
    int sp = 0;
    const double pd[3] = {p[0], p[1], p[2]};
    float best = INFINITY;
    int64_t bi = INT64_MAX;
    stack[sp++] = 0;
    while (sp > 0) {
        const int64_t node = stack[--sp];
        if (box_lb2(t->box + 6 * node, pd) * (1.0 - 4e-6) > (double)best + 1e-30) continue;
        if (t->left[node] < 0) {
            for (int64_t i = t->lo[node]; i < t->hi[node]; i++) {
                const float dx = p[0] - (float)t->pts[3 * i], dy = p[1] - (float)t->pts[3 * i + 1],
                            dz = p[2] - (float)t->pts[3 * i + 2];
                const float d = (dx * dx + dy * dy) + dz * dz;
                if (d < best || (d == best && t->id[i] < bi)) {
                    best = d;
                    bi = t->id[i];
                }
            }
            continue;
        }
        const int64_t a = t->left[node], b = t->right[node];
        const double la = box_lb2(t->box + 6 * a, pd), lb = box_lb2(t->box + 6 * b, pd);
        if (la <= lb) {
            stack[sp++] = b;
            stack[sp++] = a;
        } else {
            stack[sp++] = a;
            stack[sp++] = b;
        }
    }
    *best_id = bi;
    *best_d2 = best;
}

void or_kd_nearest_f32(const or_kdtree* t, const float* q, int64_t m, int64_t* out, float* d2, int threads)
{
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < m; i++) {
        float dd;
        kd_query_one_f32(t, q + 3 * i, &out[i], &dd);
        if (d2) d2[i] = dd;
    }
}

double or_loss_nearest_f64(const double* pos, int64_t V, const double* targets, const int64_t* nearest, double* dldp)
{
    if (V == 0) return 0.0;
    double L = 0.0;
    for (int64_t v = 0; v < V; v++) {
        const double* t = targets + 3 * nearest[v];
        double d2 = 0.0;
        for (int k = 0; k < 3; k++) {
            const double d = pos[3 * v + k] - t[k];
            d2 += d * d;
            if (dldp) dldp[3 * v + k] = 2.0 * d / (double)V;
        }
        L += d2;
    }
    return L / (double)V;
}

/* Central differences of the FP64 loss along single densities: the sum
 * over vertices of (d2(+h) - d2(-h)) / (V * 2h), with the nearest target
 * recomputed for every vertex whose position changed.  kink[s] (may be
 * NULL) is set when some vertex's nearest target at +h or -h differs from
 * the unperturbed one: the loss is not differentiable across such a switch
 * and the difference quotient there does not estimate a gradient. */
void or_fd_grad_f64(const double* dens, int dim, const uint32_t* edge, const uint8_t* boundary, int64_t V,
                    const uint32_t* idx, int64_t nI, int K, int bp, const or_kdtree* kd, const int64_t* samples,
                    int m, double h, int threads, double* out, int32_t* kink)
{
    const int64_t n3 = (int64_t)dim * dim * dim;
    double* base = (double*)malloc(sizeof(double) * (size_t)(3 * V + 3));
    or_forward_f64(dens, dim, edge, boundary, V, idx, nI, K, bp, base);
    ring_t r;
    const int smooth = K > 0 && V > 0 && nI >= 3;
    if (smooth) ring_build(&r, boundary, V, idx, nI, bp);
    if (threads < 1) threads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads)
#endif
    {
        double* d = (double*)malloc(sizeof(double) * (size_t)n3);
        double* pp = (double*)malloc(sizeof(double) * (size_t)(3 * V + 3));
        double* pm = (double*)malloc(sizeof(double) * (size_t)(3 * V + 3));
        double* cen = (double*)malloc(sizeof(double) * (size_t)(nI + 3));
        memcpy(d, dens, sizeof(double) * (size_t)n3);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int s = 0; s < m; s++) {
            const int64_t i = samples[s];
            const double d0 = d[i];
            d[i] = d0 + h;
            or_place_f64(d, dim, edge, V, pp);
            if (smooth) smooth_ring(pp, V, idx, nI / 3, &r, K, cen);
            d[i] = d0 - h;
            or_place_f64(d, dim, edge, V, pm);
            if (smooth) smooth_ring(pm, V, idx, nI / 3, &r, K, cen);
            d[i] = d0;
            double acc = 0.0;
            int switched = 0;
            for (int64_t v = 0; v < V; v++) {
                const double* a = pp + 3 * v;
                const double* b = pm + 3 * v;
                const double* c = base + 3 * v;
                if (a[0] == c[0] && a[1] == c[1] && a[2] == c[2] && b[0] == c[0] && b[1] == c[1] && b[2] == c[2])
                    continue;
                int64_t ia, ib, ic;
                double da, db, dc;
                kd_query_one(kd, a, &ia, &da);
                kd_query_one(kd, b, &ib, &db);
                kd_query_one(kd, c, &ic, &dc);
                if (ia != ic || ib != ic) switched = 1;
                acc += da - db;
            }
            out[s] = acc / ((double)V * 2.0 * h);
            if (kink) kink[s] = switched;
        }
        free(d); free(pp); free(pm); free(cen);
    }
    if (smooth) ring_free(&r);
    free(base);
}
