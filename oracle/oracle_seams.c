/*
 * oracle_seams.c -- CPU restatement of the stitcher: the marching-cubes seam
 * triangles between the chunks of a world octree (test infrastructure only;
 * never linked into the product).
 *
 * Reference.  DMCChunk::generate_octree (DMCChunk.cpp:700-778) gives every
 * chunk an octree of its boundary shell: from the root (the chunk's pos,
 * size, lattice size dim, a power of two) each node splits into the children
 * at cxyz = xyz + MCD[i] * (isize / 2) that touch a face of the chunk
 * (a coordinate 0 or dim - isize / 2), placed at pos + MCD[i] * (size / 2),
 * down to unit nodes; ChunkGenerator::extract_octrees (ChunkGenerator.cpp:
 * 217-240) hangs the chunk octree's children under the world leaf.
 * WorldStitcher::stitch_all (WorldStitcher.cpp:26-49) runs stitch_cell
 * (:184-239) on the world octree's inner nodes: the faces and edges between
 * a node's children, recursed by stitch_face_xy / zy / xz and stitch_edge_x /
 * y / z (:241-460) down to the leaves on both sides, and at every level
 * stitch_indexes (:491-572) on the 8 nodes around the configuration's centre
 * point, each descended to its leaf touching that point: marching cubes on
 * the 8 leaves (corner c = the leaf in octant c, 4*dx + 2*dy + dz; position
 * the leaf's pos, value the leaf's sample; crossing (v1 - v0) * mu + v0 with
 * mu = (0 - s0) / (s1 - s0); triangles from MCTable.h's tri_table), written
 * as a triangle soup.
 *
 * Definitions of this build where the reference's code leaves the result
 * undefined or empty (its stitcher is off on the live path: enable_stitching
 * is false, WorldOctree.cpp:31, and generate_octree is not called,
 * ChunkGenerator.cpp:103-108):
 *   - a unit node's sample is the sampler's value at the node's position
 *     (the reference's line that would read the chunk's densities is
 *     commented out, DMCChunk.cpp:755, so every sample is 0 and no triangle
 *     is ever written);
 *   - chunks without a mesh get the same shell octree (the reference gives
 *     them 8 children with unset positions, DMCChunk.cpp:710-728), and the
 *     faces between two chunks without a mesh are stitched too (the
 *     reference skips them, :249-256, leaving a crack where one chunk is
 *     all solid and the other all air).
 *
 * Two formulations, checked against each other by the tests:
 *   or_stitch_recursive  the reference's recursion, in its order;
 *   or_stitch_points     the same dual cells enumerated point by point, in
 *                        the order of the GPU kernels: for each chunk of the
 *                        leaf list, every lattice point of its closed cube
 *                        [0, dim]^3 on the cube's boundary, in x-major order;
 *                        a point is stitched when its 8 octant neighbourhoods
 *                        lie in the world but not all in one chunk, by the
 *                        finest chunk among them (the lowest node index
 *                        among equals); corner c is that chunk's unit node
 *                        containing the point's octant c.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "../include/bmf_mc_tables.h"

static const int MCD[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                              {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}}; /* Tables.hpp:7-9 */
/* corners of the 12 local edges (a, b) in the tables' numbering */
static const int EC[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                              {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};

/* Marching cubes on one dual cell; appends its triangles (9 floats each). */
static int64_t mc_cell(const float pos[8][3], const float sample[8], float* out, int64_t nt, int64_t cap)
{
    int mask = 0;
    for (int c = 0; c < 8; c++)
        if (sample[c] < 0.0f) mask |= 1 << c;
    if (mask == 0 || mask == 255) return nt;
    float cross[12][3];
    for (int e = 0; e < 12; e++) {
        const int a = EC[e][0], b = EC[e][1];
        if (((mask >> a) & 1) == ((mask >> b) & 1)) continue;
        const float mu = (0.0f - sample[a]) / (sample[b] - sample[a]);
        for (int k = 0; k < 3; k++) {
            const float d = (pos[b][k] - pos[a][k]) * mu;
            cross[e][k] = d + pos[a][k];
        }
    }
    const uint64_t row = bmf_tri_table[mask];
    const int n = (int)(row >> 60);
    for (int s = 0; s < n; s += 3, nt++) {
        if (nt >= cap) continue;
        for (int v = 0; v < 3; v++) {
            const int e = (int)((row >> (4 * (s + v))) & 15);
            memcpy(out + 9 * nt + 3 * v, cross[e], sizeof(float) * 3);
        }
    }
    return nt;
}

/* ======================= the reference's recursion ======================= */
typedef struct {
    float pos[3];
    float size;
    int world, world_leaf, leaf;
    int kid[8];
    float sample;
} snode_t;

typedef struct {
    snode_t* n;
    int count, cap;
    float* out;
    int64_t nt, cap_t;
} stitcher_t;

static int new_node(stitcher_t* S)
{
    if (S->count == S->cap) {
        S->cap = S->cap ? 2 * S->cap : 1024;
        S->n = (snode_t*)realloc(S->n, sizeof(snode_t) * (size_t)S->cap);
    }
    snode_t* q = &S->n[S->count];
    memset(q, 0, sizeof(*q));
    for (int i = 0; i < 8; i++) q->kid[i] = -1;
    return S->count++;
}

#define N(i) (S->n[i])
#define KID(i, c) (N(i).kid[c])
#define LEAF(i) (N(i).leaf)

static void indexes(stitcher_t* S, int nodes[8])
{
    static const int toward[8] = {6, 7, 4, 5, 2, 3, 0, 1};
    static const int corner_of[8] = {0, 3, 4, 7, 1, 2, 5, 6};
    for (int i = 0; i < 8; i++)
        if (nodes[i] < 0) return;
    for (int i = 0; i < 8; i++)
        while (nodes[i] >= 0 && !LEAF(nodes[i])) nodes[i] = KID(nodes[i], toward[i]);
    for (int i = 0; i < 8; i++)
        if (nodes[i] < 0) return;
    float pos[8][3], smp[8];
    for (int c = 0; c < 8; c++) {
        const snode_t* q = &N(nodes[corner_of[c]]);
        memcpy(pos[c], q->pos, sizeof(pos[c]));
        smp[c] = q->sample;
    }
    S->nt = mc_cell(pos, smp, S->out, S->nt, S->cap_t);
}

static void edge_x(stitcher_t* S, int n0, int n1, int n2, int n3)
{
    if (n0 < 0 || n1 < 0 || n2 < 0 || n3 < 0 || (LEAF(n0) && LEAF(n1) && LEAF(n2) && LEAF(n3))) return;
    int c[8] = {LEAF(n0) ? n0 : KID(n0, 7), LEAF(n0) ? n0 : KID(n0, 6), LEAF(n1) ? n1 : KID(n1, 5),
                LEAF(n1) ? n1 : KID(n1, 4), LEAF(n3) ? n3 : KID(n3, 3), LEAF(n3) ? n3 : KID(n3, 2),
                LEAF(n2) ? n2 : KID(n2, 1), LEAF(n2) ? n2 : KID(n2, 0)};
    edge_x(S, c[0], c[3], c[7], c[4]);
    edge_x(S, c[1], c[2], c[6], c[5]);
    indexes(S, c);
}

static void edge_y(stitcher_t* S, int n0, int n1, int n2, int n3)
{
    if (n0 < 0 || n1 < 0 || n2 < 0 || n3 < 0 || (LEAF(n0) && LEAF(n1) && LEAF(n2) && LEAF(n3))) return;
    int c[8] = {LEAF(n0) ? n0 : KID(n0, 2), LEAF(n1) ? n1 : KID(n1, 3), LEAF(n2) ? n2 : KID(n2, 0),
                LEAF(n3) ? n3 : KID(n3, 1), LEAF(n0) ? n0 : KID(n0, 6), LEAF(n1) ? n1 : KID(n1, 7),
                LEAF(n2) ? n2 : KID(n2, 4), LEAF(n3) ? n3 : KID(n3, 5)};
    edge_y(S, c[0], c[1], c[2], c[3]);
    edge_y(S, c[4], c[5], c[6], c[7]);
    indexes(S, c);
}

static void edge_z(stitcher_t* S, int n0, int n1, int n2, int n3)
{
    if (n0 < 0 || n1 < 0 || n2 < 0 || n3 < 0 || (LEAF(n0) && LEAF(n1) && LEAF(n2) && LEAF(n3))) return;
    int c[8] = {LEAF(n3) ? n3 : KID(n3, 5), LEAF(n2) ? n2 : KID(n2, 4), LEAF(n2) ? n2 : KID(n2, 7),
                LEAF(n3) ? n3 : KID(n3, 6), LEAF(n0) ? n0 : KID(n0, 1), LEAF(n1) ? n1 : KID(n1, 0),
                LEAF(n1) ? n1 : KID(n1, 3), LEAF(n0) ? n0 : KID(n0, 2)};
    edge_z(S, c[7], c[6], c[2], c[3]);
    edge_z(S, c[4], c[5], c[1], c[0]);
    indexes(S, c);
}

static void face_xy(stitcher_t* S, int n0, int n1)
{
    if (LEAF(n0) && LEAF(n1)) return;
    int c[8] = {LEAF(n0) ? n0 : KID(n0, 3), LEAF(n0) ? n0 : KID(n0, 2), LEAF(n1) ? n1 : KID(n1, 1),
                LEAF(n1) ? n1 : KID(n1, 0), LEAF(n0) ? n0 : KID(n0, 7), LEAF(n0) ? n0 : KID(n0, 6),
                LEAF(n1) ? n1 : KID(n1, 5), LEAF(n1) ? n1 : KID(n1, 4)};
    if (c[0] >= 0 && c[3] >= 0) face_xy(S, c[0], c[3]);
    if (c[1] >= 0 && c[2] >= 0) face_xy(S, c[1], c[2]);
    if (c[4] >= 0 && c[7] >= 0) face_xy(S, c[4], c[7]);
    if (c[5] >= 0 && c[6] >= 0) face_xy(S, c[5], c[6]);
    if (c[0] >= 0 && c[3] >= 0 && c[7] >= 0 && c[4] >= 0) edge_x(S, c[0], c[3], c[7], c[4]);
    if (c[1] >= 0 && c[2] >= 0 && c[6] >= 0 && c[5] >= 0) edge_x(S, c[1], c[2], c[6], c[5]);
    if (c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[3] >= 0) edge_y(S, c[0], c[1], c[2], c[3]);
    if (c[4] >= 0 && c[5] >= 0 && c[6] >= 0 && c[7] >= 0) edge_y(S, c[4], c[5], c[6], c[7]);
    indexes(S, c);
}

static void face_zy(stitcher_t* S, int n0, int n1)
{
    if (n0 < 0 || n1 < 0 || (LEAF(n0) && LEAF(n1))) return;
    int c[8] = {LEAF(n0) ? n0 : KID(n0, 1), LEAF(n1) ? n1 : KID(n1, 0), LEAF(n1) ? n1 : KID(n1, 3),
                LEAF(n0) ? n0 : KID(n0, 2), LEAF(n0) ? n0 : KID(n0, 5), LEAF(n1) ? n1 : KID(n1, 4),
                LEAF(n1) ? n1 : KID(n1, 7), LEAF(n0) ? n0 : KID(n0, 6)};
    face_zy(S, c[0], c[1]);
    face_zy(S, c[3], c[2]);
    face_zy(S, c[4], c[5]);
    face_zy(S, c[7], c[6]);
    edge_y(S, c[0], c[1], c[2], c[3]);
    edge_y(S, c[4], c[5], c[6], c[7]);
    edge_z(S, c[7], c[6], c[2], c[3]);
    edge_z(S, c[4], c[5], c[1], c[0]);
    indexes(S, c);
}

static void face_xz(stitcher_t* S, int n0, int n1)
{
    if (n0 < 0 || n1 < 0 || (LEAF(n0) && LEAF(n1))) return;
    int c[8] = {LEAF(n1) ? n1 : KID(n1, 4), LEAF(n1) ? n1 : KID(n1, 5), LEAF(n1) ? n1 : KID(n1, 6),
                LEAF(n1) ? n1 : KID(n1, 7), LEAF(n0) ? n0 : KID(n0, 0), LEAF(n0) ? n0 : KID(n0, 1),
                LEAF(n0) ? n0 : KID(n0, 2), LEAF(n0) ? n0 : KID(n0, 3)};
    face_xz(S, c[4], c[0]);
    face_xz(S, c[5], c[1]);
    face_xz(S, c[7], c[3]);
    face_xz(S, c[6], c[2]);
    edge_x(S, c[0], c[3], c[7], c[4]);
    edge_x(S, c[1], c[2], c[6], c[5]);
    edge_z(S, c[7], c[6], c[2], c[3]);
    edge_z(S, c[4], c[5], c[1], c[0]);
    indexes(S, c);
}

static void cell(stitcher_t* S, int n)
{
    if (LEAF(n) || !N(n).world || N(n).world_leaf) return;
    int c[8];
    memcpy(c, N(n).kid, sizeof(c));
    if (c[0] >= 0 && c[3] >= 0) face_xy(S, c[0], c[3]);
    if (c[1] >= 0 && c[2] >= 0) face_xy(S, c[1], c[2]);
    if (c[4] >= 0 && c[7] >= 0) face_xy(S, c[4], c[7]);
    if (c[5] >= 0 && c[6] >= 0) face_xy(S, c[5], c[6]);
    face_zy(S, c[0], c[1]);
    face_zy(S, c[3], c[2]);
    face_zy(S, c[4], c[5]);
    face_zy(S, c[7], c[6]);
    face_xz(S, c[4], c[0]);
    face_xz(S, c[5], c[1]);
    face_xz(S, c[7], c[3]);
    face_xz(S, c[6], c[2]);
    edge_x(S, c[0], c[3], c[7], c[4]);
    edge_x(S, c[1], c[2], c[6], c[5]);
    edge_y(S, c[0], c[1], c[2], c[3]);
    edge_y(S, c[4], c[5], c[6], c[7]);
    edge_z(S, c[7], c[6], c[2], c[3]);
    edge_z(S, c[4], c[5], c[1], c[0]);
    int cc[8];
    memcpy(cc, c, sizeof(cc));
    indexes(S, cc);
}

/* The shell octree of a chunk (generate_octree), its root's children
 * returned in kid[8]. */
static void chunk_octree(stitcher_t* S, const or_sampler_t* s, const float pos[3], float size, int dim, int kid[8])
{
    typedef struct { int node; int xyz[3]; int isize; float pos[3]; float size; } item_t;
    item_t* q = (item_t*)malloc(sizeof(item_t) * 16);
    int qcap = 16, head = 0, tail = 0;
    item_t root = {-1, {0, 0, 0}, dim, {pos[0], pos[1], pos[2]}, size};
    q[tail++] = root;
    while (head < tail) {
        item_t it = q[head++];
        const int isize = it.isize / 2;
        const float cs = it.size * 0.5f;
        for (int i = 0; i < 8; i++) {
            int cxyz[3];
            for (int a = 0; a < 3; a++) cxyz[a] = it.xyz[a] + MCD[i][a] * isize;
            int on = 0;
            for (int a = 0; a < 3; a++) on |= cxyz[a] == 0 || cxyz[a] == dim - isize;
            int child = -1;
            if (on) {
                child = new_node(S);
                snode_t* c = &S->n[child];
                for (int a = 0; a < 3; a++) c->pos[a] = it.pos[a] + (float)MCD[i][a] * cs;
                c->size = cs;
                c->leaf = 1;
                if (isize == 1) c->sample = or_sample_value_f32(s, c->pos[0], c->pos[1], c->pos[2]);
                if (isize > 1) {
                    c->leaf = 0;
                    if (tail == qcap) {
                        qcap *= 2;
                        q = (item_t*)realloc(q, sizeof(item_t) * (size_t)qcap);
                    }
                    item_t nx = {child, {cxyz[0], cxyz[1], cxyz[2]}, isize, {c->pos[0], c->pos[1], c->pos[2]}, cs};
                    q[tail++] = nx;
                }
            }
            if (it.node < 0) kid[i] = child;
            else S->n[it.node].kid[i] = child;
        }
    }
    free(q);
}

int64_t or_stitch_recursive(const or_sampler_t* s, const or_world_node_t* nodes, int n_nodes, int dim, float* out,
                            int64_t cap)
{
    stitcher_t S = {0};
    S.out = out;
    S.cap_t = cap;
    for (int i = 0; i < n_nodes; i++) {
        const int k = new_node(&S);
        snode_t* q = &S.n[k];
        memcpy(q->pos, nodes[i].pos, sizeof(q->pos));
        q->size = nodes[i].size;
        q->world = 1;
        q->world_leaf = nodes[i].world_leaf;
        q->leaf = 0;
        for (int c = 0; c < 8; c++) q->kid[c] = nodes[i].world_leaf ? -1 : nodes[i].children[c];
    }
    for (int i = 0; i < n_nodes; i++) {
        if (!nodes[i].world_leaf) continue;
        int kid[8];
        chunk_octree(&S, s, nodes[i].pos, nodes[i].size, dim, kid);
        memcpy(S.n[i].kid, kid, sizeof(kid));
    }
    /* gather_all_cells from the root, depth first in child order; every
     * inner node is stitched (the stitch flags are ignored: RESTITCH_ALL) */
    int root = -1;
    for (int i = 0; i < n_nodes && root < 0; i++)
        if (nodes[i].parent < 0) root = i;
    int* stack = (int*)malloc(sizeof(int) * (size_t)(n_nodes + 8));
    int sp = 0;
    if (root >= 0) stack[sp++] = root;
    int* order = (int*)malloc(sizeof(int) * (size_t)(n_nodes + 1));
    int no = 0;
    while (sp) {
        const int n = stack[--sp];
        if (nodes[n].world_leaf) continue;
        order[no++] = n;
        for (int c = 7; c >= 0; c--)
            if (nodes[n].children[c] >= 0) stack[sp++] = nodes[n].children[c];
    }
    for (int k = 0; k < no; k++) cell(&S, order[k]);
    free(stack);
    free(order);
    free(S.n);
    return S.nt;
}

/* ===================== the point-by-point formulation ==================== */
/* Integer geometry in units of the finest unit node: a node at level l with
 * morton code 1 d1 d2 .. dl (3-bit digits dx | dy << 1 | dz << 2, the
 * children of WorldOctree::split_node) has origin sum_k d_k 2^(L - k) dim
 * and extent 2^(L - l) dim, L the deepest level. */
typedef struct {
    const or_world_node_t* nodes;
    int n_nodes, root, dim, L;
} world_t;

static void node_box(const world_t* W, int i, int64_t o[3], int64_t* ext)
{
    const int l = W->nodes[i].level;
    const uint64_t m = W->nodes[i].morton;
    o[0] = o[1] = o[2] = 0;
    for (int k = 1; k <= l; k++) {
        const int d = (int)((m >> (3 * (l - k))) & 7u);
        const int64_t w = ((int64_t)1 << (W->L - k)) * W->dim;
        o[0] += (d & 1) * w;
        o[1] += ((d >> 1) & 1) * w;
        o[2] += ((d >> 2) & 1) * w;
    }
    *ext = ((int64_t)1 << (W->L - l)) * W->dim;
}

/* The world leaf containing the point 2p + s (half units; s = +-1 per axis),
 * or -1 outside the world. */
static int locate(const world_t* W, const int64_t p2[3])
{
    int n = W->root;
    for (;;) {
        int64_t o[3], e;
        node_box(W, n, o, &e);
        for (int a = 0; a < 3; a++)
            if (p2[a] < 2 * o[a] || p2[a] >= 2 * (o[a] + e)) return -1;
        if (W->nodes[n].world_leaf) return n;
        int d[3];
        for (int a = 0; a < 3; a++) d[a] = p2[a] >= 2 * o[a] + e;
        int next = -1;
        for (int c = 0; c < 8; c++)
            if (MCD[c][0] == d[0] && MCD[c][1] == d[1] && MCD[c][2] == d[2]) next = W->nodes[n].children[c];
        if (next < 0) return -1;
        n = next;
    }
}

/* Position of the unit node at lattice index q of chunk n: the shell
 * octree's accumulation pos + MCD * size / 2 + MCD * size / 4 + ... */
static void unit_pos(const world_t* W, int n, const int64_t q[3], float out[3])
{
    float p[3] = {W->nodes[n].pos[0], W->nodes[n].pos[1], W->nodes[n].pos[2]};
    float size = W->nodes[n].size;
    for (int half = W->dim / 2; half >= 1; half /= 2) {
        const float cs = size * 0.5f;
        for (int a = 0; a < 3; a++) p[a] = p[a] + (float)((q[a] / half) & 1) * cs;
        size = cs;
    }
    memcpy(out, p, sizeof(p));
}

/* The triangles of the points stitched by chunk L (written when out is not
 * NULL, at most cap). */
static int64_t stitch_chunk(const world_t* W, const or_sampler_t* s, int L, float* out, int64_t cap)
{
    const int dim = W->dim;
    int64_t o[3], e, nt = 0;
    node_box(W, L, o, &e);
    const int64_t u = e / dim;
    float scratch[5 * 9];
    for (int i = 0; i <= dim; i++)
        for (int j = 0; j <= dim; j++)
            for (int k = 0; k <= dim; k++) {
                if (!(i == 0 || j == 0 || k == 0 || i == dim || j == dim || k == dim)) continue;
                const int64_t P[3] = {o[0] + i * u, o[1] + j * u, o[2] + k * u};
                int oc[8];
                int same = 1, owner = -1, inside = 1;
                int64_t owner_ext = 0;
                for (int c = 0; c < 8 && inside; c++) {
                    const int64_t p2[3] = {2 * P[0] + (c & 4 ? 1 : -1), 2 * P[1] + (c & 2 ? 1 : -1),
                                           2 * P[2] + (c & 1 ? 1 : -1)};
                    oc[c] = locate(W, p2);
                    if (oc[c] < 0) {
                        inside = 0;
                        break;
                    }
                    if (oc[c] != oc[0]) same = 0;
                    int64_t oo[3], ee;
                    node_box(W, oc[c], oo, &ee);
                    if (owner < 0 || ee < owner_ext || (ee == owner_ext && oc[c] < owner)) {
                        owner = oc[c];
                        owner_ext = ee;
                    }
                }
                if (!inside || same || owner != L) continue;
                float pos[8][3], smp[8];
                for (int c = 0; c < 8; c++) {
                    int64_t oo[3], ee;
                    node_box(W, oc[c], oo, &ee);
                    const int64_t uc = ee / dim;
                    int64_t q[3];
                    const int d[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
                    for (int a = 0; a < 3; a++) {
                        /* the unit node containing P + eps in direction d[a] */
                        const int64_t rel = P[a] - oo[a];
                        q[a] = d[a] ? rel / uc : (rel + uc - 1) / uc - 1;
                    }
                    unit_pos(W, oc[c], q, pos[c]);
                    smp[c] = or_sample_value_f32(s, pos[c][0], pos[c][1], pos[c][2]);
                }
                const int64_t got = mc_cell(pos, smp, scratch, 0, 5);
                if (out)
                    for (int64_t t = 0; t < got && nt + t < cap; t++) memcpy(out + 9 * (nt + t), scratch + 9 * t, 36);
                nt += got;
            }
    return nt;
}

int64_t or_stitch_points(const or_sampler_t* s, const or_world_node_t* nodes, int n_nodes, const int32_t* leaves,
                         int n_leaves, int dim, float* out, int64_t cap)
{
    world_t W = {nodes, n_nodes, -1, dim, 0};
    for (int i = 0; i < n_nodes; i++) {
        if (nodes[i].parent < 0 && W.root < 0) W.root = i;
        if (nodes[i].level > W.L) W.L = nodes[i].level;
    }
    if (W.root < 0 || n_leaves == 0) return 0;
    /* chunks in parallel: count, then write at the chunks' offsets (in order) */
    int64_t* off = (int64_t*)calloc((size_t)n_leaves + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic)
    for (int li = 0; li < n_leaves; li++) off[li + 1] = stitch_chunk(&W, s, leaves[li], NULL, 0);
    for (int li = 0; li < n_leaves; li++) off[li + 1] += off[li];
#pragma omp parallel for schedule(dynamic)
    for (int li = 0; li < n_leaves; li++)
        if (off[li] < cap) stitch_chunk(&W, s, leaves[li], out + 9 * off[li], cap - off[li]);
    const int64_t nt = off[n_leaves];
    free(off);
    return nt;
}
