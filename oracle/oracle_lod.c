/*
 * oracle_lod.c -- CPU restatement of the watcher's LOD pass (test
 * infrastructure only; never linked into the product).
 *
 * WorldWatcher::check_leaves (WorldWatcher.cpp:155-173) walks the renderables
 * list while counter < max_gen: a node that WorldOctree::node_needs_split
 * (WorldOctree.cpp:239-249) adds 8 to the counter and goes through
 * handle_split_check (WorldWatcher.cpp:175-188); otherwise a world leaf whose
 * parent node_needs_group (WorldOctree.cpp:251-261) sends the parent through
 * handle_group_check (WorldWatcher.cpp:189-212).  process_batch (:214-238)
 * turns each pushed node into the generate batch by its flags: DESTROY none,
 * SPLIT the 8 children of WorldOctree::split_node (WorldOctree.cpp:175-210),
 * GROUP the node itself (group_node_1, WorldWatcher.cpp:383-392).  Each chunk
 * gets dim = chunk_resolution (WorldOctree::create_chunk, :263-269) and the
 * overlap of ChunkGenerator.cpp:98.  Sequential, in the reference's order,
 * with the flags updated in place as the reference does.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static const uint32_t MCDX[8] = {0, 1, 1, 0, 0, 1, 1, 0}; /* Tables.hpp:7-9 */
static const uint32_t MCDY[8] = {0, 0, 0, 0, 1, 1, 1, 1};
static const uint32_t MCDZ[8] = {0, 0, 1, 1, 0, 0, 1, 1};

enum { F_DRAW = 1, F_DESTROY = 2, F_SPLIT = 4, F_GROUP = 8, F_DRAW_CHILDREN = 64, F_SUPERCEDED = 256 };

/* glm::distance(middle, focus), middle = pos + size * 0.5 */
static float node_distance(const or_world_node_t* n, const or_lod_params_t* P)
{
    const float h = n->size * 0.5f;
    const float m[3] = {n->pos[0] + h, n->pos[1] + h, n->pos[2] + h};
    const float d[3] = {P->focus[0] - m[0], P->focus[1] - m[1], P->focus[2] - m[2]};
    return sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
}

int or_node_needs_split(const or_world_node_t* n, const or_lod_params_t* P)
{
    if (n->level >= P->max_level) return 0;
    if (n->level < P->min_level) return 1;
    const float lim = (n->size * P->split_multiplier + P->size_modifier) + n->size * 0.5f;
    return node_distance(n, P) < lim;
}

int or_node_needs_group(const or_world_node_t* n, const or_lod_params_t* P)
{
    if (n->level < P->min_level) return 0;
    if (n->level > P->max_level) return 1;
    const float lim = (n->size * P->group_multiplier + P->size_modifier) + n->size * 0.5f;
    return node_distance(n, P) > lim;
}

int or_lod_queue(const or_world_node_t* nodes, int n_nodes, const int32_t* rlist, int n_r, const or_lod_params_t* P,
                 uint32_t* flags, int32_t* dirty, or_lod_item_t* items, or_chunk_t* descs, int cap, int* n_dirty,
                 int* n_chunks)
{
    for (int i = 0; i < n_nodes; i++) flags[i] = nodes[i].flags;
    /* check_leaves */
    int nd = 0, counter = 0;
    for (int r = 0; r < n_r && counter < P->max_gen; r++) {
        const int id = rlist[r];
        const or_world_node_t* n = &nodes[id];
        if (or_node_needs_split(n, P)) {
            const uint32_t f = flags[id];
            if (n->world_leaf && !(f & F_SPLIT) && !(f & F_DRAW_CHILDREN) && (f & F_DRAW) && !(f & F_GROUP) &&
                !(f & F_SUPERCEDED)) {
                flags[id] |= F_SPLIT;
                if (flags[id] & F_GROUP) flags[id] ^= F_GROUP;
                dirty[nd++] = id;
            }
            counter += 8;
        } else if (n->world_leaf && n->parent >= 0 && or_node_needs_group(&nodes[n->parent], P)) {
            const int p = n->parent;
            const uint32_t f = flags[p];
            if (!nodes[p].world_leaf && !(f & F_GROUP) && (f & F_SUPERCEDED)) {
                int ok = 1;
                for (int i = 0; i < 8; i++) {
                    const int c = nodes[p].children[i];
                    if (c < 0 || !nodes[c].world_leaf) {
                        ok = 0;
                        break;
                    }
                }
                if (ok) {
                    flags[p] |= F_GROUP;
                    dirty[nd++] = p;
                }
            }
        }
    }
    /* process_batch -> the generate batch */
    int nc = 0;
    for (int k = 0; k < nd; k++) {
        const int id = dirty[k];
        const or_world_node_t* q = &nodes[id];
        const uint32_t f = flags[id];
        if (f & F_DESTROY) continue;
        const int count = (f & F_SPLIT) ? 8 : ((f & F_GROUP) ? 1 : 0);
        for (int i = 0; i < count; i++, nc++) {
            if (nc >= cap) continue;
            or_chunk_t* d = &descs[nc];
            or_lod_item_t* it = &items[nc];
            if (count == 8) {
                const float cs = q->size * 0.5f;
                d->pos[0] = q->pos[0] + (float)MCDX[i] * cs;
                d->pos[1] = q->pos[1] + (float)MCDY[i] * cs;
                d->pos[2] = q->pos[2] + (float)MCDZ[i] * cs;
                d->size = q->size * 0.5f;
                it->level = q->level + 1;
                it->child = i;
                it->action = 1;
                it->morton = (q->morton << 3) | (uint64_t)(MCDX[i] | (MCDY[i] << 1) | (MCDZ[i] << 2));
            } else {
                memcpy(d->pos, q->pos, sizeof(d->pos));
                d->size = q->size;
                it->level = q->level;
                it->child = -1;
                it->action = 2;
                it->morton = q->morton;
            }
            it->node = id;
            d->dim = P->chunk_resolution;
            d->iters = P->process_iters;
            d->boundary_processing = P->boundary_processing;
            d->overlap = (it->level == P->max_level && (!P->boundary_processing || P->process_iters == 0))
                             ? 0.0f
                             : P->overlap + 0.005f * (float)P->process_iters;
        }
    }
    *n_dirty = nd;
    *n_chunks = nc;
    return nc > cap ? 2 : 0;
}
