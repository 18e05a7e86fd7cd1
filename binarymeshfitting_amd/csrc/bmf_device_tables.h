// Marching-cubes tables in constant memory, one copy per translation unit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bmf_mc_tables.h"

namespace bmf {

// MarchingCubes::tri_table (MCTable.h:27-283), packed: bits 60..63 hold the
// entry count, nibble s the local edge of slot s.
static __constant__ uint64_t c_tri_table[256] = BMF_TRI_TABLE_INIT;

// For each mask and local edge, bit q set when triangle q of the cell (table
// slots 3q..3q+2) uses the edge: the cell's part of a vertex's ring, in
// ascending triangle id.  Derived from the triangle table at compile time.
struct EdgeTriTable {
    uint8_t m[256][12];
};
constexpr uint64_t kTriRows[256] = BMF_TRI_TABLE_INIT;
constexpr EdgeTriTable make_edge_tris() {
    EdgeTriTable t{};
    for (int m = 0; m < 256; m++) {
        const uint64_t row = kTriRows[m];
        const int n = (int)(row >> 60);
        for (int s = 0; s < n; s++) {
            const int e = (int)((row >> (4 * s)) & 15u);
            t.m[m][e] = (uint8_t)(t.m[m][e] | (1u << (s / 3)));
        }
    }
    return t;
}
static __constant__ EdgeTriTable c_edge_tris = make_edge_tris();

// Dual-marching-cubes polygons of each mask (Tables.hpp:96-354, EdgeTable up
// to its end marker): bits 60..63 nibble count, nibbles = local edges with
// 0xF between polygons.
static __constant__ uint64_t c_dmc_table[256] = BMF_DMC_POLY_TABLE_INIT;

// Polygons of one c_dmc_table row (edge lists separated by 15).
__device__ __forceinline__ uint32_t dmc_polygons(uint64_t row) {
    const int n = (int)(row >> 60);
    uint32_t p = 0, any = 0;
    for (int s = 0; s < n; s++) {
        if (((row >> (4 * s)) & 15ull) == 15ull) {
            p++;
            any = 0;
        } else {
            any = 1;
        }
    }
    return p + any;
}

// The four cells around a lattice edge along `axis`, in lexicographic cell
// order, as (dx, dy, dz, local edge) offsets from the edge origin; local
// edges numbered as in DMCChunk::polygonize_cell (DMCChunk.cpp:543-565).
// Visiting them in this order and each cell's table slots in order lists
// the edge's triangles in ascending id, the order in which
// MeshProcessor::init_primitives appends them (MeshProcessor.cpp:114-127).
static __constant__ int8_t c_ring[3][4][4] = {
    {{0, -1, -1, 3}, {0, -1, 0, 2}, {0, 0, -1, 1}, {0, 0, 0, 0}},
    {{-1, 0, -1, 7}, {-1, 0, 0, 6}, {0, 0, -1, 5}, {0, 0, 0, 4}},
    {{-1, -1, 0, 11}, {-1, 0, 0, 10}, {0, -1, 0, 9}, {0, 0, 0, 8}},
};

}  // namespace bmf
