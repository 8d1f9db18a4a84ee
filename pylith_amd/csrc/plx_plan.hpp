// plx_plan.hpp -- host-side tile plan (see plx_plan.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace plx {

// Per-tile bounds that size the LDS of the cell kernels.
struct TileLimits {
    int cells;   // cells per tile = threads per workgroup
    int nodes;   // max distinct nodes per tile
    int verts;   // max distinct vertices per tile
};

inline TileLimits tile_limits(int cell_type) {
    if (cell_type == 1) return {512, 768, 768};  // Q1 hex: an 8^3 brick touches 729 nodes
    return {384, 768, 192};                      // P2 tet: 4^3 cubes x 6 touch 729 nodes
}

// Read-only view of the mesh arrays a plan is built from.
struct MeshView {
    int nb, nv;
    int num_cells, num_nodes, num_vertices;
    const int32_t *closure;
    const int32_t *cone;
    const int32_t *labels;  // may be null (all 0)
    bool cone_is_closure;   // isoparametric Q1: vertices are the nodes, in closure order
    const double *coords = nullptr;  // [num_vertices][3] (spatial reordering of badly ordered keys)
};

struct HostPlan {
    int num_tiles = 0;
    int max_nodes = 0, max_verts = 0, max_colors = 0;
    int64_t halo_slots = 0;
    std::vector<int32_t> cells;  // key cells in tile order
    std::vector<int32_t> tile_cell_ptr, tile_node_ptr, tile_nown, tile_nodes, tile_vert_ptr,
        tile_verts, tile_halo_off, tile_ncolor;
    std::vector<uint16_t> lclosure, lcone;  // [nb][ncells], [nv][ncells]
    std::vector<uint8_t> color;             // [ncells]
    // With cone_is_closure the vertex list of a tile IS its node list (same order), so the
    // kernels index staged coordinates with the local node index.
    std::vector<int32_t> shared_node, shared_ptr, shared_slot;
    bool reordered = false;  // cells taken in Morton order of their centroids instead of input order
};

// sort_by_color: order each tile's cells by colour (colour-phase accumulation activates
// contiguous threads); otherwise cells keep their mesh order inside the tile (neighbouring
// threads then touch neighbouring nodes: structured LDS access for atomic accumulation).
// When the key's cells in input order tile badly (more than 1.5 halo slots per cell, e.g. a mesh
// numbered row by row), the cells are also tiled in Morton order of their centroids and the plan
// with fewer halo slots is kept.
// Permutation (new position -> point) listing n points [n][3] in Morton (Z-curve) order.
std::vector<int32_t> morton_permutation(const double *xyz, size_t n);

int build_host_plan(const MeshView &m, int label_value, const TileLimits &lim, HostPlan *hp,
                    std::string *err, bool sort_by_color = true);

}  // namespace plx
