// plx_plan.cpp -- tile plan of one weak-form key, built on the host (no device calls).
//
// The key's cells (PyLith's DSLabelAccess pointsIS: cells with the key's label value,
// feassemble/DSLabelAccess.icc:53-84) are cut, in cell order, into tiles bounded by the
// workgroup size and the LDS capacity.  Per tile: greedy cell colouring (cells sharing a node
// get different colours; cells sorted by colour), the node list with owned nodes first (numbered
// for LDS bank balance), 16-bit local closures, and halo slots for nodes other tiles also touch.
#include <algorithm>
#include <array>
#include <cmath>

#include "plx_plan.hpp"

namespace plx {

// Morton (Z-order) permutation of n points on a 2^21 grid over their bounding box (a finer grid
// than the mesh spacing, so points sharing a grid point are rare; ties keep input order).
std::vector<int32_t> morton_permutation(const double *xyz, size_t n) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], xyz[3 * i + d]);
            hi[d] = std::max(hi[d], xyz[3 * i + d]);
        }
    const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    auto spread = [](uint64_t x) {  // 21 bits -> every third bit
        x &= 0x1fffff;
        x = (x | x << 32) & 0x1f00000000ffffull;
        x = (x | x << 16) & 0x1f0000ff0000ffull;
        x = (x | x << 8) & 0x100f00f00f00f00full;
        x = (x | x << 4) & 0x10c30c30c30c30c3ull;
        x = (x | x << 2) & 0x1249249249249249ull;
        return x;
    };
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t code = 0;
        for (int d = 0; d < 3; ++d) {
            const double t = (xyz[3 * i + d] - lo[d]) / span;
            code |= spread((uint64_t)std::min(2097151.0, std::max(0.0, t * 2097151.0))) << d;
        }
        key[i] = {code, (int32_t)i};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, int32_t> &a, const std::pair<uint64_t, int32_t> &b) {
                         return a.first < b.first;
                     });
    std::vector<int32_t> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = key[i].second;
    return out;
}

// The cells in Morton order of their centroids.
static std::vector<int32_t> morton_order(const MeshView &m, const std::vector<int32_t> &cells) {
    const size_t n = cells.size();
    std::vector<double> cen(3 * n);
    for (size_t i = 0; i < n; ++i) {
        const int32_t *cn = m.cone + (size_t)cells[i] * m.nv;
        for (int d = 0; d < 3; ++d) {
            double acc = 0.0;
            for (int v = 0; v < m.nv; ++v) acc += m.coords[3 * (size_t)cn[v] + d];
            cen[3 * i + d] = acc / m.nv;
        }
    }
    std::vector<int32_t> rank = morton_permutation(cen.data(), n), out(n);
    for (size_t i = 0; i < n; ++i) out[i] = cells[rank[i]];
    return out;
}

static int plan_cells(const MeshView &m, const TileLimits &lim, HostPlan *hp, std::string *err,
                      bool sort_by_color);

// LDS bank balance of a tile's local numbering.  The cell kernels address the staged arrays as
// [3 k + c] doubles, so an 8-byte access to local index k falls on bank pair (3 k + c) mod 16 of
// the 32 four-byte banks, which k mod 16 fixes.  A half-wave (32 consecutive cells of the tile)
// accesses slot b of its 32 cells in one instruction and takes as many cycles as its most loaded
// pair.  `ids` are the tile's items (nodes or vertices), `ranges` their position ranges (owned,
// then shared: an item keeps its range), and slot[b][i] the tile-local item of slot b of cell i.
// Items get residues mod 16 that spread every such access, greedily in order of decreasing use
// (each takes the residue its range still has room for that adds least to the summed squared
// loads of its accesses).  On return `ids` is reordered so that position p holds an item of
// residue p mod 16.
constexpr int kFixShared = 1;
static void balance_banks(std::vector<int32_t> &ids, const std::vector<int> &ranges,
                          const std::vector<std::vector<int32_t>> &slot, int ncell, int nfixed = 0) {
    const int n = (int)ids.size();
    if (n <= 16) return;
    // accesses: (half-wave, slot) -> its distinct items
    std::vector<int32_t> acc_ptr(1, 0), acc_items, item_acc_ptr(n + 1, 0), item_acc;
    std::vector<int32_t> seen(n, -1);
    int nacc = 0;
    for (int g = 0; g < ncell; g += 32)
        for (size_t b = 0; b < slot.size(); ++b, ++nacc) {
            for (int i = g; i < std::min(ncell, g + 32); ++i) {
                const int32_t k = slot[b][i];
                if (seen[k] != nacc) {
                    seen[k] = nacc;
                    acc_items.push_back(k);
                    item_acc_ptr[k + 1]++;
                }
            }
            acc_ptr.push_back((int32_t)acc_items.size());
        }
    for (int k = 0; k < n; ++k) item_acc_ptr[k + 1] += item_acc_ptr[k];
    item_acc.resize(acc_items.size());
    {
        std::vector<int32_t> fill(item_acc_ptr.begin(), item_acc_ptr.end() - 1);
        for (int a = 0; a < nacc; ++a)
            for (int j = acc_ptr[a]; j < acc_ptr[a + 1]; ++j) item_acc[fill[acc_items[j]]++] = a;
    }
    // residue capacities of each position range, the item's range
    const int nr = (int)ranges.size() - 1;
    std::vector<std::array<int, 16>> cap(nr);
    std::vector<int> range_of(n);
    for (int q = 0; q < nr; ++q) {
        cap[q].fill(0);
        for (int p = ranges[q]; p < ranges[q + 1]; ++p) {
            cap[q][p & 15]++;
            range_of[p] = q;
        }
    }
    std::vector<std::array<int16_t, 16>> load(nacc);
    for (auto &l : load) l.fill(0);
    std::vector<int> res(n, -1), order(n);
    for (int k = 0; k < n; ++k) order[k] = k;
    // the last nfixed ranges keep their positions (residue p mod 16)
    for (int k = 0; k < n; ++k)
        if (range_of[k] >= nr - nfixed) {
            res[k] = k & 15;
            cap[range_of[k]][k & 15]--;
            for (int j = item_acc_ptr[k]; j < item_acc_ptr[k + 1]; ++j) load[item_acc[j]][k & 15]++;
        }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return item_acc_ptr[a + 1] - item_acc_ptr[a] > item_acc_ptr[b + 1] - item_acc_ptr[b];
    });
    for (int k : order) {
        if (res[k] >= 0) continue;
        const int q = range_of[k];
        int best = -1, best_cost = 0;
        for (int r = 0; r < 16; ++r) {
            if (cap[q][r] == 0) continue;
            int cost = 0;
            for (int j = item_acc_ptr[k]; j < item_acc_ptr[k + 1]; ++j) cost += 2 * load[item_acc[j]][r] + 1;
            if (best < 0 || cost < best_cost) {
                best = r;
                best_cost = cost;
            }
        }
        res[k] = best;
        cap[q][best]--;
        for (int j = item_acc_ptr[k]; j < item_acc_ptr[k + 1]; ++j) load[item_acc[j]][best]++;
    }
    // positions: within each range, residue r takes positions p = r (mod 16)
    std::vector<int32_t> out(n);
    for (int q = 0; q < nr; ++q) {
        std::array<std::vector<int>, 16> byres;
        for (int k = 0; k < n; ++k)
            if (range_of[k] == q) byres[res[k]].push_back(k);
        std::array<size_t, 16> next{};
        for (int p = ranges[q]; p < ranges[q + 1]; ++p) out[p] = ids[byres[p & 15][next[p & 15]++]];
    }
    ids.swap(out);
}

int build_host_plan(const MeshView &m, int label_value, const TileLimits &lim, HostPlan *hp,
                    std::string *err, bool sort_by_color) {
    std::vector<int32_t> &cells = hp->cells;
    cells.clear();
    for (int c = 0; c < m.num_cells; ++c)
        if ((m.labels ? m.labels[c] : 0) == label_value) cells.push_back(c);
    if (cells.empty()) {
        *err = "no cells carry the key's label value";
        return 1;
    }
    const std::vector<int32_t> input = cells;
    int e = plan_cells(m, lim, hp, err, sort_by_color);
    hp->reordered = false;
    if (e || !m.coords || (double)hp->halo_slots <= 1.5 * (double)input.size()) return e;
    HostPlan alt;
    alt.cells = morton_order(m, input);
    std::string err2;
    if (plan_cells(m, lim, &alt, &err2, sort_by_color) == 0 && alt.halo_slots < hp->halo_slots) {
        *hp = std::move(alt);
        hp->reordered = true;
    }
    return 0;
}

// The plan of hp->cells taken in their given order.
static int plan_cells(const MeshView &m, const TileLimits &lim, HostPlan *hp, std::string *err,
                      bool sort_by_color) {
    const int nb = m.nb, nv = m.nv;
    std::vector<int32_t> &cells = hp->cells;
    const int ncells = (int)cells.size();

    // Greedy tiling in cell order, bounded by LDS capacity.
    std::vector<int32_t> node_mark(m.num_nodes, -1), vert_mark(m.num_vertices, -1);
    std::vector<int32_t> &tile_cell_ptr = hp->tile_cell_ptr;
    tile_cell_ptr.assign(1, 0);
    {
        int tcells = 0, tnodes = 0, tverts = 0, tile = 0;
        for (int i = 0; i < ncells; ++i) {
            const int32_t *cl = m.closure + (size_t)cells[i] * nb;
            const int32_t *cn = m.cone + (size_t)cells[i] * nv;
            int newn = 0, newv = 0;
            for (int b = 0; b < nb; ++b) newn += node_mark[cl[b]] != tile;
            for (int b = 0; b < nv; ++b) newv += vert_mark[cn[b]] != tile;
            if (tcells == lim.cells || tnodes + newn > lim.nodes || tverts + newv > lim.verts) {
                tile_cell_ptr.push_back(i);
                tcells = tnodes = tverts = 0;
                ++tile;
            }
            for (int b = 0; b < nb; ++b)
                if (node_mark[cl[b]] != tile) {
                    node_mark[cl[b]] = tile;
                    ++tnodes;
                }
            for (int b = 0; b < nv; ++b)
                if (vert_mark[cn[b]] != tile) {
                    vert_mark[cn[b]] = tile;
                    ++tverts;
                }
            ++tcells;
        }
        tile_cell_ptr.push_back(ncells);
    }
    const int ntiles = (int)tile_cell_ptr.size() - 1;
    hp->num_tiles = ntiles;

    // How many tiles touch each node.
    std::vector<int32_t> tile_count(m.num_nodes, 0);
    std::fill(node_mark.begin(), node_mark.end(), -1);
    for (int t = 0; t < ntiles; ++t)
        for (int i = tile_cell_ptr[t]; i < tile_cell_ptr[t + 1]; ++i) {
            const int32_t *cl = m.closure + (size_t)cells[i] * nb;
            for (int b = 0; b < nb; ++b)
                if (node_mark[cl[b]] != t) {
                    node_mark[cl[b]] = t;
                    ++tile_count[cl[b]];
                }
        }

    // Greedy colouring inside each tile, then cells sorted by colour.
    hp->color.assign(ncells, 0);
    hp->tile_ncolor.assign(ntiles, 0);
    hp->max_colors = 0;
    {
        std::vector<uint64_t> used(2 * (size_t)m.num_nodes, 0);
        std::vector<std::pair<int, int32_t>> order;
        for (int t = 0; t < ntiles; ++t) {
            const int b0 = tile_cell_ptr[t], b1 = tile_cell_ptr[t + 1];
            order.clear();
            int nc = 0;
            for (int i = b0; i < b1; ++i) {
                const int32_t *cl = m.closure + (size_t)cells[i] * nb;
                uint64_t m0 = 0, m1 = 0;
                for (int b = 0; b < nb; ++b) {
                    m0 |= used[2 * (size_t)cl[b]];
                    m1 |= used[2 * (size_t)cl[b] + 1];
                }
                int c = 0;
                while (c < 128 && ((c < 64 ? (m0 >> c) : (m1 >> (c - 64))) & 1)) ++c;
                if (c >= 128) {
                    *err = "tile colouring needs more than 128 colours";
                    return 1;
                }
                for (int b = 0; b < nb; ++b)
                    used[2 * (size_t)cl[b] + (c >> 6)] |= (uint64_t)1 << (c & 63);
                order.emplace_back(c, cells[i]);
                nc = std::max(nc, c + 1);
            }
            for (int i = b0; i < b1; ++i) {
                const int32_t *cl = m.closure + (size_t)cells[i] * nb;
                for (int b = 0; b < nb; ++b) used[2 * (size_t)cl[b]] = used[2 * (size_t)cl[b] + 1] = 0;
            }
            if (sort_by_color)
                std::stable_sort(order.begin(), order.end(),
                                 [](const std::pair<int, int32_t> &a, const std::pair<int, int32_t> &b) {
                                     return a.first < b.first;
                                 });
            for (int i = b0; i < b1; ++i) {
                hp->color[i] = (uint8_t)order[i - b0].first;
                cells[i] = order[i - b0].second;
            }
            hp->tile_ncolor[t] = nc;
            hp->max_colors = std::max(hp->max_colors, nc);
        }
    }

    // Per-tile node / vertex lists and local closures.
    hp->tile_node_ptr.assign(1, 0);
    hp->tile_vert_ptr.assign(1, 0);
    hp->tile_nown.clear();
    hp->tile_nodes.clear();
    hp->tile_verts.clear();
    hp->tile_halo_off.clear();
    hp->lclosure.assign((size_t)nb * ncells, 0);
    hp->lcone.assign((size_t)nv * ncells, 0);
    std::vector<int32_t> local_of(m.num_nodes, -1), vlocal_of(m.num_vertices, -1);
    std::vector<std::pair<int32_t, int32_t>> shared_pairs;  // (node, halo slot)
    int64_t halo = 0;
    int max_nodes = 0, max_verts = 0;
    std::vector<int32_t> own, sh, vl;
    for (int t = 0; t < ntiles; ++t) {
        own.clear();
        sh.clear();
        vl.clear();
        for (int i = tile_cell_ptr[t]; i < tile_cell_ptr[t + 1]; ++i) {
            const int32_t *cl = m.closure + (size_t)cells[i] * nb;
            const int32_t *cn = m.cone + (size_t)cells[i] * nv;
            for (int b = 0; b < nb; ++b)
                if (local_of[cl[b]] == -1) {
                    local_of[cl[b]] = 0;
                    (tile_count[cl[b]] == 1 ? own : sh).push_back(cl[b]);
                }
            for (int b = 0; b < nv; ++b)
                if (vlocal_of[cn[b]] == -1) {
                    vlocal_of[cn[b]] = 0;
                    vl.push_back(cn[b]);
                }
        }
        std::sort(own.begin(), own.end());
        std::sort(sh.begin(), sh.end());
        std::sort(vl.begin(), vl.end());
        const int c0 = tile_cell_ptr[t], tc = tile_cell_ptr[t + 1] - c0;
        // local numbering: owned nodes, then shared ones.  Colour-sorted tiles put scattered cells in
        // each half-wave: their owned nodes are numbered for LDS bank balance around the shared
        // ones, which keep node order (the halo reduction reads their partials in node order).
        // Tiles in mesh order keep node order: neighbouring lanes already touch spread-out banks.
        std::vector<int32_t> ids(own);
        ids.insert(ids.end(), sh.begin(), sh.end());
        if (sort_by_color) {
            for (size_t q = 0; q < ids.size(); ++q) local_of[ids[q]] = (int32_t)q;
            std::vector<std::vector<int32_t>> slot(nb, std::vector<int32_t>(tc));
            for (int i = 0; i < tc; ++i)
                for (int b = 0; b < nb; ++b) slot[b][i] = local_of[m.closure[(size_t)cells[c0 + i] * nb + b]];
            balance_banks(ids, {0, (int)own.size(), (int)ids.size()}, slot, tc, kFixShared);
        }
        if (sort_by_color && !m.cone_is_closure && vl.size() > 16) {
            for (size_t q = 0; q < vl.size(); ++q) vlocal_of[vl[q]] = (int32_t)q;
            std::vector<std::vector<int32_t>> slot(nv, std::vector<int32_t>(tc));
            for (int i = 0; i < tc; ++i)
                for (int b = 0; b < nv; ++b) slot[b][i] = vlocal_of[m.cone[(size_t)cells[c0 + i] * nv + b]];
            balance_banks(vl, {0, (int)vl.size()}, slot, tc);
        }
        int k = 0;
        for (int32_t n : ids) {
            if (k >= (int)own.size()) shared_pairs.emplace_back(n, (int32_t)(halo + (k - (int)own.size())));
            local_of[n] = k++;
            hp->tile_nodes.push_back(n);
        }
        hp->tile_nown.push_back((int32_t)own.size());
        hp->tile_halo_off.push_back((int32_t)halo);
        halo += (int64_t)sh.size();
        hp->tile_node_ptr.push_back((int32_t)hp->tile_nodes.size());
        int kv = 0;
        for (int32_t v : vl) {
            vlocal_of[v] = kv++;
            hp->tile_verts.push_back(v);
        }
        hp->tile_vert_ptr.push_back((int32_t)hp->tile_verts.size());
        max_nodes = std::max(max_nodes, k);
        max_verts = std::max(max_verts, kv);
        for (int i = tile_cell_ptr[t]; i < tile_cell_ptr[t + 1]; ++i) {
            const int32_t *cl = m.closure + (size_t)cells[i] * nb;
            const int32_t *cn = m.cone + (size_t)cells[i] * nv;
            for (int b = 0; b < nb; ++b) hp->lclosure[(size_t)b * ncells + i] = (uint16_t)local_of[cl[b]];
            for (int b = 0; b < nv; ++b) hp->lcone[(size_t)b * ncells + i] = (uint16_t)vlocal_of[cn[b]];
        }
        for (int32_t n : own) local_of[n] = -1;
        for (int32_t n : sh) local_of[n] = -1;
        for (int32_t v : vl) vlocal_of[v] = -1;
    }
    if (m.cone_is_closure) {
        hp->tile_verts = hp->tile_nodes;
        hp->tile_vert_ptr = hp->tile_node_ptr;
        hp->lcone = hp->lclosure;
        max_verts = max_nodes;
    }
    if (halo > INT32_MAX) {
        *err = "halo too large";
        return 1;
    }
    hp->halo_slots = halo;
    hp->max_nodes = max_nodes;
    hp->max_verts = max_verts;

    // Shared-node reduction map: nodes ascending, slots in tile order (stable sort).
    std::stable_sort(shared_pairs.begin(), shared_pairs.end(),
                     [](const std::pair<int32_t, int32_t> &a, const std::pair<int32_t, int32_t> &b) {
                         return a.first < b.first;
                     });
    hp->shared_node.clear();
    hp->shared_slot.clear();
    hp->shared_ptr.assign(1, 0);
    for (size_t i = 0; i < shared_pairs.size(); ++i) {
        if (i > 0 && shared_pairs[i].first != shared_pairs[i - 1].first)
            hp->shared_ptr.push_back((int32_t)i);
        if (i == 0 || shared_pairs[i].first != shared_pairs[i - 1].first)
            hp->shared_node.push_back(shared_pairs[i].first);
        hp->shared_slot.push_back(shared_pairs[i].second);
    }
    if (!shared_pairs.empty()) hp->shared_ptr.push_back((int32_t)shared_pairs.size());
    return 0;
}

}  // namespace plx
