// plx_emulate.cpp -- host emulation of the tile kernels (apply_tiles + reduce_halo).
//
// Runs the same plan, the same LDS staging, the same element math (plx_element.hpp) and the
// same write-out as the gfx950 kernels, on the CPU, with every LDS / global / halo index
// bounds-checked.  The CPU test suite compares it with the oracle; it is how kernel indexing
// is checked before code reaches a GPU.
#include <vector>

#include "../../include/pylith_amd_debug.h"
#include "plx_element.hpp"
#include "plx_plan.hpp"

namespace {

struct HostTab {
    static plx::P2Tables &t() {
        static plx::P2Tables tab = [] {
            plx::P2Tables x;
            plx::p2_fill_tables(&x);
            return x;
        }();
        return tab;
    }
    static double w(int q) { return t().w[q]; }
    static double dphi(int q, int b, int k) { return t().dphi[q][b][k]; }
    static double lam(int q, int v) { return t().lam[q][v]; }
};

#define CHECK(cond)                                                                              \
    do {                                                                                         \
        if (!(cond)) return 3;                                                                   \
    } while (0)

int xcd_tile(int i, int n) {
    const int per = n >> 3, rem = n & 7;
    const int x = i & 7, j = i >> 3;
    return x * per + (x < rem ? x : rem) + j;
}

template <typename Real>
int emulate(const plx::MeshView &mv, const double *coords, const plx::HostPlan &hp,
            const double *lam_mu, const Real *u, Real *r, int mode, int cell_type, bool linear) {
    const int nb = mv.nb, nv = mv.nv;
    const int stride = (int)hp.cells.size();
    const plx::TileLimits lim = plx::tile_limits(cell_type);
    std::vector<Real> halo((size_t)hp.halo_slots * 3, Real(0));
    std::vector<char> tile_seen(hp.num_tiles, 0);
    for (int blk = 0; blk < hp.num_tiles; ++blk) {
        const int tile = xcd_tile(blk, hp.num_tiles);
        CHECK(tile >= 0 && tile < hp.num_tiles && !tile_seen[tile]);
        tile_seen[tile] = 1;
        const int c0 = hp.tile_cell_ptr[tile], nc = hp.tile_cell_ptr[tile + 1] - c0;
        const int n0 = hp.tile_node_ptr[tile], nn = hp.tile_node_ptr[tile + 1] - n0;
        const int v0 = hp.tile_vert_ptr[tile], nvt = hp.tile_vert_ptr[tile + 1] - v0;
        const int nown = hp.tile_nown[tile];
        CHECK(nc > 0 && nc <= lim.cells && nn <= lim.nodes && nvt <= lim.verts && nown <= nn);
        std::vector<double> sX(3 * (size_t)nvt);
        std::vector<Real> sU(3 * (size_t)nn), sR(3 * (size_t)nn, Real(0));
        for (int k = 0; k < nn; ++k) {
            const int g = hp.tile_nodes[n0 + k];
            CHECK(g >= 0 && g < mv.num_nodes);
            for (int c = 0; c < 3; ++c) sU[3 * k + c] = u[3 * (size_t)g + c];
        }
        for (int k = 0; k < nvt; ++k) {
            const int g = hp.tile_verts[v0 + k];
            CHECK(g >= 0 && g < mv.num_vertices);
            for (int d = 0; d < 3; ++d) sX[3 * k + d] = coords[3 * (size_t)g + d];
        }
        std::vector<Real> re((size_t)nc * 3 * nb);
        std::vector<int> lcs((size_t)nc * nb);
        for (int tid = 0; tid < nc; ++tid) {
            const int pos = c0 + tid;
            CHECK(pos < stride);
            const int cell = hp.cells[pos];
            const Real lam = (Real)lam_mu[2 * (size_t)cell], mu = (Real)lam_mu[2 * (size_t)cell + 1];
            int *lc = &lcs[(size_t)tid * nb];
            for (int b = 0; b < nb; ++b) {
                lc[b] = hp.lclosure[(size_t)b * stride + pos];
                CHECK(lc[b] < nn);
            }
            Real *out = &re[(size_t)tid * 3 * nb];
            if (cell_type == PLX_HEX_Q1) {
                double X[3][8];
                Real U[3][8], rl[3][8];
                for (int l = 0; l < 8; ++l) {
                    const int v = lc[plx::q1_lex(l)];
                    for (int d = 0; d < 3; ++d) {
                        X[d][l] = sX[3 * v + d];
                        U[d][l] = sU[3 * v + d];
                    }
                }
                // the kernels' paths: affine cells factorized (or at the quadrature points), the
                // rest at the quadrature points with the Jacobian at each point
                double TX[3][8];
                for (int d = 0; d < 3; ++d) plx::wht8<double>(X[d], TX[d]);
                const plx::ArrayAccX xacc{X};
                const plx::ArrayAcc3Q1<Real> uacc{U};
                if (plx::q1_affine(TX)) {
                    if (linear) {
                        double J[9];
                        plx::q1_affine_jacobian(xacc, J);
                        plx::q1_element_affine<Real>(J, uacc, lam, mu, rl);
                    } else {
                        plx::q1_element_qp<Real, true>(xacc, uacc, lam, mu, rl);
                    }
                } else {
                    plx::q1_element_qp<Real, false>(xacc, uacc, lam, mu, rl);
                }
                for (int l = 0; l < 8; ++l)
                    for (int c = 0; c < 3; ++c) out[c * nb + plx::q1_lex(l)] = rl[c][l];
            } else {
                double X[4][3];
                Real U[3][10], rl[3][10];
                for (int v = 0; v < nv; ++v) {
                    const int lv = hp.lcone[(size_t)v * stride + pos];
                    CHECK(lv < nvt);
                    for (int d = 0; d < 3; ++d) X[v][d] = sX[3 * lv + d];
                }
                for (int b = 0; b < nb; ++b)
                    for (int c = 0; c < 3; ++c) U[c][b] = sU[3 * lc[b] + c];
                if (linear)
                    plx::p2_element_linear<Real>(X, plx::ArrayAcc3<Real>{U}, lam, mu, HostTab::t().wdiag, HostTab::t().woff, rl);
                else
                    plx::p2_element<Real, HostTab>(X, U, lam, mu, rl);
                for (int b = 0; b < nb; ++b)
                    for (int c = 0; c < 3; ++c) out[c * nb + b] = rl[c][b];
            }
        }
        // colour phases; within a phase no two cells may touch the same LDS node
        const int ncolor = hp.tile_ncolor[tile];
        std::vector<int> stamp(nn, -1);
        for (int col = 0; col < ncolor; ++col)
            for (int tid = 0; tid < nc; ++tid) {
                if (hp.color[c0 + tid] != col) continue;
                const int *lc = &lcs[(size_t)tid * nb];
                for (int b = 0; b < nb; ++b) {
                    CHECK(stamp[lc[b]] != col);  // a race in the GPU kernel
                    stamp[lc[b]] = col;
                    for (int c = 0; c < 3; ++c) sR[3 * lc[b] + c] += re[(size_t)tid * 3 * nb + c * nb + b];
                }
            }
        for (int k = 0; k < nown; ++k) {
            const size_t g = (size_t)hp.tile_nodes[n0 + k];
            for (int c = 0; c < 3; ++c)
                r[3 * g + c] = mode == PLX_INSERT_VALUES ? sR[3 * k + c] : r[3 * g + c] + sR[3 * k + c];
        }
        const size_t hoff = (size_t)hp.tile_halo_off[tile];
        for (int k = nown; k < nn; ++k) {
            const size_t s = hoff + (size_t)(k - nown);
            CHECK((int64_t)s < hp.halo_slots);
            for (int c = 0; c < 3; ++c) halo[3 * s + c] = sR[3 * k + c];
        }
    }
    const int nshared = (int)hp.shared_node.size();
    for (int s = 0; s < nshared; ++s) {
        Real a[3] = {0, 0, 0};
        for (int j = hp.shared_ptr[s]; j < hp.shared_ptr[s + 1]; ++j) {
            CHECK(hp.shared_slot[j] >= 0 && hp.shared_slot[j] < hp.halo_slots);
            for (int c = 0; c < 3; ++c) a[c] += halo[3 * (size_t)hp.shared_slot[j] + c];
        }
        const size_t g = 3 * (size_t)hp.shared_node[s];
        for (int c = 0; c < 3; ++c) r[g + c] = mode == PLX_INSERT_VALUES ? a[c] : r[g + c] + a[c];
    }
    return 0;
}

}  // namespace

extern "C" int plx_debug_emulate_apply(int cell_type, int num_cells, int num_vertices,
                                       int num_nodes, const int32_t *closure, const int32_t *cone,
                                       const double *coords, const int32_t *labels,
                                       int label_value, const double *lam_mu, int scalar_type,
                                       const void *u, void *r, int mode, int algorithm) {
    if ((cell_type != PLX_HEX_Q1 && cell_type != PLX_TET_P2) || !closure || !cone || !coords ||
        !lam_mu || !u || !r)
        return 1;
    const int nb = cell_type == PLX_HEX_Q1 ? 8 : 10;
    if (cell_type == PLX_HEX_Q1)
        for (size_t i = 0; i < (size_t)num_cells * 8; ++i)
            if (cone[i] != closure[i]) return 1;
    plx::MeshView mv{nb, cell_type == PLX_HEX_Q1 ? 8 : 4, num_cells, num_nodes, num_vertices,
                     closure, cone, labels, cell_type == PLX_HEX_Q1, coords};
    plx::HostPlan hp;
    std::string msg;
    if (plx::build_host_plan(mv, label_value, plx::tile_limits(cell_type), &hp, &msg)) return 1;
    if (scalar_type == PLX_FP64)
        return emulate<double>(mv, coords, hp, lam_mu, (const double *)u, (double *)r, mode, cell_type,
                               algorithm != PLX_ALGO_QUADRATURE_POINTS);
    if (scalar_type == PLX_FP32)
        return emulate<float>(mv, coords, hp, lam_mu, (const float *)u, (float *)r, mode, cell_type,
                              algorithm != PLX_ALGO_QUADRATURE_POINTS);
    return 1;
}
