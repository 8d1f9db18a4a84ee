// plx_assembly.hip -- gfx950 kernels of the residual / Jacobian-action apply.
//
// One workgroup per tile of cells, one cell per thread:
//   1. the tile's nodal field and vertex coordinates are staged in LDS through the tile's
//      node / vertex lists (coalesced where node numbers run contiguously);
//   2. each thread evaluates its element vector (Q1 hex: sum-factorized 2x2x2 Gauss;
//      P2 tet: 14-point rule) calling the device twin of the pointwise kernel per
//      quadrature point;
//   3. element vectors are accumulated into an LDS node buffer, colour by colour, so the
//      summation order is fixed;
//   4. nodes owned by the tile are written straight to the output vector; shared nodes
//      leave as partial sums in the halo buffer, reduced by a node-ordered second kernel.
// The result is deterministic: no global atomics, fixed order everywhere.
#include "plx_element.hpp"
#include "plx_internal.hpp"

namespace plx {

// Reference tables of the P2 element in constant memory.
__constant__ P2Tables c_p2;
static bool g_tables_ready = false;

struct DevTab {
    __device__ static double w(int q) { return c_p2.w[q]; }
    __device__ static double dphi(int q, int b, int k) { return c_p2.dphi[q][b][k]; }
};

static int ensure_tables(plx_dm dm) {
    if (g_tables_ready) return 0;
    P2Tables t;
    p2_fill_tables(&t);
    PLX_HIP(dm, hipMemcpyToSymbol(HIP_SYMBOL(c_p2), &t, sizeof(t)));
    g_tables_ready = true;
    return 0;
}

// --------------------------------------------------------------------------------------------
// Tile kernel.
// --------------------------------------------------------------------------------------------
struct ApplyArgs {
    TilePlanDev p;
    const uint8_t *color;   // [stride] colour of each key cell (tile-local greedy colouring)
    const int32_t *tile_ncolor;
    const void *mat;        // [2][stride] lambda, mu
    const void *u;
    void *r;
    void *halo;
    int mode;
};

// Map workgroup i to a tile so that each of the 8 XCDs (round-robin dispatch) walks a
// contiguous range of tiles: neighbouring tiles then share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int i, int n) {
    const int per = n >> 3, rem = n & 7;
    const int x = i & 7, j = i >> 3;
    return x * per + min(x, rem) + j;
}

template <typename Real, int CT, int NTHREADS>
__global__ void __launch_bounds__(NTHREADS) apply_tiles(ApplyArgs A) {
    constexpr int NB = CT == PLX_HEX_Q1 ? 8 : 10;
    constexpr int NV = CT == PLX_HEX_Q1 ? 8 : 4;
    extern __shared__ __align__(16) unsigned char smem[];
    const TilePlanDev &P = A.p;
    const int tile = xcd_tile(blockIdx.x, P.num_tiles);
    const int c0 = P.tile_cell_ptr[tile], nc = P.tile_cell_ptr[tile + 1] - c0;
    const int n0 = P.tile_node_ptr[tile], nn = P.tile_node_ptr[tile + 1] - n0;
    const int v0 = P.tile_vert_ptr[tile], nvt = P.tile_vert_ptr[tile + 1] - v0;
    const int nown = P.tile_nown[tile];
    const int tid = threadIdx.x;

    double *sX = reinterpret_cast<double *>(smem);
    Real *sU = reinterpret_cast<Real *>(sX + 3 * nvt);
    Real *sR = sU + 3 * nn;
    const Real *u = static_cast<const Real *>(A.u);
    Real *r = static_cast<Real *>(A.r);

    for (int k = tid; k < nn; k += NTHREADS) {
        const int g = P.tile_nodes[n0 + k];
        sU[3 * k + 0] = u[3 * (size_t)g + 0];
        sU[3 * k + 1] = u[3 * (size_t)g + 1];
        sU[3 * k + 2] = u[3 * (size_t)g + 2];
        sR[3 * k + 0] = Real(0);
        sR[3 * k + 1] = Real(0);
        sR[3 * k + 2] = Real(0);
    }
    for (int k = tid; k < nvt; k += NTHREADS) {
        const int g = P.tile_verts[v0 + k];
        sX[3 * k + 0] = P.coords[3 * (size_t)g + 0];
        sX[3 * k + 1] = P.coords[3 * (size_t)g + 1];
        sX[3 * k + 2] = P.coords[3 * (size_t)g + 2];
    }
    __syncthreads();

    const bool active = tid < nc;
    const int pos = c0 + tid;
    int lc[NB];
    Real re[3][NB];
    int color = 0;
    if (active) {
        const Real *mat = static_cast<const Real *>(A.mat);
        const Real lam = mat[pos], mu = mat[P.stride + pos];
        color = A.color[pos];
#pragma unroll
        for (int b = 0; b < NB; ++b) lc[b] = P.lclosure[(size_t)b * P.stride + pos];
        if constexpr (CT == PLX_HEX_Q1) {
            double X[3][8];
            Real U[3][8];
#pragma unroll
            for (int l = 0; l < 8; ++l) {
                const int v = lc[q1_lex(l)];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    X[d][l] = sX[3 * v + d];
                    U[d][l] = sU[3 * v + d];
                }
            }
            Real rl[3][8];
            q1_element<Real>(X, U, lam, mu, rl);
#pragma unroll
            for (int l = 0; l < 8; ++l)
#pragma unroll
                for (int c = 0; c < 3; ++c) re[c][q1_lex(l)] = rl[c][l];
        } else {
            double X[4][3];
            Real U[3][10];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int lv = P.lcone[(size_t)v * P.stride + pos];
#pragma unroll
                for (int d = 0; d < 3; ++d) X[v][d] = sX[3 * lv + d];
            }
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int c = 0; c < 3; ++c) U[c][b] = sU[3 * lc[b] + c];
            p2_element<Real, DevTab>(X, U, lam, mu, re);
        }
    }
    // Colour-ordered accumulation: cells of one colour share no node.
    const int ncolor = A.tile_ncolor[tile];
    for (int col = 0; col < ncolor; ++col) {
        __syncthreads();
        if (active && color == col) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int c = 0; c < 3; ++c) sR[3 * lc[b] + c] += re[c][b];
        }
    }
    __syncthreads();

    const bool insert = A.mode == PLX_INSERT_VALUES;
    for (int k = tid; k < nown; k += NTHREADS) {
        const size_t g = (size_t)P.tile_nodes[n0 + k];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const Real v = sR[3 * k + c];
            r[3 * g + c] = insert ? v : r[3 * g + c] + v;
        }
    }
    Real *halo = static_cast<Real *>(A.halo);
    const size_t hoff = (size_t)P.tile_halo_off[tile];
    for (int k = nown + tid; k < nn; k += NTHREADS) {
        const size_t s = hoff + (size_t)(k - nown);
#pragma unroll
        for (int c = 0; c < 3; ++c) halo[3 * s + c] = sR[3 * k + c];
    }
}

// Shared nodes: sum the tiles' partials in tile order.
template <typename Real>
__global__ void __launch_bounds__(256) reduce_halo(int nshared, const int32_t *__restrict__ node,
                                                   const int32_t *__restrict__ ptr,
                                                   const int32_t *__restrict__ slot,
                                                   const Real *__restrict__ halo, Real *r, int mode) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshared) return;
    Real a0 = Real(0), a1 = Real(0), a2 = Real(0);
    for (int j = ptr[s]; j < ptr[s + 1]; ++j) {
        const size_t h = 3 * (size_t)slot[j];
        a0 += halo[h + 0];
        a1 += halo[h + 1];
        a2 += halo[h + 2];
    }
    const size_t g = 3 * (size_t)node[s];
    if (mode == PLX_INSERT_VALUES) {
        r[g + 0] = a0;
        r[g + 1] = a1;
        r[g + 2] = a2;
    } else {
        r[g + 0] += a0;
        r[g + 1] += a1;
        r[g + 2] += a2;
    }
}

static TilePlanDev plan_dev(plx_dm dm, const KeyPlan *kp) {
    TilePlanDev p;
    p.num_tiles = kp->num_tiles;
    p.stride = (int)kp->cells.size();
    p.tile_cell_ptr = kp->d_tile_cell_ptr;
    p.tile_node_ptr = kp->d_tile_node_ptr;
    p.tile_nown = kp->d_tile_nown;
    p.tile_nodes = kp->d_tile_nodes;
    p.tile_vert_ptr = kp->d_tile_vert_ptr;
    p.tile_verts = kp->d_tile_verts;
    p.tile_halo_off = kp->d_tile_halo_off;
    p.lclosure = kp->d_lclosure;
    p.lcone = kp->d_lcone;
    p.coords = dm->d_coords;
    return p;
}

template <typename Real>
static int apply_typed(plx_dm dm, KeyPlan *kp, const void *u, void *r, int mode) {
    ApplyArgs A;
    A.p = plan_dev(dm, kp);
    A.color = kp->d_color;
    A.tile_ncolor = kp->d_tile_ncolor;
    A.mat = std::is_same<Real, double>::value ? (const void *)kp->d_mat64 : (const void *)kp->d_mat32;
    A.u = u;
    A.r = r;
    A.halo = kp->d_halo;
    A.mode = mode;
    const size_t lds = (size_t)kp->max_verts * 3 * sizeof(double) + (size_t)kp->max_nodes * 6 * sizeof(Real);
    {
    ScopedEvent ev(dm, "plx:apply_tiles");
    if (dm->cell_type == PLX_HEX_Q1) {
        hipLaunchKernelGGL((apply_tiles<Real, PLX_HEX_Q1, 512>), dim3(kp->num_tiles), dim3(512), lds,
                           dm->stream, A);
    } else {
        hipLaunchKernelGGL((apply_tiles<Real, PLX_TET_P2, 384>), dim3(kp->num_tiles), dim3(384), lds,
                           dm->stream, A);
    }
    }
    PLX_HIP(dm, hipGetLastError());
    if (kp->num_shared > 0) {
        ScopedEvent ev(dm, "plx:reduce_halo");
        hipLaunchKernelGGL(reduce_halo<Real>, dim3((kp->num_shared + 255) / 256), dim3(256), 0,
                           dm->stream, kp->num_shared, kp->d_shared_node, kp->d_shared_ptr,
                           kp->d_shared_slot, static_cast<const Real *>(kp->d_halo),
                           static_cast<Real *>(r), mode);
        PLX_HIP(dm, hipGetLastError());
    }
    return 0;
}

int launch_apply(plx_dm dm, KeyPlan *kp, int scalar_type, bool jacobian, const void *u, void *r,
                 int mode) {
    (void)jacobian;  // linear isotropic elasticity: J v uses the same moduli as the residual
    int err = ensure_tables(dm);
    if (err) return err;
    if (mode != PLX_ADD_VALUES && mode != PLX_INSERT_VALUES) return set_error(dm, 1, "bad insert mode");
    if (scalar_type == PLX_FP64) return apply_typed<double>(dm, kp, u, r, mode);
    if (scalar_type == PLX_FP32) return apply_typed<float>(dm, kp, u, r, mode);
    return set_error(dm, 1, "bad scalar type");
}

// --------------------------------------------------------------------------------------------
// Point-block diagonal of J (FP64): D_b[f][g] = -sum_q w|J| [(lambda+mu) dphi_f dphi_g
//   + mu delta_fg |grad phi|^2] for isotropic C, i.e. the diagonal blocks of the element
// matrix of Jf3 = -C.  One thread per cell, global FP64 atomics (set-up path, not timed).
// --------------------------------------------------------------------------------------------
template <int CT>
__global__ void __launch_bounds__(256) pbdiag_cells(TilePlanDev P, const int32_t *cells,
                                                    const int32_t *closure, const int32_t *cone,
                                                    const double *mat, int ncells, double *blocks) {
    constexpr int NB = CT == PLX_HEX_Q1 ? 8 : 10;
    constexpr int NV = CT == PLX_HEX_Q1 ? 8 : 4;
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= ncells) return;
    const int cell = cells[pos];
    const double lam = mat[pos], mu = mat[P.stride + pos];
    double X[NV][3];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) X[v][d] = P.coords[3 * (size_t)cone[(size_t)cell * NV + v] + d];
    double D[NB][6];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int i = 0; i < 6; ++i) D[b][i] = 0.0;
    const int nq = CT == PLX_HEX_Q1 ? 8 : kMaxQuad;
    const double g = 0.57735026918962576451;
    const double ref[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    for (int q = 0; q < nq; ++q) {
        double dphi[NB][3], J[9] = {0}, Ji[9], w;
        if constexpr (CT == PLX_HEX_Q1) {
            const double xi[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
            for (int b = 0; b < 8; ++b) {
                const double a0 = 1 + ref[b][0] * xi[0], a1 = 1 + ref[b][1] * xi[1], a2 = 1 + ref[b][2] * xi[2];
                dphi[b][0] = 0.125 * ref[b][0] * a1 * a2;
                dphi[b][1] = 0.125 * a0 * ref[b][1] * a2;
                dphi[b][2] = 0.125 * a0 * a1 * ref[b][2];
            }
            for (int v = 0; v < 8; ++v)
                for (int d = 0; d < 3; ++d)
                    for (int k = 0; k < 3; ++k) J[3 * d + k] += X[v][d] * dphi[v][k];
            w = 1.0;
        } else {
            for (int b = 0; b < 10; ++b)
                for (int k = 0; k < 3; ++k) dphi[b][k] = c_p2.dphi[q][b][k];
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
            w = c_p2.w[q];
        }
        const double wd = w * inv3(J, Ji);
        for (int b = 0; b < NB; ++b) {
            double gp[3];
            for (int d = 0; d < 3; ++d) gp[d] = dphi[b][0] * Ji[d] + dphi[b][1] * Ji[3 + d] + dphi[b][2] * Ji[6 + d];
            const double n2 = gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2];
            const double lm = lam + mu;
            D[b][0] -= wd * (lm * gp[0] * gp[0] + mu * n2);
            D[b][1] -= wd * (lm * gp[1] * gp[1] + mu * n2);
            D[b][2] -= wd * (lm * gp[2] * gp[2] + mu * n2);
            D[b][3] -= wd * lm * gp[0] * gp[1];
            D[b][4] -= wd * lm * gp[1] * gp[2];
            D[b][5] -= wd * lm * gp[0] * gp[2];
        }
    }
    for (int b = 0; b < NB; ++b) {
        double *B = blocks + 9 * (size_t)closure[(size_t)cell * NB + b];
        atomicAdd(&B[0], D[b][0]);
        atomicAdd(&B[4], D[b][1]);
        atomicAdd(&B[8], D[b][2]);
        atomicAdd(&B[1], D[b][3]);
        atomicAdd(&B[3], D[b][3]);
        atomicAdd(&B[5], D[b][4]);
        atomicAdd(&B[7], D[b][4]);
        atomicAdd(&B[2], D[b][5]);
        atomicAdd(&B[6], D[b][5]);
    }
}

int launch_pbdiag(plx_dm dm, KeyPlan *kp, double *blocks) {
    int err = ensure_tables(dm);
    if (err) return err;
    if (!dm->d_closure) {
        PLX_HIP(dm, hipMalloc(&dm->d_closure, dm->closure.size() * sizeof(int32_t)));
        PLX_HIP(dm, hipMemcpy(dm->d_closure, dm->closure.data(), dm->closure.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        PLX_HIP(dm, hipMalloc(&dm->d_cone, dm->cone.size() * sizeof(int32_t)));
        PLX_HIP(dm, hipMemcpy(dm->d_cone, dm->cone.data(), dm->cone.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    const int n = (int)kp->cells.size();
    const TilePlanDev P = plan_dev(dm, kp);
    if (dm->cell_type == PLX_HEX_Q1)
        hipLaunchKernelGGL(pbdiag_cells<PLX_HEX_Q1>, dim3((n + 255) / 256), dim3(256), 0, dm->stream, P,
                           kp->d_cells, dm->d_closure, dm->d_cone, kp->d_mat64, n, blocks);
    else
        hipLaunchKernelGGL(pbdiag_cells<PLX_TET_P2>, dim3((n + 255) / 256), dim3(256), 0, dm->stream, P,
                           kp->d_cells, dm->d_closure, dm->d_cone, kp->d_mat64, n, blocks);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

// --------------------------------------------------------------------------------------------
// Geometry validation: count cells with a non-positive Jacobian determinant at a quadrature
// point (DMPlexComputeCellGeometryFEM rejects them).
// --------------------------------------------------------------------------------------------
template <int CT>
__global__ void __launch_bounds__(256) validate_cells(int ncells, const int32_t *cone,
                                                      const double *coords, int *bad) {
    constexpr int NV = CT == PLX_HEX_Q1 ? 8 : 4;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncells) return;
    double X[NV][3];
    for (int v = 0; v < NV; ++v)
        for (int d = 0; d < 3; ++d) X[v][d] = coords[3 * (size_t)cone[(size_t)cell * NV + v] + d];
    double minDet = 1e300;
    if constexpr (CT == PLX_HEX_Q1) {
        const double g = 0.57735026918962576451;
        const double ref[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        for (int q = 0; q < 8; ++q) {
            const double xi[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
            double J[9] = {0}, Ji[9];
            for (int b = 0; b < 8; ++b) {
                const double a0 = 1 + ref[b][0] * xi[0], a1 = 1 + ref[b][1] * xi[1], a2 = 1 + ref[b][2] * xi[2];
                const double dp[3] = {0.125 * ref[b][0] * a1 * a2, 0.125 * a0 * ref[b][1] * a2,
                                      0.125 * a0 * a1 * ref[b][2]};
                for (int d = 0; d < 3; ++d)
                    for (int k = 0; k < 3; ++k) J[3 * d + k] += X[b][d] * dp[k];
            }
            minDet = fmin(minDet, inv3(J, Ji));
        }
    } else {
        double J[9], Ji[9];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
        minDet = inv3(J, Ji);
    }
    if (!(minDet > 0.0)) atomicAdd(bad, 1);
}

int validate_geometry(plx_dm dm, int *num_bad) {
    int32_t *d_cone = nullptr;
    int *d_bad = nullptr;
    PLX_HIP(dm, hipMalloc(&d_cone, dm->cone.size() * sizeof(int32_t)));
    PLX_HIP(dm, hipMemcpy(d_cone, dm->cone.data(), dm->cone.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    PLX_HIP(dm, hipMalloc(&d_bad, sizeof(int)));
    PLX_HIP(dm, hipMemset(d_bad, 0, sizeof(int)));
    const int n = dm->num_cells;
    if (dm->cell_type == PLX_HEX_Q1)
        hipLaunchKernelGGL(validate_cells<PLX_HEX_Q1>, dim3((n + 255) / 256), dim3(256), 0, 0, n, d_cone, dm->d_coords, d_bad);
    else
        hipLaunchKernelGGL(validate_cells<PLX_TET_P2>, dim3((n + 255) / 256), dim3(256), 0, 0, n, d_cone, dm->d_coords, d_bad);
    PLX_HIP(dm, hipGetLastError());
    PLX_HIP(dm, hipMemcpy(num_bad, d_bad, sizeof(int), hipMemcpyDeviceToHost));
    (void)hipFree(d_cone);
    (void)hipFree(d_bad);
    return 0;
}

}  // namespace plx
