// plx_fault.hip -- cohesive cells of a fault with prescribed slip (FP64).
//
// Device twins of FaultCohesiveKin::f0u_neg / f0u_pos / f0l_slip and of the Jacobian blocks
// Jf0ul_neg = -I, Jf0ul_pos = +I, Jf0lu = [-I, +I] (FaultCohesiveKin.hh:73-224,298-417),
// integrated over each cohesive cell's Q1 face with 2x2 Gauss points as
// DMPlexComputeResidualHybridByKey does for IntegratorInterface (IntegratorInterface.cc:603-722).
// One thread per cohesive cell, one launch per colour: cells of one colour share no node, so the
// scatter needs no atomics and the summation order is fixed.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "plx_internal.hpp"

namespace plx {

struct CohesiveArgs {
    const int32_t *cells;  // cohesive cells of this colour
    int ncells;
    const int32_t *neg, *pos, *lag, *cone;  // [ncoh][4]
    const double *coords;
    const double *slip;    // [num_lag][3]
    int64_t lag_offset;    // Lagrange rows of x
    int64_t lag_base;      // Lagrange rows of f (lag_offset, or 0 for a Lagrange-only array)
    double r1[3], r2[3];   // reference directions (kernel constants)
};

enum { kCohResidual = 0, kCohJacobian = 1, kCohSlipOnly = 2 };

__host__ __device__ inline void unit_cross(double r[3], const double u[3], const double v[3]) {
    r[0] = u[1] * v[2] - u[2] * v[1];
    r[1] = u[2] * v[0] - u[0] * v[2];
    r[2] = u[0] * v[1] - u[1] * v[0];
    const double m = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    r[0] /= m;
    r[1] /= m;
    r[2] /= m;
}

// BoundaryDirections::tangential_directions (BoundaryDirections.hh:50-82), including the sign
// test as written there (refDir2[1]*n[1] counted twice, refDir2[2]*n[2] never).
__host__ __device__ inline void tangents(double t1[3], double t2[3], const double r1[3],
                                         const double r2[3], const double n[3]) {
    double r[3] = {r1[0], r1[1], r1[2]};
    if (fabs(r[0] * n[0] + r[1] * n[1] + r[2] * n[2]) > 0.999) {
        r[0] = r2[0];
        r[1] = r2[1];
        r[2] = r2[2];
    }
    unit_cross(t1, r, n);
    unit_cross(t2, n, t1);
    if (r2[0] * n[0] + r2[1] * n[1] + r2[1] * n[1] < 0.0) {
        t2[0] = -t2[0];
        t2[1] = -t2[1];
        t2[2] = -t2[2];
    }
}

// Face basis / geometry at Gauss point q of a cohesive cell (vertex order (-1,-1),(1,-1),(1,1),(-1,1)).
// Evaluated without FMA contraction, each operation rounded as written: the sign test of the
// fault frame (tangents) is decided by the exact rounding of the normal, e.g. +0 vs -1e-17 for a
// fault plane containing the second reference direction.
__host__ __device__ inline void face_point(const double X[4][3], int q, double psi[4], double *area,
                                           double n[3]) {
#pragma clang fp contract(off)
    const double g = 0.57735026918962576451;
    const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g;
    const double sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
    double dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double a = 1.0 + sx[v] * xi, b = 1.0 + sy[v] * eta;
        psi[v] = 0.25 * a * b;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            dxi[d] += X[v][d] * 0.25 * sx[v] * b;
            deta[d] += X[v][d] * 0.25 * a * sy[v];
        }
    }
    n[0] = dxi[1] * deta[2] - dxi[2] * deta[1];
    n[1] = dxi[2] * deta[0] - dxi[0] * deta[2];
    n[2] = dxi[0] * deta[1] - dxi[1] * deta[0];
    const double m = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    *area = m;
    n[0] /= m;
    n[1] /= m;
    n[2] /= m;
}

// kCohResidual: residual of (u, lambda) with the slip term; kCohJacobian: J v (no slip term);
// kCohSlipOnly: the Lagrange rows of the residual at u = 0, lambda = 0 (- integral psi d).
template <int MODE>
__global__ void __launch_bounds__(256) cohesive_kernel(CohesiveArgs A, const double *__restrict__ x,
                                                       double *f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.ncells) return;
    const int c = A.cells[i];
    int nn[4], pn[4];
    int64_t ln[4], lo[4];
    double X[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        nn[v] = A.neg[4 * c + v];
        pn[v] = A.pos[4 * c + v];
        ln[v] = A.lag_offset + 3 * (int64_t)A.lag[4 * c + v];
        lo[v] = A.lag_base + 3 * (int64_t)A.lag[4 * c + v];
        const int cv = A.cone[4 * c + v];
#pragma unroll
        for (int d = 0; d < 3; ++d) X[v][d] = A.coords[3 * (size_t)cv + d];
    }
    double rn[4][3], rp[4][3], rl[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int k = 0; k < 3; ++k) rn[v][k] = rp[v][k] = rl[v][k] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double psi[4], area, n[3];
        face_point(X, q, psi, &area, n);
        double um[3] = {0, 0, 0}, up[3] = {0, 0, 0}, lam[3] = {0, 0, 0}, s[3] = {0, 0, 0};
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (MODE != kCohSlipOnly) {
                    um[k] += psi[v] * x[3 * (size_t)nn[v] + k];
                    up[k] += psi[v] * x[3 * (size_t)pn[v] + k];
                    lam[k] += psi[v] * x[ln[v] + k];
                }
                if (MODE != kCohJacobian) s[k] += psi[v] * A.slip[3 * (size_t)A.lag[4 * c + v] + k];
            }
        double d[3] = {0, 0, 0};
        if (MODE != kCohJacobian) {
            double t1[3], t2[3];
            tangents(t1, t2, A.r1, A.r2, n);
#pragma unroll
            for (int k = 0; k < 3; ++k) d[k] = n[k] * s[0] + t1[k] * s[1] + t2[k] * s[2];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double w = area * psi[v];
                rn[v][k] += w * (-lam[k]);              // f0u_neg / Jf0ul_neg
                rp[v][k] += w * (+lam[k]);              // f0u_pos / Jf0ul_pos
                rl[v][k] += w * (up[k] - um[k] - d[k]); // f0l_slip / Jf0lu
            }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (MODE != kCohSlipOnly) {
                f[3 * (size_t)nn[v] + k] += rn[v][k];
                f[3 * (size_t)pn[v] + k] += rp[v][k];
            }
            f[lo[v] + k] += rl[v][k];
        }
}

__global__ void __launch_bounds__(256) fill_range(int64_t n, double *a, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

static int cohesive_launches(plx_dm dm, int kind, const double *x, double *f, int mode) {
    FaultData &F = *dm->fault;
    const int64_t lag_base = kind == kCohSlipOnly ? 0 : F.lag_offset;
    if (mode == PLX_INSERT_VALUES) {
        const int64_t n = 3 * (int64_t)F.num_lag;
        hipLaunchKernelGGL(fill_range, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, dm->stream, n,
                           f + lag_base, 0.0);
    }
    CohesiveArgs A;
    A.neg = F.d_neg;
    A.pos = F.d_pos;
    A.lag = F.d_lag;
    A.cone = F.d_cone;
    A.coords = dm->d_coords;
    A.slip = F.d_slip;
    A.lag_offset = F.lag_offset;
    A.lag_base = lag_base;
    for (int k = 0; k < 3; ++k) {
        A.r1[k] = F.ref_dir_1[k];
        A.r2[k] = F.ref_dir_2[k];
    }
    ScopedEvent ev(dm, "plx:cohesive");
    for (size_t col = 0; col + 1 < F.color_ptr.size(); ++col) {
        A.cells = F.d_color_cells + F.color_ptr[col];
        A.ncells = F.color_ptr[col + 1] - F.color_ptr[col];
        if (A.ncells == 0) continue;
        const dim3 g((A.ncells + 255) / 256);
        if (kind == kCohJacobian)
            hipLaunchKernelGGL(cohesive_kernel<kCohJacobian>, g, dim3(256), 0, dm->stream, A, x, f);
        else if (kind == kCohResidual)
            hipLaunchKernelGGL(cohesive_kernel<kCohResidual>, g, dim3(256), 0, dm->stream, A, x, f);
        else
            hipLaunchKernelGGL(cohesive_kernel<kCohSlipOnly>, g, dim3(256), 0, dm->stream, A, x, f);
    }
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

int launch_cohesive(plx_dm dm, bool jacobian, const double *x, double *f, int mode) {
    return cohesive_launches(dm, jacobian ? kCohJacobian : kCohResidual, x, f, mode);
}

// lag_out [num_lag][3] = - integral psi d (the Lagrange rows of the cohesive residual at zero state)
int launch_fault_slip_rhs(plx_dm dm, double *lag_out) {
    return cohesive_launches(dm, kCohSlipOnly, nullptr, lag_out, PLX_INSERT_VALUES);
}

// --------------------------------------------------------------------------------------------
// Split-node maps used by the constraint-eliminating solve.
// --------------------------------------------------------------------------------------------
// Split-node maps of the merged (null-space) representation: op 0 folds (v[neg] += v[pos],
// v[pos] = 0), op 1 unfolds (v[pos] = v[neg]), op 2 clears (v[pos] = 0).
template <typename Real>
__global__ void __launch_bounds__(256) split_node_kernel(int n, const int32_t *__restrict__ negn,
                                                         const int32_t *__restrict__ posn, Real *v, int op) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t a = 3 * (size_t)negn[i], b = 3 * (size_t)posn[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (op == 0) {
            v[a + k] += v[b + k];
            v[b + k] = Real(0);
        } else if (op == 1) {
            v[b + k] = v[a + k];
        } else {
            v[b + k] = Real(0);
        }
    }
}

int launch_fold(plx_dm dm, int st, void *v, int op) {
    FaultData &F = *dm->fault;
    const dim3 g((F.num_lag + 255) / 256);
    if (st == PLX_FP64)
        hipLaunchKernelGGL(split_node_kernel<double>, g, dim3(256), 0, dm->stream, F.num_lag, F.d_neg_node,
                           F.d_pos_node, (double *)v, op);
    else
        hipLaunchKernelGGL(split_node_kernel<float>, g, dim3(256), 0, dm->stream, F.num_lag, F.d_neg_node,
                           F.d_pos_node, (float *)v, op);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

// out[i] = v[positive-side node of Lagrange node i] (a local vector's rows to a Lagrange-space array).
__global__ void __launch_bounds__(256) gather_pos_kernel(int n, const int32_t *__restrict__ posn,
                                                         const double *__restrict__ v, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = v[3 * (size_t)posn[i] + k];
}

// v[pos] = v[neg] + jump
__global__ void __launch_bounds__(256) sync_pos_kernel(int n, const int32_t *__restrict__ negn,
                                                       const int32_t *__restrict__ posn,
                                                       const double *__restrict__ jump, double *v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        v[3 * (size_t)posn[i] + k] = v[3 * (size_t)negn[i] + k] + jump[3 * (size_t)i + k];
}

// y = M x for the three components (x, y: [num_lag][3])
__global__ void __launch_bounds__(256) fault_mass_kernel(int n, const int32_t *__restrict__ ptr,
                                                         const int32_t *__restrict__ col,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ x, double *y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
        const double m = val[e];
        const size_t o = 3 * (size_t)col[e];
        a0 += m * x[o];
        a1 += m * x[o + 1];
        a2 += m * x[o + 2];
    }
    y[3 * (size_t)i] = a0;
    y[3 * (size_t)i + 1] = a1;
    y[3 * (size_t)i + 2] = a2;
}

__global__ void __launch_bounds__(256) fault_jacobi_kernel(int n, const double *__restrict__ dinv,
                                                           const double *__restrict__ r, double *z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    z[i] = dinv[i / 3] * r[i];
}

// jump[i] = R slip[i] for k slip sets (planar fault): slip [k][num_lag][3] -> jump [k][num_lag][3]
__global__ void __launch_bounds__(256) planar_jump_kernel(int64_t n, const double *__restrict__ slip,
                                                          double R0, double R1, double R2, double R3,
                                                          double R4, double R5, double R6, double R7,
                                                          double R8, double *jump) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s0 = slip[3 * i], s1 = slip[3 * i + 1], s2 = slip[3 * i + 2];
    jump[3 * i] = R0 * s0 + R1 * s1 + R2 * s2;
    jump[3 * i + 1] = R3 * s0 + R4 * s1 + R5 * s2;
    jump[3 * i + 2] = R6 * s0 + R7 * s1 + R8 * s2;
}

int launch_planar_jump(plx_dm dm, int64_t count, const double *slip, double *jump) {
    const double *R = dm->fault->R;
    hipLaunchKernelGGL(planar_jump_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, dm->stream, count,
                       slip, R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], R[8], jump);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

int launch_sync_pos(plx_dm dm, const double *jump, double *v) {
    FaultData &F = *dm->fault;
    hipLaunchKernelGGL(sync_pos_kernel, dim3((F.num_lag + 255) / 256), dim3(256), 0, dm->stream, F.num_lag,
                       F.d_neg_node, F.d_pos_node, jump, v);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

int launch_fault_mass(plx_dm dm, const double *x, double *y) {
    FaultData &F = *dm->fault;
    hipLaunchKernelGGL(fault_mass_kernel, dim3((F.num_lag + 255) / 256), dim3(256), 0, dm->stream, F.num_lag,
                       F.d_mass_ptr, F.d_mass_col, F.d_mass_val, x, y);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

int launch_fault_jacobi(plx_dm dm, const double *r, double *z) {
    FaultData &F = *dm->fault;
    hipLaunchKernelGGL(fault_jacobi_kernel, dim3((3 * F.num_lag + 255) / 256), dim3(256), 0, dm->stream,
                       F.num_lag, F.d_mass_dinv, r, z);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

// Scalar fault-face mass matrix on the host (setup): M_ab = sum_q |x_xi x x_eta| psi_a psi_b.
static void assemble_fault_mass(const std::vector<double> &coords, const std::vector<int32_t> &lag,
                                const std::vector<int32_t> &cone, int num_lag, std::vector<int32_t> &ptr,
                                std::vector<int32_t> &col, std::vector<double> &val,
                                std::vector<double> &dinv) {
    std::vector<std::map<int32_t, double>> rows(num_lag);
    const size_t ncoh = lag.size() / 4;
    const double g = 0.57735026918962576451;
    const double sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
    for (size_t c = 0; c < ncoh; ++c) {
        double m[4][4] = {};
        for (int q = 0; q < 4; ++q) {
            const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g;
            double psi[4], dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0};
            for (int v = 0; v < 4; ++v) {
                const double a = 1.0 + sx[v] * xi, b = 1.0 + sy[v] * eta;
                psi[v] = 0.25 * a * b;
                for (int d = 0; d < 3; ++d) {
                    const double X = coords[3 * (size_t)cone[4 * c + v] + d];
                    dxi[d] += X * 0.25 * sx[v] * b;
                    deta[d] += X * 0.25 * a * sy[v];
                }
            }
            const double n0 = dxi[1] * deta[2] - dxi[2] * deta[1], n1 = dxi[2] * deta[0] - dxi[0] * deta[2],
                         n2 = dxi[0] * deta[1] - dxi[1] * deta[0];
            const double area = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) m[a][b] += area * psi[a] * psi[b];
        }
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) rows[lag[4 * c + a]][lag[4 * c + b]] += m[a][b];
    }
    ptr.assign(1, 0);
    col.clear();
    val.clear();
    dinv.assign(num_lag, 0.0);
    for (int i = 0; i < num_lag; ++i) {
        for (const auto &e : rows[i]) {
            col.push_back(e.first);
            val.push_back(e.second);
            if (e.first == i) dinv[i] = 1.0 / e.second;
        }
        ptr.push_back((int32_t)col.size());
    }
}

int launch_gather_pos(plx_dm dm, const double *v, double *out) {
    FaultData &F = *dm->fault;
    hipLaunchKernelGGL(gather_pos_kernel, dim3((F.num_lag + 255) / 256), dim3(256), 0, dm->stream, F.num_lag,
                       F.d_pos_node, v, out);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

void free_fault(FaultData *F) {
    void *ptrs[] = {F->d_neg,      F->d_pos,      F->d_lag,        F->d_cone,      F->d_color_cells,
                    F->d_neg_node, F->d_pos_node, F->d_slip,       F->d_mass_ptr,  F->d_mass_col,
                    F->d_mass_val, F->d_mass_dinv};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete F;
}

}  // namespace plx

using namespace plx;

extern "C" {

int plx_dm_add_fault(plx_dm dm, int num_cohesive, const int32_t *closure_neg, const int32_t *closure_pos,
                     const int32_t *closure_lag, const int32_t *face_cone, int num_lagrange,
                     const double ref_dir_1[3], const double ref_dir_2[3]) {
    if (dm) dm->epoch++, dm->op_epoch++;
    if (!dm || num_cohesive <= 0 || num_lagrange <= 0 || !closure_neg || !closure_pos || !closure_lag ||
        !face_cone || !ref_dir_1 || !ref_dir_2)
        return set_error(dm, 1, "bad fault arguments");
    if (dm->fault) return set_error(dm, 1, "the DM already has a fault");
    if (dm->cell_type != PLX_HEX_Q1) return set_error(dm, 1, "cohesive cells are implemented for Q1 hex meshes");
    const size_t n4 = 4 * (size_t)num_cohesive;
    std::vector<int32_t> neg(closure_neg, closure_neg + n4), pos(closure_pos, closure_pos + n4),
        lag(closure_lag, closure_lag + n4), cone(face_cone, face_cone + n4);
    std::vector<int32_t> neg_node(num_lagrange, -1), pos_node(num_lagrange, -1);
    for (size_t i = 0; i < n4; ++i) {
        if (neg[i] < 0 || neg[i] >= dm->num_nodes || pos[i] < 0 || pos[i] >= dm->num_nodes ||
            lag[i] < 0 || lag[i] >= num_lagrange || cone[i] < 0 || cone[i] >= dm->num_vertices)
            return set_error(dm, 1, "cohesive cell index out of range");
        if ((neg_node[lag[i]] >= 0 && neg_node[lag[i]] != neg[i]) ||
            (pos_node[lag[i]] >= 0 && pos_node[lag[i]] != pos[i]))
            return set_error(dm, 1, "a Lagrange node must pair one negative and one positive node");
        neg_node[lag[i]] = neg[i];
        pos_node[lag[i]] = pos[i];
        if (neg[i] == pos[i]) return set_error(dm, 1, "fault nodes must be split");
    }
    for (int i = 0; i < num_lagrange; ++i)
        if (neg_node[i] < 0) return set_error(dm, 1, "a Lagrange node has no cohesive cell");
    // greedy colouring: cells sharing a node get different colours
    std::vector<int> color(num_cohesive, 0);
    {
        std::vector<std::vector<int>> node_cells(dm->num_nodes);
        for (int c = 0; c < num_cohesive; ++c)
            for (int v = 0; v < 4; ++v) node_cells[neg[4 * c + v]].push_back(c);
        for (int c = 0; c < num_cohesive; ++c) {
            uint64_t used = 0;
            for (int v = 0; v < 4; ++v)
                for (int o : node_cells[neg[4 * c + v]])
                    if (o < c) used |= (uint64_t)1 << std::min(color[o], 63);
            int k = 0;
            while ((used >> k) & 1) ++k;
            color[c] = k;
        }
    }
    const int ncol = *std::max_element(color.begin(), color.end()) + 1;
    std::vector<int32_t> color_cells;
    std::vector<int32_t> color_ptr{0};
    for (int k = 0; k < ncol; ++k) {
        for (int c = 0; c < num_cohesive; ++c)
            if (color[c] == k) color_cells.push_back(c);
        color_ptr.push_back((int32_t)color_cells.size());
    }
    FaultData *F = new FaultData();
    F->num_cohesive = num_cohesive;
    F->num_lag = num_lagrange;
    F->lag_offset = 3 * (int64_t)dm->num_nodes;
    for (int k = 0; k < 3; ++k) {
        F->ref_dir_1[k] = ref_dir_1[k];
        F->ref_dir_2[k] = ref_dir_2[k];
    }
    F->color_ptr = color_ptr;
    F->neg_node = neg_node;
    F->pos_node = pos_node;
    dm->fault = F;
    auto up = [&](int32_t **d, const std::vector<int32_t> &h) -> int {
        if (hipMalloc(d, h.size() * sizeof(int32_t)) != hipSuccess ||
            hipMemcpy(*d, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
            return set_error(dm, 76, "fault upload failed");
        return 0;
    };
    int err = 0;
    if ((err = up(&F->d_neg, neg)) || (err = up(&F->d_pos, pos)) || (err = up(&F->d_lag, lag)) ||
        (err = up(&F->d_cone, cone)) || (err = up(&F->d_color_cells, color_cells)) ||
        (err = up(&F->d_neg_node, neg_node)) || (err = up(&F->d_pos_node, pos_node)))
        return err;
    {
        std::vector<int32_t> mptr, mcol;
        std::vector<double> mval, mdinv;
        assemble_fault_mass(dm->coords, lag, cone, num_lagrange, mptr, mcol, mval, mdinv);
        for (double d : mdinv)
            if (!(d > 0.0)) return set_error(dm, 1, "degenerate fault face");
        if ((err = up(&F->d_mass_ptr, mptr)) || (err = up(&F->d_mass_col, mcol))) return err;
        PLX_HIP(dm, hipMalloc(&F->d_mass_val, mval.size() * sizeof(double)));
        PLX_HIP(dm, hipMemcpy(F->d_mass_val, mval.data(), mval.size() * sizeof(double), hipMemcpyHostToDevice));
        PLX_HIP(dm, hipMalloc(&F->d_mass_dinv, mdinv.size() * sizeof(double)));
        PLX_HIP(dm, hipMemcpy(F->d_mass_dinv, mdinv.data(), mdinv.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    {
        // planar fault: one frame R = [n t1 t2] for every point, so M^-1 integral(psi R s) = R s
        double n0[3] = {0, 0, 0};
        bool planar = true;
        for (int c = 0; c < num_cohesive && planar; ++c) {
            double X[4][3];
            for (int v = 0; v < 4; ++v)
                for (int d = 0; d < 3; ++d) X[v][d] = dm->coords[3 * (size_t)cone[4 * c + v] + d];
            for (int q = 0; q < 4; ++q) {
                double psi[4], area, n[3];
                face_point(X, q, psi, &area, n);
                if (c == 0 && q == 0)
                    for (int d = 0; d < 3; ++d) n0[d] = n[d];
                else if (std::fabs(n[0] - n0[0]) + std::fabs(n[1] - n0[1]) + std::fabs(n[2] - n0[2]) > 1e-12)
                    planar = false;
            }
        }
        F->planar = planar;
        if (planar) {
            double t1[3], t2[3];
            tangents(t1, t2, F->ref_dir_1, F->ref_dir_2, n0);
            for (int d = 0; d < 3; ++d) {
                F->R[3 * d + 0] = n0[d];
                F->R[3 * d + 1] = t1[d];
                F->R[3 * d + 2] = t2[d];
            }
        }
    }
    PLX_HIP(dm, hipMalloc(&F->d_slip, 3 * (size_t)num_lagrange * sizeof(double)));
    PLX_HIP(dm, hipMemset(F->d_slip, 0, 3 * (size_t)num_lagrange * sizeof(double)));
    // the Lagrange-multiplier field follows the displacement field in the local vector
    dm->local_size = 3 * (int64_t)dm->num_nodes + 3 * (int64_t)num_lagrange;
    const std::vector<int32_t> dofs(dm->con.dofs);
    return plx_dm_set_constraints(dm, (int)dofs.size(), dofs.data());
}

int plx_dm_set_fault_slip(plx_dm dm, const double *slip) {
    if (!dm || !slip) return set_error(dm, 1, "null slip");
    if (!dm->fault) return set_error(dm, 1, "the DM has no fault");
    PLX_HIP(dm, hipMemcpy(dm->fault->d_slip, slip, 3 * (size_t)dm->fault->num_lag * sizeof(double),
                          hipMemcpyHostToDevice));
    return 0;
}

int plx_dm_set_fault_slip_device(plx_dm dm, const double *slip) {
    if (!dm || !slip) return set_error(dm, 1, "null slip");
    if (!dm->fault) return set_error(dm, 1, "the DM has no fault");
    PLX_HIP(dm, hipMemcpyAsync(dm->fault->d_slip, slip, 3 * (size_t)dm->fault->num_lag * sizeof(double),
                               hipMemcpyDeviceToDevice, dm->stream));
    return 0;
}

int plx_dm_get_fault_info(plx_dm dm, int64_t *lagrange_offset, int *num_lagrange) {
    if (!dm) return 1;
    if (!dm->fault) return set_error(dm, 1, "the DM has no fault");
    if (lagrange_offset) *lagrange_offset = dm->fault->lag_offset;
    if (num_lagrange) *num_lagrange = dm->fault->num_lag;
    return 0;
}

int plx_compute_residual_cohesive(plx_dm dm, double t, const double *locX, double *locF, int mode) {
    (void)t;
    if (!dm || !locX || !locF) return set_error(dm, 1, "null vector");
    if (!dm->fault) return set_error(dm, 1, "the DM has no fault");
    return launch_cohesive(dm, false, locX, locF, mode);
}

int plx_compute_jacobian_action_cohesive(plx_dm dm, const double *locY, double *locF, int mode) {
    if (!dm || !locY || !locF) return set_error(dm, 1, "null vector");
    if (!dm->fault) return set_error(dm, 1, "the DM has no fault");
    return launch_cohesive(dm, true, locY, locF, mode);
}

}  // extern "C"
