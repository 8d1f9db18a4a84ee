// pylith_amd_pointfn.hpp -- pointwise kernels in PETSc's PetscPointFn / PetscPointJacFn form on
// the device, and the launchers that integrate them over the cells of a key (header-only HIP;
// include it from a hipcc translation unit).
//
// A pointwise kernel keeps the reference's parameter list exactly
// (docs/developer/implementation/petsc-fem.md:447-490; e.g. IsotropicLinearElasticity.hh:
// 804-821 and Jf3vu_infinitesimalStrain, :890-1035):
//
//   __device__ void f1(const PylithInt dim, const PylithInt numS, const PylithInt numA,
//                      const PylithInt sOff[], const PylithInt sOff_x[], const PylithScalar s[],
//                      const PylithScalar s_t[], const PylithScalar s_x[],
//                      const PylithInt aOff[], const PylithInt aOff_x[], const PylithScalar a[],
//                      const PylithScalar a_t[], const PylithScalar a_x[], const PylithReal t,
//                      const PylithScalar x[], const PylithInt numConstants,
//                      const PylithScalar constants[], PylithScalar f1[]);
//
// and the Jacobian kernels add `const PylithReal s_tshift` after t.  PLX_POINTFN / PLX_POINTJAC
// wrap such a function into the functor type the launchers take; `None` is an absent term.
// Then, e.g.,
//
//   PLX_POINTFN(MyF1, f1v_myRheology);
//   PLX_POINTJAC(MyJf3, Jf3vu_myRheology);
//   plx_dm_add_residual_pointfn(dm, label, LHS, pylith_amd::pointfn::launch_residual<None, MyF1>);
//   plx_dm_add_jacobian_pointfn(dm, label, LHS,
//       pylith_amd::pointfn::launch_jacobian_action<None, None, None, MyJf3>,
//       pylith_amd::pointfn::launch_jacobian_assembly<None, None, None, MyJf3>);
//
// registers them on the key (label, part) in place of built-in twins.
//
// What the kernels receive at each quadrature point, as PETSc's PetscFEIntegrateResidual passes
// them: one solution subfield (displacement, numS = 1, sOff = sOff_x = {0}); s[3] the values,
// s_x[3 * c + d] = du_c / dx_d, s_t the time derivative (zero without locX_t); the auxiliary
// record of the cell (order 0) or of the quadrature point with the offsets aOff of the key's
// subfields (a_x and a_t zero: piecewise-constant aux); x the physical point; t, s_tshift and
// the key's constants.  Output layouts are PETSc's: f0[c], f1[c * 3 + d], Jf0[f * 3 + g],
// Jf1[(f * 3 + g) * 3 + e] (trial derivative e), Jf2[(f * 3 + g) * 3 + d] (test derivative d),
// Jf3[((f * 3 + g) * 3 + d) * 3 + e]; kernels accumulate into zeroed outputs.
//
// Integration: Q1 hexahedra with 2x2x2 Gauss points (PyLith's vertex order: 0-3 one face
// cyclic, 4-7 the opposite face), P2 tetrahedra with the 14-point degree-5 rule (vertices, then
// edge nodes (0,1), (0,2), (0,3), (1,2), (1,3), (2,3)).  Residual and action run over the key's
// tiles (nodes staged in LDS, element vectors added in LDS, shared nodes through halo slots);
// with args->deterministic they run one thread per cell colour by colour instead: the cells of
// one colour share no node, so the scatter-add needs no atomics and the summation order is
// fixed (bitwise reproducible).  The assembled Jacobian always takes the colour passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "pylith_amd.h"

namespace pylith_amd {
namespace pointfn {

typedef int32_t PylithInt;
typedef double PylithReal;
typedef double PylithScalar;

// An absent term (no f0, no Jf1, ...).
struct None {};

#define PLX_POINTFN(Name, fn)                                                                      \
    struct Name {                                                                                  \
        __device__ static void f(const pylith_amd::pointfn::PylithInt dim,                         \
                                 const pylith_amd::pointfn::PylithInt numS,                        \
                                 const pylith_amd::pointfn::PylithInt numA,                        \
                                 const pylith_amd::pointfn::PylithInt sOff[],                      \
                                 const pylith_amd::pointfn::PylithInt sOff_x[],                    \
                                 const pylith_amd::pointfn::PylithScalar s[],                      \
                                 const pylith_amd::pointfn::PylithScalar s_t[],                    \
                                 const pylith_amd::pointfn::PylithScalar s_x[],                    \
                                 const pylith_amd::pointfn::PylithInt aOff[],                      \
                                 const pylith_amd::pointfn::PylithInt aOff_x[],                    \
                                 const pylith_amd::pointfn::PylithScalar a[],                      \
                                 const pylith_amd::pointfn::PylithScalar a_t[],                    \
                                 const pylith_amd::pointfn::PylithScalar a_x[],                    \
                                 const pylith_amd::pointfn::PylithReal t,                          \
                                 const pylith_amd::pointfn::PylithScalar x[],                      \
                                 const pylith_amd::pointfn::PylithInt numConstants,                \
                                 const pylith_amd::pointfn::PylithScalar constants[],              \
                                 pylith_amd::pointfn::PylithScalar out[]) {                        \
            fn(dim, numS, numA, sOff, sOff_x, s, s_t, s_x, aOff, aOff_x, a, a_t, a_x, t, x,        \
               numConstants, constants, out);                                                      \
        }                                                                                          \
    }

#define PLX_POINTJAC(Name, fn)                                                                     \
    struct Name {                                                                                  \
        __device__ static void f(const pylith_amd::pointfn::PylithInt dim,                         \
                                 const pylith_amd::pointfn::PylithInt numS,                        \
                                 const pylith_amd::pointfn::PylithInt numA,                        \
                                 const pylith_amd::pointfn::PylithInt sOff[],                      \
                                 const pylith_amd::pointfn::PylithInt sOff_x[],                    \
                                 const pylith_amd::pointfn::PylithScalar s[],                      \
                                 const pylith_amd::pointfn::PylithScalar s_t[],                    \
                                 const pylith_amd::pointfn::PylithScalar s_x[],                    \
                                 const pylith_amd::pointfn::PylithInt aOff[],                      \
                                 const pylith_amd::pointfn::PylithInt aOff_x[],                    \
                                 const pylith_amd::pointfn::PylithScalar a[],                      \
                                 const pylith_amd::pointfn::PylithScalar a_t[],                    \
                                 const pylith_amd::pointfn::PylithScalar a_x[],                    \
                                 const pylith_amd::pointfn::PylithReal t,                          \
                                 const pylith_amd::pointfn::PylithReal s_tshift,                   \
                                 const pylith_amd::pointfn::PylithScalar x[],                      \
                                 const pylith_amd::pointfn::PylithInt numConstants,                \
                                 const pylith_amd::pointfn::PylithScalar constants[],              \
                                 pylith_amd::pointfn::PylithScalar out[]) {                        \
            fn(dim, numS, numA, sOff, sOff_x, s, s_t, s_x, aOff, aOff_x, a, a_t, a_x, t, s_tshift, \
               x, numConstants, constants, out);                                                   \
        }                                                                                          \
    }

constexpr int kMaxAuxValues = 64;  // doubles per auxiliary record

// ---------------------------------------------------------------------------------------------
// Basis and geometry at a quadrature point.
// ---------------------------------------------------------------------------------------------
template <int CT>
struct Basis;

template <>
struct Basis<PLX_HEX_Q1> {
    static constexpr int NB = 8, NV = 8, NQ = 8;
    // reference basis phi[b], d phi / d xi [b][k], weight; the geometry basis is the same
    __device__ static void point(int q, double &w, double phi[8], double dref[8][3]) {
        const double g = 0.57735026918962576451;
        const double xi[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
        const double sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const double a0 = 1.0 + sg[b][0] * xi[0], a1 = 1.0 + sg[b][1] * xi[1], a2 = 1.0 + sg[b][2] * xi[2];
            phi[b] = 0.125 * a0 * a1 * a2;
            dref[b][0] = 0.125 * sg[b][0] * a1 * a2;
            dref[b][1] = 0.125 * a0 * sg[b][1] * a2;
            dref[b][2] = 0.125 * a0 * a1 * sg[b][2];
        }
        w = 1.0;
    }
    // isoparametric map: x = sum_v phi_v X_v, J[d][k] = sum_v X_v[d] dphi_v[k]
    __device__ static void geometry(const double X[8][3], const double phi[8], const double dref[8][3],
                                    double x[3], double J[9]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x[d] = 0.0;
            J[3 * d] = J[3 * d + 1] = J[3 * d + 2] = 0.0;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                x[d] += phi[v] * X[v][d];
                J[3 * d] += X[v][d] * dref[v][0];
                J[3 * d + 1] += X[v][d] * dref[v][1];
                J[3 * d + 2] += X[v][d] * dref[v][2];
            }
        }
    }
};

template <>
struct Basis<PLX_TET_P2> {
    static constexpr int NB = 10, NV = 4, NQ = 14;
    // barycentric coordinates and weight of point q of the 14-point degree-5 rule (two S31
    // orbits and one S22 orbit of the tetrahedral group; weights sum to the volume 1/6)
    __device__ static void rule(int q, double l[4], double &w) {
        const double A0 = 0.09273525031089125323737967, A1 = 0.3108859192633006188837448;
        const double W0 = 0.01224884051939366598892406, W1 = 0.0187813209530026585737831;
        const double B = 0.4544962958743504682934192, W3 = 0.007091003462846894735973002;
        const int edges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        if (q < 8) {
            const double a = q < 4 ? A0 : A1;
            for (int j = 0; j < 4; ++j) l[j] = a;
            l[q & 3] = 1.0 - 3.0 * a;
            w = q < 4 ? W0 : W1;
        } else {
            for (int j = 0; j < 4; ++j) l[j] = 0.5 - B;
            l[edges[q - 8][0]] = B;
            l[edges[q - 8][1]] = B;
            w = W3;
        }
    }
    __device__ static void point(int q, double &w, double phi[10], double dref[10][3], double l[4]) {
        const int edges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        rule(q, l, w);
        for (int v = 0; v < 4; ++v) {
            phi[v] = l[v] * (2.0 * l[v] - 1.0);
            for (int k = 0; k < 3; ++k) dref[v][k] = (4.0 * l[v] - 1.0) * dl[v][k];
        }
        for (int e = 0; e < 6; ++e) {
            const int a = edges[e][0], b = edges[e][1];
            phi[4 + e] = 4.0 * l[a] * l[b];
            for (int k = 0; k < 3; ++k) dref[4 + e][k] = 4.0 * (l[a] * dl[b][k] + l[b] * dl[a][k]);
        }
    }
    // affine map from the four vertices: x = sum_v l_v X_v, J[d][k] = X[k+1][d] - X[0][d]
    __device__ static void geometry(const double X[4][3], const double l[4], double x[3], double J[9]) {
        for (int d = 0; d < 3; ++d) {
            x[d] = l[0] * X[0][d] + l[1] * X[1][d] + l[2] * X[2][d] + l[3] * X[3][d];
            for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
        }
    }
};

__device__ inline double inverse3(const double J[9], double Ji[9]) {
    const double c0 = J[4] * J[8] - J[5] * J[7];
    const double c1 = J[5] * J[6] - J[3] * J[8];
    const double c2 = J[3] * J[7] - J[4] * J[6];
    const double det = J[0] * c0 + J[1] * c1 + J[2] * c2;
    const double id = 1.0 / det;
    Ji[0] = c0 * id;
    Ji[1] = (J[2] * J[7] - J[1] * J[8]) * id;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) * id;
    Ji[3] = c1 * id;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) * id;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) * id;
    Ji[6] = c2 * id;
    Ji[7] = (J[1] * J[6] - J[0] * J[7]) * id;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) * id;
    return det;
}

// Quadrature point q of one cell: weight w |J|, basis values, physical gradients and the point.
template <int CT>
struct PointData {
    static constexpr int NB = Basis<CT>::NB;
    double wdet;
    double phi[NB];
    double grad[NB][3];
    double x[3];
};

template <int CT>
__device__ inline void point_data(const double X[Basis<CT>::NV][3], int q, PointData<CT> &P) {
    constexpr int NB = Basis<CT>::NB;
    double w, dref[NB][3], J[9], Ji[9];
    if constexpr (CT == PLX_HEX_Q1) {
        Basis<CT>::point(q, w, P.phi, dref);
        Basis<CT>::geometry(X, P.phi, dref, P.x, J);
    } else {
        double l[4];
        Basis<CT>::point(q, w, P.phi, dref, l);
        Basis<CT>::geometry(X, l, P.x, J);
    }
    P.wdet = w * inverse3(J, Ji);
    for (int b = 0; b < NB; ++b)
        for (int d = 0; d < 3; ++d)
            P.grad[b][d] = dref[b][0] * Ji[d] + dref[b][1] * Ji[3 + d] + dref[b][2] * Ji[6 + d];
}

// The auxiliary record of (cell, q), and the zero gradient / time derivative arrays.
__device__ inline const double *aux_record(const plx_pointfn_args &A, int cell, int q, int nq) {
    return A.aux + (size_t)(A.aux_per_qp ? (size_t)cell * nq + q : (size_t)cell) * A.aux_stride;
}

// Calls one term when it is registered.
template <class F>
__device__ inline void call_fn(const plx_pointfn_args &A, const int *aOff, const int *aOff_x,
                               const double *a, const double *zero, const double s[3], const double s_t[3],
                               const double s_x[9], const double x[3], double *out) {
    if constexpr (!std::is_same<F, None>::value) {
        const PylithInt sOff[1] = {0}, sOff_x[1] = {0};
        F::f(3, 1, A.num_aux, sOff, sOff_x, s, s_t, s_x, aOff, aOff_x, a, zero, zero, A.t, x, A.num_constants,
             A.constants, out);
    }
}

template <class F>
__device__ inline void call_jac(const plx_pointfn_args &A, const int *aOff, const int *aOff_x,
                                const double *a, const double *zero, const double s[3], const double s_t[3],
                                const double s_x[9], const double x[3], double *out) {
    if constexpr (!std::is_same<F, None>::value) {
        const PylithInt sOff[1] = {0}, sOff_x[1] = {0};
        F::f(3, 1, A.num_aux, sOff, sOff_x, s, s_t, s_x, aOff, aOff_x, a, zero, zero, A.t, A.s_tshift, x,
             A.num_constants, A.constants, out);
    }
}

// Gathers the cell's vertex coordinates and a local vector's nodal values.
template <int CT>
__device__ inline void gather_cell(const plx_pointfn_args &A, int cell, double X[Basis<CT>::NV][3],
                                   const double *vec, double V[Basis<CT>::NB][3]) {
    constexpr int NB = Basis<CT>::NB, NV = Basis<CT>::NV;
    for (int v = 0; v < NV; ++v) {
        const size_t g = (size_t)A.cone[(size_t)cell * NV + v];
        for (int d = 0; d < 3; ++d) X[v][d] = A.coords[3 * g + d];
    }
    for (int b = 0; b < NB; ++b) {
        const size_t g = (size_t)A.closure[(size_t)cell * NB + b];
        for (int c = 0; c < 3; ++c) V[b][c] = vec ? vec[3 * g + c] : 0.0;
    }
}

// Values and gradient of a nodal field at a point.
template <int CT>
__device__ inline void eval_field(const PointData<CT> &P, const double V[Basis<CT>::NB][3], double s[3],
                                  double s_x[9]) {
    constexpr int NB = Basis<CT>::NB;
    for (int c = 0; c < 3; ++c) {
        s[c] = 0.0;
        s_x[3 * c] = s_x[3 * c + 1] = s_x[3 * c + 2] = 0.0;
        for (int b = 0; b < NB; ++b) {
            s[c] += V[b][c] * P.phi[b];
            for (int d = 0; d < 3; ++d) s_x[3 * c + d] += V[b][c] * P.grad[b][d];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------------------------
// INSERT_VALUES: the closure dofs of the key's cells are zeroed before the colour passes.
template <int CT>
__global__ void __launch_bounds__(256) zero_closure(plx_pointfn_args A) {
    constexpr int NB = Basis<CT>::NB;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.num_cells) return;
    const int cell = A.cells[i];
    for (int b = 0; b < NB; ++b) {
        const size_t g = (size_t)A.closure[(size_t)cell * NB + b];
        A.f[3 * g] = A.f[3 * g + 1] = A.f[3 * g + 2] = 0.0;
    }
}

// ---------------------------------------------------------------------------------------------
// Residual and Jacobian action through the structure of the bases (one thread per cell).
//   P2 (affine map): J^-1 once per cell; at a point the values from the 10 basis values and the
//   reference gradient from the barycentric form of the basis gradients; the f1 test sum
//   collects physical moments M_v = sum_q w|J| l_v(q) f1(q) at the four vertices (the P2
//   gradients are linear in the barycentric coordinates), r_b = sum_v (M_v J^-T) : dref_b(v).
//   Q1: the modal (Walsh-Hadamard) form u = sum_b T_b m_b, m_b the products of the reference
//   coordinates in b; values, reference gradients and the Jacobian at a point from the modal
//   coefficients, and the test sums collect modal moments R_b = sum_q (f0 m_b + P : grad m_b).
// Both are the quadrature sums of the tables (identical up to rounding).  The kernels are
// inlined: a term's inputs it never reads (s, s_t, x, ...) cost nothing.
// ---------------------------------------------------------------------------------------------
struct CellP2 {
    static constexpr int NB = 10, NV = 4, NQ = 14;
    double Ji[9], det;
    __device__ void setup(const double X[4][3]) {
        double J[9];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
        det = inverse3(J, Ji);
    }
    // basis values and the reference gradient of a nodal field at barycentric point l
    __device__ static void values(const double l[4], double phi[10]) {
        const int edges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        for (int v = 0; v < 4; ++v) phi[v] = l[v] * (2.0 * l[v] - 1.0);
        for (int e = 0; e < 6; ++e) phi[4 + e] = 4.0 * l[edges[e][0]] * l[edges[e][1]];
    }
    __device__ static void ref_grad(const double l[4], const double V[10][3], double G[9]) {
        const double d0 = 4.0 * l[0] - 1.0, d1 = 4.0 * l[1] - 1.0, d2 = 4.0 * l[2] - 1.0, d3 = 4.0 * l[3] - 1.0;
        for (int c = 0; c < 3; ++c) {
            const double v0 = V[0][c], e01 = V[4][c], e02 = V[5][c], e03 = V[6][c], e12 = V[7][c], e13 = V[8][c],
                         e23 = V[9][c];
            const double base = -d0 * v0;
            G[3 * c + 0] = base + d1 * V[1][c] +
                           4.0 * ((l[0] - l[1]) * e01 - l[2] * e02 - l[3] * e03 + l[2] * e12 + l[3] * e13);
            G[3 * c + 1] = base + d2 * V[2][c] +
                           4.0 * (-l[1] * e01 + (l[0] - l[2]) * e02 - l[3] * e03 + l[1] * e12 + l[3] * e23);
            G[3 * c + 2] = base + d3 * V[3][c] +
                           4.0 * (-l[1] * e01 - l[2] * e02 + (l[0] - l[3]) * e03 + l[1] * e13 + l[2] * e23);
        }
    }
};

// re += the P2 element vector of the moments: f1 part from M[v][c * 3 + d] (physical),
// f0 part R0[b][c] given per node.
__device__ inline void p2_finish(const double Ji[9], const double M[4][9], double re[10][3]) {
    double Mr[4][9], S[9];  // reference moments (component c, direction k) and their sum
    for (int v = 0; v < 4; ++v)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k)
                Mr[v][3 * c + k] = M[v][3 * c] * Ji[3 * k] + M[v][3 * c + 1] * Ji[3 * k + 1] +
                                   M[v][3 * c + 2] * Ji[3 * k + 2];
    for (int i = 0; i < 9; ++i) S[i] = Mr[0][i] + Mr[1][i] + Mr[2][i] + Mr[3][i];
    // vertex a: (4 M_a - S) . e_a; edge (a, b): 4 (M_a . e_b + M_b . e_a); e_0 = -(1, 1, 1)
    for (int c = 0; c < 3; ++c) {
        double me0[4];
        for (int v = 0; v < 4; ++v) me0[v] = -(Mr[v][3 * c] + Mr[v][3 * c + 1] + Mr[v][3 * c + 2]);
        const double se0 = -(S[3 * c] + S[3 * c + 1] + S[3 * c + 2]);
        re[0][c] += 4.0 * me0[0] - se0;
        re[1][c] += 4.0 * Mr[1][3 * c + 0] - S[3 * c + 0];
        re[2][c] += 4.0 * Mr[2][3 * c + 1] - S[3 * c + 1];
        re[3][c] += 4.0 * Mr[3][3 * c + 2] - S[3 * c + 2];
        re[4][c] += 4.0 * (Mr[0][3 * c + 0] + me0[1]);
        re[5][c] += 4.0 * (Mr[0][3 * c + 1] + me0[2]);
        re[6][c] += 4.0 * (Mr[0][3 * c + 2] + me0[3]);
        re[7][c] += 4.0 * (Mr[1][3 * c + 1] + Mr[2][3 * c + 0]);
        re[8][c] += 4.0 * (Mr[1][3 * c + 2] + Mr[3][3 * c + 0]);
        re[9][c] += 4.0 * (Mr[2][3 * c + 2] + Mr[3][3 * c + 1]);
    }
}

// Q1 modal transform: T[b] = sum over corners (PyLith vertex order) of F times the signs of the
// corner's coordinates in b, b = xi + 2 eta + 4 zeta bits (butterflies in lexicographic order).
// The transform is its own transpose, and applied twice gives 8 times the input.
__device__ inline void q1_modal(const double F[8], double T[8]) {
    const double f[8] = {F[0], F[1], F[3], F[2], F[4], F[5], F[7], F[6]};  // lexicographic corners
    double a[8], t[8];
    for (int p = 0; p < 4; ++p) {
        a[2 * p] = f[2 * p] + f[2 * p + 1];
        a[2 * p + 1] = f[2 * p + 1] - f[2 * p];
    }
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < 2; ++i) {
            t[i + 4 * k] = a[i + 4 * k] + a[i + 2 + 4 * k];
            t[i + 2 + 4 * k] = a[i + 2 + 4 * k] - a[i + 4 * k];
        }
    for (int i = 0; i < 4; ++i) {
        T[i] = t[i] + t[i + 4];
        T[i + 4] = t[i + 4] - t[i];
    }
}

// nodal values (PyLith vertex order) of the modal sums: r_v = sum_b R_b S_vb
__device__ inline void q1_unmodal(const double R[8], double r[8]) {
    double a[8], t[8];
    for (int i = 0; i < 4; ++i) {  // zeta
        a[i] = R[i] - R[i + 4];
        a[i + 4] = R[i] + R[i + 4];
    }
    for (int k = 0; k < 2; ++k)  // eta
        for (int i = 0; i < 2; ++i) {
            t[i + 4 * k] = a[i + 4 * k] - a[i + 2 + 4 * k];
            t[i + 2 + 4 * k] = a[i + 4 * k] + a[i + 2 + 4 * k];
        }
    double lx[8];
    for (int p = 0; p < 4; ++p) {  // xi
        lx[2 * p] = t[2 * p] - t[2 * p + 1];
        lx[2 * p + 1] = t[2 * p] + t[2 * p + 1];
    }
    r[0] = lx[0];
    r[1] = lx[1];
    r[2] = lx[3];
    r[3] = lx[2];
    r[4] = lx[4];
    r[5] = lx[5];
    r[6] = lx[7];
    r[7] = lx[6];
}

// Point q of the 2x2x2 rule: modal functions m_b, reference point; weight 1.
struct Q1Point {
    double xi, eta, zeta, m[8];
    __device__ Q1Point(int q) {
        const double g = 0.57735026918962576451;
        xi = (q & 1) ? g : -g;
        eta = (q & 2) ? g : -g;
        zeta = (q & 4) ? g : -g;
        m[0] = 1.0;
        m[1] = xi;
        m[2] = eta;
        m[3] = xi * eta;
        m[4] = zeta;
        m[5] = xi * zeta;
        m[6] = eta * zeta;
        m[7] = xi * eta * zeta;
    }
    // value and reference gradient of a field from its modal coefficients over 8
    __device__ double value(const double T[8]) const {
        double v = 0.0;
        for (int b = 0; b < 8; ++b) v += T[b] * m[b];
        return v;
    }
    __device__ void grad(const double T[8], double g[3]) const {
        g[0] = T[1] + T[3] * eta + T[5] * zeta + T[7] * m[6];
        g[1] = T[2] + T[3] * xi + T[6] * zeta + T[7] * m[5];
        g[2] = T[4] + T[5] * xi + T[6] * eta + T[7] * m[3];
    }
    // R_b += P . grad m_b (reference flux P[k])
    __device__ void add_flux(const double P[3], double R[8]) const {
        R[1] += P[0];
        R[2] += P[1];
        R[4] += P[2];
        R[3] += eta * P[0] + xi * P[1];
        R[5] += zeta * P[0] + xi * P[2];
        R[6] += zeta * P[1] + eta * P[2];
        R[7] += m[6] * P[0] + m[5] * P[1] + m[3] * P[2];
    }
};

// Modal coefficients over 8 of the three components of a nodal field V[8][3].
__device__ inline void q1_modal3(const double V[8][3], double T[3][8]) {
    for (int c = 0; c < 3; ++c) {
        double f[8];
        for (int v = 0; v < 8; ++v) f[v] = V[v][c];
        q1_modal(f, T[c]);
        for (int b = 0; b < 8; ++b) T[c][b] *= 0.125;
    }
}

// The pointwise Jacobian at a point: Jf0 [9], Jf1 [27], Jf2 [27], Jf3 [81], zeroed then filled.
template <class J0, class J1, class J2, class J3>
__device__ inline void point_jacobian(const plx_pointfn_args &A, const int *aOff, const int *aOff_x,
                                      const double *a, const double *zero, const double s[3], const double s_t[3],
                                      const double s_x[9], const double x[3], double g0[9], double g1[27],
                                      double g2[27], double g3[81]) {
    for (int k = 0; k < 9; ++k) g0[k] = 0.0;
    for (int k = 0; k < 27; ++k) g1[k] = g2[k] = 0.0;
    for (int k = 0; k < 81; ++k) g3[k] = 0.0;
    call_jac<J0>(A, aOff, aOff_x, a, zero, s, s_t, s_x, x, g0);
    call_jac<J1>(A, aOff, aOff_x, a, zero, s, s_t, s_x, x, g1);
    call_jac<J2>(A, aOff, aOff_x, a, zero, s, s_t, s_x, x, g2);
    call_jac<J3>(A, aOff, aOff_x, a, zero, s, s_t, s_x, x, g3);
}

// A term's output at a point: the residual's f0 / f1, or the action's t0 / t1 built from the
// pointwise Jacobians and the direction v.
template <class F0, class F1>
struct ResidualTerms {
    static constexpr bool has0 = !std::is_same<F0, None>::value;
    __device__ static void eval(const plx_pointfn_args &A, const int *aOff, const int *aOff_x, const double *a,
                                const double *zero, const double s[3], const double s_t[3], const double s_x[9],
                                const double *, const double *, const double x[3], double f0[3], double f1[9]) {
        for (int c = 0; c < 3; ++c) f0[c] = 0.0;
        for (int k = 0; k < 9; ++k) f1[k] = 0.0;
        call_fn<F0>(A, aOff, aOff_x, a, zero, s, s_t, s_x, x, f0);
        call_fn<F1>(A, aOff, aOff_x, a, zero, s, s_t, s_x, x, f1);
    }
};

template <class J0, class J1, class J2, class J3>
struct ActionTerms {
    static constexpr bool has0 = !std::is_same<J0, None>::value || !std::is_same<J1, None>::value;
    __device__ static void eval(const plx_pointfn_args &A, const int *aOff, const int *aOff_x, const double *a,
                                const double *zero, const double s[3], const double s_t[3], const double s_x[9],
                                const double *v, const double *v_x, const double x[3], double t0[3], double t1[9]) {
        double g0[9], g1[27], g2[27], g3[81];
        point_jacobian<J0, J1, J2, J3>(A, aOff, aOff_x, a, zero, s, s_t, s_x, x, g0, g1, g2, g3);
        for (int f = 0; f < 3; ++f) {
            t0[f] = 0.0;
            for (int g = 0; g < 3; ++g) {
                t0[f] += g0[3 * f + g] * v[g];
                for (int e = 0; e < 3; ++e) t0[f] += g1[(3 * f + g) * 3 + e] * v_x[3 * g + e];
            }
            for (int d = 0; d < 3; ++d) {
                double acc = 0.0;
                for (int g = 0; g < 3; ++g) {
                    acc += g2[(3 * f + g) * 3 + d] * v[g];
                    for (int e = 0; e < 3; ++e) acc += g3[((3 * f + g) * 3 + d) * 3 + e] * v_x[3 * g + e];
                }
                t1[3 * f + d] = acc;
            }
        }
    }
};

// Zero arrays for the a_t / a_x arguments (piecewise-constant aux without time derivative).
static __constant__ double kZeroAux[3 * kMaxAuxValues];

// f += the element vectors of the cells [c0, c1) of one colour: TERMS gives (f0, f1) at a point
// from the solution u (and, for the action, the direction v).
// Element vector re[b][c] of one P2 cell from its vertex coordinates and nodal values (Ut read
// only when A.u_t is set, V only for the action).
template <class TERMS, bool ACTION>
__device__ inline void integrate_p2_cell(const plx_pointfn_args &A, int cell, const double X[4][3],
                                         const double U[10][3], const double Ut[10][3], const double V[10][3],
                                         double re[10][3]) {
    int aOff[16], aOff_x[16];
    for (int k = 0; k < A.num_aux && k < 16; ++k) {
        aOff[k] = A.aux_off[k];
        aOff_x[k] = 3 * A.aux_off[k];
    }
    CellP2 C;
    C.setup(X);
    double M[4][9];
    for (int v = 0; v < 4; ++v)
        for (int i = 0; i < 9; ++i) M[v][i] = 0.0;
    for (int b = 0; b < 10; ++b) re[b][0] = re[b][1] = re[b][2] = 0.0;
#pragma unroll 1
    for (int q = 0; q < CellP2::NQ; ++q) {
        double l[4], w, phi[10];
        Basis<PLX_TET_P2>::rule(q, l, w);
        CellP2::values(l, phi);
        const double wdet = w * C.det;
        double s[3] = {0, 0, 0}, s_t[3] = {0, 0, 0}, x[3], G[9], s_x[9];
        for (int b = 0; b < 10; ++b)
            for (int c = 0; c < 3; ++c) s[c] += U[b][c] * phi[b];
        if (A.u_t)
            for (int b = 0; b < 10; ++b)
                for (int c = 0; c < 3; ++c) s_t[c] += Ut[b][c] * phi[b];
        for (int d = 0; d < 3; ++d) x[d] = l[0] * X[0][d] + l[1] * X[1][d] + l[2] * X[2][d] + l[3] * X[3][d];
        CellP2::ref_grad(l, U, G);
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
                s_x[3 * c + d] = G[3 * c] * C.Ji[d] + G[3 * c + 1] * C.Ji[3 + d] + G[3 * c + 2] * C.Ji[6 + d];
        double v[3] = {0, 0, 0}, v_x[9];
        if (ACTION) {
            for (int b = 0; b < 10; ++b)
                for (int c = 0; c < 3; ++c) v[c] += V[b][c] * phi[b];
            CellP2::ref_grad(l, V, G);
            for (int c = 0; c < 3; ++c)
                for (int d = 0; d < 3; ++d)
                    v_x[3 * c + d] = G[3 * c] * C.Ji[d] + G[3 * c + 1] * C.Ji[3 + d] + G[3 * c + 2] * C.Ji[6 + d];
        }
        double f0[3], f1[9];
        TERMS::eval(A, aOff, aOff_x, aux_record(A, cell, q, CellP2::NQ), kZeroAux, s, s_t, s_x, v, v_x, x, f0, f1);
        for (int i = 0; i < 9; ++i) {
            const double wf = wdet * f1[i];
            M[0][i] += l[0] * wf;
            M[1][i] += l[1] * wf;
            M[2][i] += l[2] * wf;
            M[3][i] += l[3] * wf;
        }
        if (TERMS::has0)
            for (int b = 0; b < 10; ++b)
                for (int c = 0; c < 3; ++c) re[b][c] += wdet * f0[c] * phi[b];
    }
    p2_finish(C.Ji, M, re);
}

template <class TERMS, bool ACTION>
__device__ inline void integrate_p2(const plx_pointfn_args &A, int cell) {
    double X[4][3], U[10][3], Ut[10][3], V[10][3], re[10][3];
    gather_cell<PLX_TET_P2>(A, cell, X, A.u, U);
    if (A.u_t) gather_cell<PLX_TET_P2>(A, cell, X, A.u_t, Ut);
    if (ACTION) gather_cell<PLX_TET_P2>(A, cell, X, A.v, V);
    integrate_p2_cell<TERMS, ACTION>(A, cell, X, U, Ut, V, re);
    for (int b = 0; b < 10; ++b) {
        const size_t g = (size_t)A.closure[(size_t)cell * 10 + b];
        for (int c = 0; c < 3; ++c) A.f[3 * g + c] += re[b][c];
    }
}

// Element vector re[v][c] of one Q1 cell from its corner coordinates and nodal values (Ut read
// only when A.u_t is set, V only for the action).
template <class TERMS, bool ACTION>
__device__ inline void integrate_q1_cell(const plx_pointfn_args &A, int cell, const double X[8][3],
                                         const double U[8][3], const double Ut[8][3], const double V[8][3],
                                         double re[8][3]) {
    double TX[3][8], T[3][8], Tt[3][8], TV[3][8];
    q1_modal3(U, T);
    if (A.u_t) q1_modal3(Ut, Tt);
    if (ACTION) q1_modal3(V, TV);
    q1_modal3(X, TX);
    int aOff[16], aOff_x[16];
    for (int k = 0; k < A.num_aux && k < 16; ++k) {
        aOff[k] = A.aux_off[k];
        aOff_x[k] = 3 * A.aux_off[k];
    }
    double R[3][8];
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 8; ++b) R[c][b] = 0.0;
#pragma unroll 1
    for (int q = 0; q < 8; ++q) {
        const Q1Point P(q);
        double J[9], Ji[9];
        for (int d = 0; d < 3; ++d) P.grad(TX[d], J + 3 * d);
        const double wdet = inverse3(J, Ji);  // weight 1
        double s[3], s_t[3] = {0, 0, 0}, x[3], s_x[9];
        for (int c = 0; c < 3; ++c) {
            double G[3];
            s[c] = P.value(T[c]);
            if (A.u_t) s_t[c] = P.value(Tt[c]);
            P.grad(T[c], G);
            for (int d = 0; d < 3; ++d) s_x[3 * c + d] = G[0] * Ji[d] + G[1] * Ji[3 + d] + G[2] * Ji[6 + d];
            x[c] = P.value(TX[c]);
        }
        double v[3] = {0, 0, 0}, v_x[9];
        if (ACTION)
            for (int c = 0; c < 3; ++c) {
                double G[3];
                v[c] = P.value(TV[c]);
                P.grad(TV[c], G);
                for (int d = 0; d < 3; ++d) v_x[3 * c + d] = G[0] * Ji[d] + G[1] * Ji[3 + d] + G[2] * Ji[6 + d];
            }
        double f0[3], f1[9];
        TERMS::eval(A, aOff, aOff_x, aux_record(A, cell, q, 8), kZeroAux, s, s_t, s_x, v, v_x, x, f0, f1);
        for (int c = 0; c < 3; ++c) {
            // reference flux P_k = w|J| sum_d f1[c][d] Jinv[k][d]
            double Pk[3];
            for (int k = 0; k < 3; ++k)
                Pk[k] = wdet * (f1[3 * c] * Ji[3 * k] + f1[3 * c + 1] * Ji[3 * k + 1] + f1[3 * c + 2] * Ji[3 * k + 2]);
            P.add_flux(Pk, R[c]);
            if (TERMS::has0)
                for (int b = 0; b < 8; ++b) R[c][b] += wdet * f0[c] * P.m[b];
        }
    }
    for (int c = 0; c < 3; ++c) {
        double r[8];
        q1_unmodal(R[c], r);  // the modes of the test functions carry the factor 1/8
        for (int v = 0; v < 8; ++v) re[v][c] = 0.125 * r[v];
    }
}

template <class TERMS, bool ACTION>
__device__ inline void integrate_q1(const plx_pointfn_args &A, int cell) {
    double X[8][3], U[8][3], Ut[8][3], V[8][3], re[8][3];
    gather_cell<PLX_HEX_Q1>(A, cell, X, A.u, U);
    if (A.u_t) gather_cell<PLX_HEX_Q1>(A, cell, X, A.u_t, Ut);
    if (ACTION) gather_cell<PLX_HEX_Q1>(A, cell, X, A.v, V);
    integrate_q1_cell<TERMS, ACTION>(A, cell, X, U, Ut, V, re);
    for (int v = 0; v < 8; ++v) {
        const size_t g = (size_t)A.closure[(size_t)cell * 8 + v];
        for (int c = 0; c < 3; ++c) A.f[3 * g + c] += re[v][c];
    }
}

// One tile per workgroup: the tile's vertex coordinates and nodal values are staged in LDS,
// each thread integrates cells of the tile through the pointwise kernels and adds their element
// vectors into an LDS node accumulator (FP64 LDS atomics), then the tile writes its owned nodes
// and leaves the shared nodes' partial sums in their halo slots (reduce_shared adds them).
// Consecutive tiles run on one XCD (workgroups are dispatched round-robin over the 8 XCDs), so
// nodes shared by neighbouring tiles meet in that XCD's L2.
// Threads per tile workgroup: Q1 tiles hold up to 512 cells (2 per thread), P2 tiles up to 384
// (3 per thread in 2-wave workgroups, ~4 resident per CU: 5 % faster than 256 threads with one
// or two cells each).
template <int CT>
struct TileThreads {
    static constexpr int NT = CT == PLX_TET_P2 ? 128 : 256;
};

template <class TERMS, bool ACTION, int CT>
__global__ void __launch_bounds__(TileThreads<CT>::NT) tile_cells(plx_pointfn_args A) {
    constexpr int NB = Basis<CT>::NB, NV = Basis<CT>::NV, kTileThreads = TileThreads<CT>::NT;
    extern __shared__ double pf_lds[];
    const int tid = threadIdx.x;
    const int n = A.num_tiles, blk = blockIdx.x;
    const int per = n >> 3, rem = n & 7, xcd = blk & 7;
    const int tile = xcd * per + (xcd < rem ? xcd : rem) + (blk >> 3);
    const int c0 = A.tile_cell_ptr[tile], nc = A.tile_cell_ptr[tile + 1] - c0;
    const int n0 = A.tile_node_ptr[tile], nn = A.tile_node_ptr[tile + 1] - n0;
    const int v0 = A.tile_vert_ptr[tile], nvt = A.tile_vert_ptr[tile + 1] - v0;
    const int nown = A.tile_nown[tile];
    const bool has_t = A.u_t != nullptr;
    double *sX = pf_lds;
    double *sU = sX + 3 * A.tile_max_verts;
    double *sUt = sU + 3 * A.tile_max_nodes;
    double *sV = sUt + (has_t ? 3 * A.tile_max_nodes : 0);
    double *sR = sV + (ACTION ? 3 * A.tile_max_nodes : 0);
    for (int k = tid; k < nvt; k += kTileThreads) {
        const size_t g = (size_t)A.tile_verts[v0 + k];
        for (int d = 0; d < 3; ++d) sX[3 * k + d] = A.coords[3 * g + d];
    }
    for (int k = tid; k < nn; k += kTileThreads) {
        const size_t g = (size_t)A.tile_nodes[n0 + k];
        for (int c = 0; c < 3; ++c) {
            sU[3 * k + c] = A.u ? A.u[3 * g + c] : 0.0;  // an action may come without a state (zero)
            if (has_t) sUt[3 * k + c] = A.u_t[3 * g + c];
            if (ACTION) sV[3 * k + c] = A.v[3 * g + c];
            sR[3 * k + c] = 0.0;
        }
    }
    __syncthreads();
    for (int i = tid; i < nc; i += kTileThreads) {
        const size_t p = (size_t)(c0 + i);
        const int cell = A.tile_cells[p];
        int lc[NB];
        double X[NV][3], U[NB][3], Ut[NB][3], V[NB][3], re[NB][3];
        for (int v = 0; v < NV; ++v) {
            const int lv = A.lcone[(size_t)v * A.num_cells + p];
            for (int d = 0; d < 3; ++d) X[v][d] = sX[3 * lv + d];
        }
        for (int b = 0; b < NB; ++b) {
            lc[b] = A.lclosure[(size_t)b * A.num_cells + p];
            for (int c = 0; c < 3; ++c) {
                U[b][c] = sU[3 * lc[b] + c];
                Ut[b][c] = has_t ? sUt[3 * lc[b] + c] : 0.0;
                V[b][c] = ACTION ? sV[3 * lc[b] + c] : 0.0;
            }
        }
        if constexpr (CT == PLX_TET_P2) integrate_p2_cell<TERMS, ACTION>(A, cell, X, U, Ut, V, re);
        else integrate_q1_cell<TERMS, ACTION>(A, cell, X, U, Ut, V, re);
        for (int b = 0; b < NB; ++b)
            for (int c = 0; c < 3; ++c) atomicAdd(&sR[3 * lc[b] + c], re[b][c]);
    }
    __syncthreads();
    const bool insert = A.mode == PLX_INSERT_VALUES;
    for (int k = tid; k < nown; k += kTileThreads) {
        const size_t g = (size_t)A.tile_nodes[n0 + k];
        for (int c = 0; c < 3; ++c) A.f[3 * g + c] = insert ? sR[3 * k + c] : A.f[3 * g + c] + sR[3 * k + c];
    }
    const size_t hoff = (size_t)A.tile_halo_off[tile];
    for (int k = nown + tid; k < nn; k += kTileThreads) {
        const size_t sl = hoff + (size_t)(k - nown);
        for (int c = 0; c < 3; ++c) A.halo[3 * sl + c] = sR[3 * k + c];
    }
}

// Shared nodes: the tiles' partial sums added in slot (tile) order.
static __global__ void __launch_bounds__(256) reduce_shared(plx_pointfn_args A) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= A.num_shared) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (int j = A.shared_ptr[s]; j < A.shared_ptr[s + 1]; ++j)
        for (int c = 0; c < 3; ++c) a[c] += A.halo[3 * (size_t)A.shared_slot[j] + c];
    const size_t g = 3 * (size_t)A.shared_node[s];
    const bool insert = A.mode == PLX_INSERT_VALUES;
    for (int c = 0; c < 3; ++c) A.f[g + c] = insert ? a[c] : A.f[g + c] + a[c];
}

// f += sum_q w|J| (f0 . phi + f1 : grad phi) over the cells [c0, c1) of one colour.
template <class F0, class F1, int CT>
__global__ void __launch_bounds__(128) residual_cells(plx_pointfn_args A, int c0, int c1) {
    const int i = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c1) return;
    if constexpr (CT == PLX_TET_P2) integrate_p2<ResidualTerms<F0, F1>, false>(A, A.cells[i]);
    else integrate_q1<ResidualTerms<F0, F1>, false>(A, A.cells[i]);
}

// f += J(u) v over the cells [c0, c1) of one colour.
template <class J0, class J1, class J2, class J3, int CT>
__global__ void __launch_bounds__(128) action_cells(plx_pointfn_args A, int c0, int c1) {
    const int i = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c1) return;
    if constexpr (CT == PLX_TET_P2) integrate_p2<ActionTerms<J0, J1, J2, J3>, true>(A, A.cells[i]);
    else integrate_q1<ActionTerms<J0, J1, J2, J3>, true>(A, A.cells[i]);
}

// values += the element blocks of the cells [c0, c1) of one colour; thread = (cell, row node a).
template <class J0, class J1, class J2, class J3, int CT>
__global__ void __launch_bounds__(128) assemble_cells(plx_pointfn_args A, int c0, int c1) {
    constexpr int NB = Basis<CT>::NB, NV = Basis<CT>::NV, NQ = Basis<CT>::NQ;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = c0 + t / NB, a_row = t % NB;
    if (i >= c1) return;
    const int cell = A.cells[i];
    double X[NV][3], U[NB][3], Ut[NB][3];
    gather_cell<CT>(A, cell, X, A.u, U);
    gather_cell<CT>(A, cell, X, A.u_t, Ut);
    int aOff[16], aOff_x[16];
    for (int k = 0; k < A.num_aux && k < 16; ++k) {
        aOff[k] = A.aux_off[k];
        aOff_x[k] = 3 * A.aux_off[k];
    }
    double zero[3 * kMaxAuxValues];
    for (int k = 0; k < 3 * kMaxAuxValues; ++k) zero[k] = 0.0;
    double K[NB][9];
    for (int b = 0; b < NB; ++b)
        for (int m = 0; m < 9; ++m) K[b][m] = 0.0;
    for (int q = 0; q < NQ; ++q) {
        PointData<CT> P;
        point_data<CT>(X, q, P);
        double s[3], s_x[9], s_t[3], st_x[9];
        eval_field<CT>(P, U, s, s_x);
        eval_field<CT>(P, Ut, s_t, st_x);
        double g0[9], g1[27], g2[27], g3[81];
        point_jacobian<J0, J1, J2, J3>(A, aOff, aOff_x, aux_record(A, cell, q, NQ), zero, s, s_t, s_x, P.x, g0, g1,
                                       g2, g3);
        const double pa = P.phi[a_row];
        const double *ga = P.grad[a_row];
        for (int b = 0; b < NB; ++b) {
            const double pb = P.phi[b];
            const double *gb = P.grad[b];
            for (int f = 0; f < 3; ++f)
                for (int g = 0; g < 3; ++g) {
                    const int fg = 3 * f + g;
                    double acc = pa * g0[fg] * pb;
                    for (int e = 0; e < 3; ++e) acc += pa * g1[fg * 3 + e] * gb[e] + ga[e] * g2[fg * 3 + e] * pb;
                    for (int d = 0; d < 3; ++d)
                        for (int e = 0; e < 3; ++e) acc += ga[d] * g3[(fg * 3 + d) * 3 + e] * gb[e];
                    K[b][fg] += P.wdet * acc;
                }
        }
    }
    const int row = A.closure[(size_t)cell * NB + a_row];
    const int e0 = A.row_ptr[row], e1 = A.row_ptr[row + 1];
    for (int b = 0; b < NB; ++b) {
        const int colnode = A.closure[(size_t)cell * NB + b];
        int lo = e0, hi = e1;  // the block of column colnode in the sorted row
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (A.col[mid] < colnode) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= e1 || A.col[lo] != colnode) continue;  // outside the pattern (cannot happen)
        double *blk = A.values + 9 * (size_t)lo;
        for (int m = 0; m < 9; ++m) blk[m] += K[b][m];
    }
}

static __global__ void __launch_bounds__(256) fill_values(int64_t n, double *v) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        v[t] = 0.0;
}

// ---------------------------------------------------------------------------------------------
// Host launchers (plx_pointfn_launcher).
// ---------------------------------------------------------------------------------------------
inline bool pointfn_args_ok(const plx_pointfn_args *A) {
    return A && A->num_aux >= 0 && A->num_aux <= 16 && A->aux_stride <= kMaxAuxValues &&
           (A->cell_type == PLX_HEX_Q1 || A->cell_type == PLX_TET_P2);
}

// The tile path applies when the key's plan is there and the caller did not ask for fixed
// summation order.
inline bool use_tiles(const plx_pointfn_args *A) {
    return !A->deterministic && A->num_tiles > 0 && A->tile_cell_ptr && A->tile_nodes && A->lclosure &&
           A->lcone && A->tile_cells && (A->num_shared == 0 || (A->halo && A->shared_ptr));
}

template <class TERMS, bool ACTION, int CT>
inline int launch_tiles(const plx_pointfn_args *A, hipStream_t s) {
    const int fields = 2 + (A->u_t ? 1 : 0) + (ACTION ? 1 : 0);  // u, accumulator, u_t, v
    const size_t lds = sizeof(double) * 3 * ((size_t)A->tile_max_verts + (size_t)fields * A->tile_max_nodes);
    if (lds > 65536 && hipFuncSetAttribute(reinterpret_cast<const void *>(&tile_cells<TERMS, ACTION, CT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return 76;
    hipLaunchKernelGGL((tile_cells<TERMS, ACTION, CT>), dim3(A->num_tiles), dim3(TileThreads<CT>::NT), lds, s, *A);
    if (A->num_shared > 0)
        hipLaunchKernelGGL(reduce_shared, dim3((A->num_shared + 255) / 256), dim3(256), 0, s, *A);
    return hipGetLastError() == hipSuccess ? 0 : 76;
}

template <class F0, class F1, int CT>
inline int launch_residual_ct(const plx_pointfn_args *A, hipStream_t s) {
    if (use_tiles(A)) return launch_tiles<ResidualTerms<F0, F1>, false, CT>(A, s);
    if (A->mode == PLX_INSERT_VALUES && A->num_cells > 0)
        hipLaunchKernelGGL(zero_closure<CT>, dim3((A->num_cells + 255) / 256), dim3(256), 0, s, *A);
    for (int k = 0; k < A->num_colors; ++k) {
        const int c0 = A->color_ptr[k], c1 = A->color_ptr[k + 1];
        if (c1 > c0)
            hipLaunchKernelGGL((residual_cells<F0, F1, CT>), dim3((c1 - c0 + 127) / 128), dim3(128), 0, s, *A, c0, c1);
    }
    return hipGetLastError() == hipSuccess ? 0 : 76;
}

template <class F0, class F1>
int launch_residual(const plx_pointfn_args *A, void *stream) {
    if (!pointfn_args_ok(A) || !A->f || !A->u) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return A->cell_type == PLX_HEX_Q1 ? launch_residual_ct<F0, F1, PLX_HEX_Q1>(A, s)
                                      : launch_residual_ct<F0, F1, PLX_TET_P2>(A, s);
}

template <class J0, class J1, class J2, class J3, int CT>
inline int launch_action_ct(const plx_pointfn_args *A, hipStream_t s) {
    if (use_tiles(A)) return launch_tiles<ActionTerms<J0, J1, J2, J3>, true, CT>(A, s);
    if (A->mode == PLX_INSERT_VALUES && A->num_cells > 0)
        hipLaunchKernelGGL(zero_closure<CT>, dim3((A->num_cells + 255) / 256), dim3(256), 0, s, *A);
    for (int k = 0; k < A->num_colors; ++k) {
        const int c0 = A->color_ptr[k], c1 = A->color_ptr[k + 1];
        if (c1 > c0)
            hipLaunchKernelGGL((action_cells<J0, J1, J2, J3, CT>), dim3((c1 - c0 + 127) / 128), dim3(128), 0, s, *A,
                               c0, c1);
    }
    return hipGetLastError() == hipSuccess ? 0 : 76;
}

template <class J0, class J1, class J2, class J3>
int launch_jacobian_action(const plx_pointfn_args *A, void *stream) {
    if (!pointfn_args_ok(A) || !A->f || !A->v) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return A->cell_type == PLX_HEX_Q1 ? launch_action_ct<J0, J1, J2, J3, PLX_HEX_Q1>(A, s)
                                      : launch_action_ct<J0, J1, J2, J3, PLX_TET_P2>(A, s);
}

template <class J0, class J1, class J2, class J3, int CT>
inline int launch_assembly_ct(const plx_pointfn_args *A, hipStream_t s) {
    constexpr int NB = Basis<CT>::NB;
    for (int k = 0; k < A->num_colors; ++k) {
        const int c0 = A->color_ptr[k], c1 = A->color_ptr[k + 1];
        if (c1 > c0)
            hipLaunchKernelGGL((assemble_cells<J0, J1, J2, J3, CT>), dim3(((c1 - c0) * NB + 127) / 128), dim3(128), 0,
                               s, *A, c0, c1);
    }
    return hipGetLastError() == hipSuccess ? 0 : 76;
}

// values (+)= the element matrices; INSERT_VALUES zeroing is the caller's (the library zeroes
// the whole matrix once before the first key).
template <class J0, class J1, class J2, class J3>
int launch_jacobian_assembly(const plx_pointfn_args *A, void *stream) {
    if (!pointfn_args_ok(A) || !A->values || !A->row_ptr || !A->col) return 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return A->cell_type == PLX_HEX_Q1 ? launch_assembly_ct<J0, J1, J2, J3, PLX_HEX_Q1>(A, s)
                                      : launch_assembly_ct<J0, J1, J2, J3, PLX_TET_P2>(A, s);
}

}  // namespace pointfn
}  // namespace pylith_amd
