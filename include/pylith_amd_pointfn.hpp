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
// edge nodes (0,1), (0,2), (0,3), (1,2), (1,3), (2,3)).  One thread per cell; the cells of one
// colour share no node, so the scatter-add needs no atomics and the summation order is fixed
// (bitwise reproducible).
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

// f += sum_q w|J| (f0 . phi + f1 : grad phi) over the cells [c0, c1) of one colour.
template <class F0, class F1, int CT>
__global__ void __launch_bounds__(128) residual_cells(plx_pointfn_args A, int c0, int c1) {
    constexpr int NB = Basis<CT>::NB, NV = Basis<CT>::NV, NQ = Basis<CT>::NQ;
    const int i = c0 + blockIdx.x * blockDim.x + threadIdx.x;
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
    double re[NB][3];
    for (int b = 0; b < NB; ++b) re[b][0] = re[b][1] = re[b][2] = 0.0;
    for (int q = 0; q < NQ; ++q) {
        PointData<CT> P;
        point_data<CT>(X, q, P);
        double s[3], s_x[9], s_t[3], st_x[9];
        eval_field<CT>(P, U, s, s_x);
        eval_field<CT>(P, Ut, s_t, st_x);
        const double *a = aux_record(A, cell, q, NQ);
        double f0[3] = {0, 0, 0}, f1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        call_fn<F0>(A, aOff, aOff_x, a, zero, s, s_t, s_x, P.x, f0);
        call_fn<F1>(A, aOff, aOff_x, a, zero, s, s_t, s_x, P.x, f1);
        for (int b = 0; b < NB; ++b)
            for (int c = 0; c < 3; ++c)
                re[b][c] += P.wdet * (f0[c] * P.phi[b] + f1[3 * c] * P.grad[b][0] + f1[3 * c + 1] * P.grad[b][1] +
                                      f1[3 * c + 2] * P.grad[b][2]);
    }
    for (int b = 0; b < NB; ++b) {
        const size_t g = (size_t)A.closure[(size_t)cell * NB + b];
        for (int c = 0; c < 3; ++c) A.f[3 * g + c] += re[b][c];
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

// f += J(u) v over the cells [c0, c1) of one colour.
template <class J0, class J1, class J2, class J3, int CT>
__global__ void __launch_bounds__(128) action_cells(plx_pointfn_args A, int c0, int c1) {
    constexpr int NB = Basis<CT>::NB, NV = Basis<CT>::NV, NQ = Basis<CT>::NQ;
    const int i = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c1) return;
    const int cell = A.cells[i];
    double X[NV][3], U[NB][3], Ut[NB][3], V[NB][3];
    gather_cell<CT>(A, cell, X, A.u, U);
    gather_cell<CT>(A, cell, X, A.u_t, Ut);
    gather_cell<CT>(A, cell, X, A.v, V);
    int aOff[16], aOff_x[16];
    for (int k = 0; k < A.num_aux && k < 16; ++k) {
        aOff[k] = A.aux_off[k];
        aOff_x[k] = 3 * A.aux_off[k];
    }
    double zero[3 * kMaxAuxValues];
    for (int k = 0; k < 3 * kMaxAuxValues; ++k) zero[k] = 0.0;
    double re[NB][3];
    for (int b = 0; b < NB; ++b) re[b][0] = re[b][1] = re[b][2] = 0.0;
    for (int q = 0; q < NQ; ++q) {
        PointData<CT> P;
        point_data<CT>(X, q, P);
        double s[3], s_x[9], s_t[3], st_x[9], v[3], v_x[9];
        eval_field<CT>(P, U, s, s_x);
        eval_field<CT>(P, Ut, s_t, st_x);
        eval_field<CT>(P, V, v, v_x);
        double g0[9], g1[27], g2[27], g3[81];
        point_jacobian<J0, J1, J2, J3>(A, aOff, aOff_x, aux_record(A, cell, q, NQ), zero, s, s_t, s_x, P.x, g0, g1,
                                       g2, g3);
        double t0[3], t1[9];
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
        for (int b = 0; b < NB; ++b)
            for (int c = 0; c < 3; ++c)
                re[b][c] += P.wdet * (t0[c] * P.phi[b] + t1[3 * c] * P.grad[b][0] + t1[3 * c + 1] * P.grad[b][1] +
                                      t1[3 * c + 2] * P.grad[b][2]);
    }
    for (int b = 0; b < NB; ++b) {
        const size_t g = (size_t)A.closure[(size_t)cell * NB + b];
        for (int c = 0; c < 3; ++c) A.f[3 * g + c] += re[b][c];
    }
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

template <class F0, class F1, int CT>
inline int launch_residual_ct(const plx_pointfn_args *A, hipStream_t s) {
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
