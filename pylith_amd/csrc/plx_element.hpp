// plx_element.hpp -- element math of the residual / Jacobian-action apply, compiled for both
// the gfx950 kernels and the host emulation used by the CPU test suite.
#pragma once

#include <hip/hip_runtime.h>

namespace plx {

constexpr int kP2Quad = 14;

struct P2Tables {
    double w[kP2Quad];
    double dphi[kP2Quad][10][3];
    double lam[kP2Quad][4];  // barycentric coordinates of the points
    double wdiag, woff;      // P1 moments of the rule: sum_q w_q l_c l_d = wdiag (c == d), woff (c != d)
};

// Reference tables of the P2 tetrahedron at the 14-point degree-5 rule (two S31 orbits and
// one S22 orbit of the symmetric tetrahedral group; weights sum to the volume 1/6).
inline void p2_fill_tables(P2Tables *t) {
    const double A[2] = {0.09273525031089125323737967, 0.3108859192633006188837448};
    const double W[2] = {0.01224884051939366598892406, 0.0187813209530026585737831};
    const double B = 0.4544962958743504682934192, W3 = 0.007091003462846894735973002;
    const int edges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    double lam[kP2Quad][4];
    int q = 0;
    for (int o = 0; o < 2; ++o)
        for (int i = 0; i < 4; ++i, ++q) {
            for (int j = 0; j < 4; ++j) lam[q][j] = A[o];
            lam[q][i] = 1.0 - 3.0 * A[o];
            t->w[q] = W[o];
        }
    for (int e = 0; e < 6; ++e, ++q) {
        for (int j = 0; j < 4; ++j) lam[q][j] = 0.5 - B;
        lam[q][edges[e][0]] = B;
        lam[q][edges[e][1]] = B;
        t->w[q] = W3;
    }
    const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    t->wdiag = t->woff = 0.0;
    for (q = 0; q < kP2Quad; ++q) {
        for (int v = 0; v < 4; ++v) t->lam[q][v] = lam[q][v];
        t->wdiag += t->w[q] * lam[q][0] * lam[q][0];
        t->woff += t->w[q] * lam[q][0] * lam[q][1];
    }
    for (q = 0; q < kP2Quad; ++q) {
        for (int v = 0; v < 4; ++v)
            for (int d = 0; d < 3; ++d) t->dphi[q][v][d] = (4.0 * lam[q][v] - 1.0) * dl[v][d];
        for (int e = 0; e < 6; ++e) {
            const int a = edges[e][0], b = edges[e][1];
            for (int d = 0; d < 3; ++d)
                t->dphi[q][4 + e][d] = 4.0 * (lam[q][a] * dl[b][d] + lam[q][b] * dl[a][d]);
        }
    }
}

__host__ __device__ inline double inv3(const double J[9], double Ji[9]) {
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

// Device twin of the isotropic stress of IsotropicLinearElasticity::cauchyStress
// (IsotropicLinearElasticity.hh:182-255) contracted into the reference-space flux
//   P[c][k] = -scale * sum_d sigma[c][d] Jinv[k][d]
// where sigma = lambda tr(eps) I + 2 mu eps, eps = sym(grad u), grad u = Ghat Jinv.
// The same map is the Jf3 contraction of Jf3vu_infinitesimalStrain (21 nonzeros of -C,
// IsotropicLinearElasticity.hh:1014-1034) applied to grad v.
template <typename Real>
__host__ __device__ inline void iso_flux(const Real G[9], const Real Ji[9], Real lam, Real mu,
                                         Real scale, Real P[9]) {
    Real g[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            g[3 * c + d] = G[3 * c + 0] * Ji[0 * 3 + d] + G[3 * c + 1] * Ji[1 * 3 + d] +
                           G[3 * c + 2] * Ji[2 * 3 + d];
    const Real tr = g[0] + g[4] + g[8];
    const Real ls = -scale * lam * tr;
    const Real m2 = -scale * Real(2) * mu;
    const Real m1 = -scale * mu;
    Real s[9];
    s[0] = ls + m2 * g[0];
    s[4] = ls + m2 * g[4];
    s[8] = ls + m2 * g[8];
    s[1] = s[3] = m1 * (g[1] + g[3]);
    s[5] = s[7] = m1 * (g[5] + g[7]);
    s[2] = s[6] = m1 * (g[2] + g[6]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            P[3 * c + k] = s[3 * c + 0] * Ji[3 * k + 0] + s[3 * c + 1] * Ji[3 * k + 1] +
                           s[3 * c + 2] * Ji[3 * k + 2];
}

// Sum of the products a_k b_k over the k set in the 3-bit mask M (terms of structurally zero
// operands are not formed: FP64 0 * x cannot be folded away by the compiler).
template <int M, typename Real>
__host__ __device__ inline Real msum3(Real a0, Real b0, Real a1, Real b1, Real a2, Real b2) {
    if constexpr (M == 1) return a0 * b0;
    else if constexpr (M == 2) return a1 * b1;
    else if constexpr (M == 4) return a2 * b2;
    else if constexpr (M == 3) return a0 * b0 + a1 * b1;
    else if constexpr (M == 5) return a0 * b0 + a2 * b2;
    else if constexpr (M == 6) return a1 * b1 + a2 * b2;
    else return a0 * b0 + a1 * b1 + a2 * b2;
}

// iso_flux for a reference gradient whose columns outside IN are zero, forming only the flux
// columns in OUT (bit k: reference direction k).
template <typename Real, int IN, int OUT>
__host__ __device__ inline void iso_flux_cols(const Real G[9], const Real Ji[9], Real lam, Real mu, Real scale,
                                              Real P[9]) {
    Real g[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            g[3 * c + d] = msum3<IN, Real>(G[3 * c + 0], Ji[0 * 3 + d], G[3 * c + 1], Ji[1 * 3 + d], G[3 * c + 2],
                                           Ji[2 * 3 + d]);
    const Real tr = g[0] + g[4] + g[8];
    const Real ls = -scale * lam * tr;
    const Real m2 = -scale * Real(2) * mu;
    const Real m1 = -scale * mu;
    Real s[9];
    s[0] = ls + m2 * g[0];
    s[4] = ls + m2 * g[4];
    s[8] = ls + m2 * g[8];
    s[1] = s[3] = m1 * (g[1] + g[3]);
    s[5] = s[7] = m1 * (g[5] + g[7]);
    s[2] = s[6] = m1 * (g[2] + g[6]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if ((OUT >> k) & 1)
                P[3 * c + k] = s[3 * c + 0] * Ji[3 * k + 0] + s[3 * c + 1] * Ji[3 * k + 1] +
                               s[3 * c + 2] * Ji[3 * k + 2];
}

// iso_flux plus a given stress ps (vector order xx, yy, zz, xy, yz, xz) added to sigma.
template <typename Real>
__host__ __device__ inline void iso_flux_pre(const Real G[9], const Real Ji[9], Real lam, Real mu,
                                             Real scale, const Real ps[6], Real P[9]) {
    Real g[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            g[3 * c + d] = G[3 * c + 0] * Ji[0 * 3 + d] + G[3 * c + 1] * Ji[1 * 3 + d] +
                           G[3 * c + 2] * Ji[2 * 3 + d];
    const Real tr = g[0] + g[4] + g[8];
    Real s[9];
    s[0] = -scale * (lam * tr + Real(2) * mu * g[0] + ps[0]);
    s[4] = -scale * (lam * tr + Real(2) * mu * g[4] + ps[1]);
    s[8] = -scale * (lam * tr + Real(2) * mu * g[8] + ps[2]);
    s[1] = s[3] = -scale * (mu * (g[1] + g[3]) + ps[3]);
    s[5] = s[7] = -scale * (mu * (g[5] + g[7]) + ps[4]);
    s[2] = s[6] = -scale * (mu * (g[2] + g[6]) + ps[5]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            P[3 * c + k] = s[3 * c + 0] * Ji[3 * k + 0] + s[3 * c + 1] * Ji[3 * k + 1] +
                           s[3 * c + 2] * Ji[3 * k + 2];
}

// --------------------------------------------------------------------------------------------
// Q1 hexahedron, 2x2x2 Gauss.  Lexicographic corner index l = i + 2j + 4k maps to PyLith's
// vertex order through q1_lex.
// --------------------------------------------------------------------------------------------
__host__ __device__ constexpr int q1_lex(int l) {
    return l == 2 ? 3 : l == 3 ? 2 : l == 6 ? 7 : l == 7 ? 6 : l;
}

// Reference gradients of a P2 field at the four vertices, one solution component at a time:
// S = (u1-u0, u2-u0, u3-u0), G_v = 4 d_v - S with d_0 = (u01-u0, u02-u0, u03-u0),
// d_1 = (u1-u01, u12-u01, u13-u01), d_2 = (u12-u02, u2-u02, u23-u02), d_3 = (u13-u03, u23-u03, u3-u03).
// G[v][3 c + k]: component c, reference direction k.
template <typename Real, typename UAcc>
__host__ __device__ inline void p2_vertex_grads(const UAcc &U, Real G[4][9]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Real u[10];
#pragma unroll
        for (int b = 0; b < 10; ++b) u[b] = U(c, b);
        const Real S0 = u[1] - u[0], S1 = u[2] - u[0], S2 = u[3] - u[0];
        G[0][3 * c + 0] = Real(4) * (u[4] - u[0]) - S0;
        G[0][3 * c + 1] = Real(4) * (u[5] - u[0]) - S1;
        G[0][3 * c + 2] = Real(4) * (u[6] - u[0]) - S2;
        G[1][3 * c + 0] = Real(4) * (u[1] - u[4]) - S0;
        G[1][3 * c + 1] = Real(4) * (u[7] - u[4]) - S1;
        G[1][3 * c + 2] = Real(4) * (u[8] - u[4]) - S2;
        G[2][3 * c + 0] = Real(4) * (u[7] - u[5]) - S0;
        G[2][3 * c + 1] = Real(4) * (u[2] - u[5]) - S1;
        G[2][3 * c + 2] = Real(4) * (u[9] - u[5]) - S2;
        G[3][3 * c + 0] = Real(4) * (u[8] - u[6]) - S0;
        G[3][3 * c + 1] = Real(4) * (u[9] - u[6]) - S1;
        G[3][3 * c + 2] = Real(4) * (u[3] - u[6]) - S2;
    }
}

// r_i = sum_v M_v : dphi_i(vertex v) for the P2 basis (reference gradients linear in the
// barycentric coordinates): vertex a: (4 M_a - sum M) . e_a; edge (a, b): 4 (M_a . e_b + M_b . e_a),
// with e_0 = (-1, -1, -1), e_k = unit k (k = 1, 2, 3).  SM = sum_v M_v.
template <typename Real>
__host__ __device__ inline void p2_vertex_contract(const Real M[4][9], const Real SM[9], Real re[3][10]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Real Me0[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) Me0[v] = -(M[v][3 * c] + M[v][3 * c + 1] + M[v][3 * c + 2]);
        const Real SMe0 = -(SM[3 * c] + SM[3 * c + 1] + SM[3 * c + 2]);
        re[c][0] = Real(4) * Me0[0] - SMe0;
        re[c][1] = Real(4) * M[1][3 * c + 0] - SM[3 * c + 0];
        re[c][2] = Real(4) * M[2][3 * c + 1] - SM[3 * c + 1];
        re[c][3] = Real(4) * M[3][3 * c + 2] - SM[3 * c + 2];
        re[c][4] = Real(4) * (M[0][3 * c + 0] + Me0[1]);           // (0,1)
        re[c][5] = Real(4) * (M[0][3 * c + 1] + Me0[2]);           // (0,2)
        re[c][6] = Real(4) * (M[0][3 * c + 2] + Me0[3]);           // (0,3)
        re[c][7] = Real(4) * (M[1][3 * c + 1] + M[2][3 * c + 0]);  // (1,2)
        re[c][8] = Real(4) * (M[1][3 * c + 2] + M[3][3 * c + 0]);  // (1,3)
        re[c][9] = Real(4) * (M[2][3 * c + 2] + M[3][3 * c + 1]);  // (2,3)
    }
}

// Element vector of one P2 tet (affine map, 14-point rule): re[c][b], with the stress evaluated at
// every quadrature point (any small-strain constitutive law sigma(eps) fits in place of the
// isotropic one).  On an affine tet the P2 gradients are linear: the strain at a point is
// interpolated from the strains at the four vertices, eps(q) = sum_v l_v(q) eps_v, and the test
// sum collects vertex moments of the stress, r_i = sum_v M_v : dphi_i(v) with
// M_v = -det (sum_q w_q l_v(q) sigma(q)) Jinv^T -- the same quadrature sum as through the basis
// tables (identical up to rounding), with 6 + 6 products per point and vertex instead of 90 + 90.
// Tab::w(q) and Tab::lam(q, v) give the rule's weights and barycentric coordinates.
template <typename Real, typename Tab>
__host__ __device__ inline void p2_element(const double X[4][3], const Real U[3][10], Real lam,
                                           Real mu, Real re[3][10]) {
    double J[9], Jid[9];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
    const double det = inv3(J, Jid);
    Real Ji[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ji[i] = (Real)Jid[i];
    // strains at the vertices (xx, yy, zz, xy, yz, xz) from the physical vertex gradients
    Real ev[4][6];
    {
        struct Acc {
            const Real (*U)[10];
            __host__ __device__ Real operator()(int c, int b) const { return U[c][b]; }
        };
        Real Gv[4][9];
        p2_vertex_grads<Real>(Acc{U}, Gv);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Real g[9];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    g[3 * c + d] = Gv[v][3 * c + 0] * Ji[0 * 3 + d] + Gv[v][3 * c + 1] * Ji[1 * 3 + d] +
                                   Gv[v][3 * c + 2] * Ji[2 * 3 + d];
            ev[v][0] = g[0];
            ev[v][1] = g[4];
            ev[v][2] = g[8];
            ev[v][3] = Real(0.5) * (g[1] + g[3]);
            ev[v][4] = Real(0.5) * (g[5] + g[7]);
            ev[v][5] = Real(0.5) * (g[2] + g[6]);
        }
    }
    Real Ms[4][6];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < 6; ++i) Ms[v][i] = Real(0);
#pragma unroll 2
    for (int q = 0; q < kP2Quad; ++q) {
        const Real l0 = (Real)Tab::lam(q, 0), l1 = (Real)Tab::lam(q, 1), l2 = (Real)Tab::lam(q, 2),
                   l3 = (Real)Tab::lam(q, 3), w = (Real)Tab::w(q);
        Real e[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) e[i] = l0 * ev[0][i] + l1 * ev[1][i] + l2 * ev[2][i] + l3 * ev[3][i];
        // pointwise stress (isotropic linear elasticity), weighted
        const Real ls = w * lam * (e[0] + e[1] + e[2]), m2 = w * Real(2) * mu;
        Real sg[6];
        sg[0] = ls + m2 * e[0];
        sg[1] = ls + m2 * e[1];
        sg[2] = ls + m2 * e[2];
        sg[3] = m2 * e[3];
        sg[4] = m2 * e[4];
        sg[5] = m2 * e[5];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            Ms[0][i] += l0 * sg[i];
            Ms[1][i] += l1 * sg[i];
            Ms[2][i] += l2 * sg[i];
            Ms[3][i] += l3 * sg[i];
        }
    }
    // M_v = -det S_v Jinv^T (S_v the symmetric stress moment), then the vertex contraction
    Real M[4][9], SM[9];
    const Real nd = -(Real)det;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const Real S[9] = {Ms[v][0], Ms[v][3], Ms[v][5], Ms[v][3], Ms[v][1], Ms[v][4], Ms[v][5], Ms[v][4], Ms[v][2]};
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                M[v][3 * c + k] = nd * (S[3 * c + 0] * Ji[3 * k + 0] + S[3 * c + 1] * Ji[3 * k + 1] +
                                        S[3 * c + 2] * Ji[3 * k + 2]);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) SM[i] = M[0][i] + M[1][i] + M[2][i] + M[3][i];
    p2_vertex_contract<Real>(M, SM, re);
}

// --------------------------------------------------------------------------------------------
// Factorized element paths for pointwise kernels that are linear in the solution gradient with
// a per-cell constant coefficient (isotropic linear elasticity: f1 = -C : grad u and its Jf3),
// on affine cells.  With P(q) = L(G(q)) linear, the quadrature sum
//     r_i = sum_q w_q P(q) : dphi_i(q)
// is evaluated exactly through the structure of the basis: the pointwise map is applied to the
// polynomial components of the gradient instead of to its values at the points.  The result is
// the same quadrature sum (identical up to rounding), at a fraction of the arithmetic.
// --------------------------------------------------------------------------------------------

// Solution values of one cell read on demand (component c, local basis b).
template <typename Real>
struct ArrayAcc3 {
    const Real (*U)[10];
    __host__ __device__ Real operator()(int c, int b) const { return U[c][b]; }
};
template <typename Real>
struct ArrayAcc3Q1 {
    const Real (*U)[8];
    __host__ __device__ Real operator()(int c, int b) const { return U[c][b]; }
};
struct ArrayAccX {
    const double (*X)[8];
    __host__ __device__ double operator()(int d, int l) const { return X[d][l]; }
};

// Walsh-Hadamard transform of 8 lexicographic corner values: T[b] for b = bxi + 2 beta + 4 bzeta,
// T[b] = sum_l F[l] prod(sign of the odd coordinates of corner l).
template <typename Real>
__host__ __device__ inline void wht8(const Real F[8], Real T[8]) {
    Real a[8];
#pragma unroll
    for (int p = 0; p < 4; ++p) {  // xi
        a[2 * p] = F[2 * p] + F[2 * p + 1];
        a[2 * p + 1] = F[2 * p + 1] - F[2 * p];
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)  // eta
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            T[i + 4 * k] = a[i + 4 * k] + a[i + 2 + 4 * k];
            T[i + 2 + 4 * k] = a[i + 2 + 4 * k] - a[i + 4 * k];
        }
#pragma unroll
    for (int i = 0; i < 4; ++i) {  // zeta
        a[i] = T[i] + T[i + 4];
        a[i + 4] = T[i + 4] - T[i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) T[i] = a[i];
}

// Inverse structure: nodal values r[l] = sum_{b != 0} R[b] prod(signs of corner l in b).
template <typename Real>
__host__ __device__ inline void iwht8(const Real R[8], Real r[8]) {
    Real a[8], t[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {  // zeta
        a[i] = R[i] - R[i + 4];
        a[i + 4] = R[i] + R[i + 4];
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)  // eta
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            t[i + 4 * k] = a[i + 4 * k] - a[i + 2 + 4 * k];
            t[i + 2 + 4 * k] = a[i + 4 * k] + a[i + 2 + 4 * k];
        }
#pragma unroll
    for (int p = 0; p < 4; ++p) {  // xi
        r[2 * p] = t[2 * p] - t[2 * p + 1];
        r[2 * p + 1] = t[2 * p] + t[2 * p + 1];
    }
}

// Affine test of a Q1 cell from the coordinate transform: the bilinear and trilinear modes of
// the corner coordinates vanish (relative to the cell size).
__host__ __device__ inline bool q1_affine(const double TX[3][8]) {
    double size = 0.0, higher = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        size = fmax(size, fmax(fabs(TX[d][1]), fmax(fabs(TX[d][2]), fabs(TX[d][4]))));
        higher = fmax(higher, fmax(fmax(fabs(TX[d][3]), fabs(TX[d][5])), fmax(fabs(TX[d][6]), fabs(TX[d][7]))));
    }
    return higher <= 1e-13 * size;
}

// Q1 affine element: modal form of the 2x2x2 Gauss sum (Gauss orthogonality of the modes:
// sum_q w_q m(q) m'(q) = 8 delta nu_m, nu = 1, 1/3, 1/9 for constant, linear, bilinear modes).
// re[c][l] in lexicographic corner order.  TX: WHT of the corner coordinates (affine cell).
// Constant Jacobian of an affine Q1 cell (a parallelepiped: every mode above the linear ones
// vanishes) from the edge vectors at lexicographic corner 0: J[:, k] = (x(e_k) - x(0)) / 2 with
// e = 1 (xi), 2 (eta), 4 (zeta).  The plan routes only cells that pass q1_affine here.
template <typename XAcc>
__host__ __device__ inline void q1_affine_jacobian(const XAcc &X, double J[9]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double x0 = X(d, 0);
        J[3 * d + 0] = 0.5 * (X(d, 1) - x0);
        J[3 * d + 1] = 0.5 * (X(d, 2) - x0);
        J[3 * d + 2] = 0.5 * (X(d, 4) - x0);
    }
}

template <typename Real, typename UAcc>
__host__ __device__ inline void q1_element_affine(const double J[9], const UAcc &U, Real lam, Real mu,
                                                  Real re[3][8]) {
    double Jid[9];
    const double det = inv3(J, Jid);
    Real Ji[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ji[i] = (Real)Jid[i];
    // modal coefficients of u times 8: T[c][b]
    Real T[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Real u[8];
#pragma unroll
        for (int l = 0; l < 8; ++l) u[l] = U(c, l);
        wht8<Real>(u, T[c]);
    }
    // Gradient modes (times 8), then fluxes scaled by det * nu_m / 8:
    //   const: [T1, T2, T4];  xi: [0, T3, T5];  eta: [T3, 0, T6];  zeta: [T5, T6, 0];
    //   eta zeta: [T7, 0, 0]; xi zeta: [0, T7, 0]; xi eta: [0, 0, T7]  (T3 = xi eta, T5 = xi zeta, T6 = eta zeta)
    // Only the nonzero gradient columns enter and only the flux columns a mode feeds are formed.
    Real R[3][8];
    // nu / 8 as constants: an FP64 division costs about a dozen instructions per cell
    constexpr double c1 = 0.125 / 3.0, c2 = 0.125 / 9.0;
    const Real s0 = (Real)(0.125 * det), s1 = (Real)(c1 * det), s2 = (Real)(c2 * det);
    Real G[9], P[9];
    // constant mode -> R_xi (b=1), R_eta (b=2), R_zeta (b=4): column k of P
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        G[3 * c + 0] = T[c][1];
        G[3 * c + 1] = T[c][2];
        G[3 * c + 2] = T[c][4];
    }
    iso_flux_cols<Real, 7, 7>(G, Ji, lam, mu, s0, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        R[c][0] = Real(0);
        R[c][1] = P[3 * c + 0];
        R[c][2] = P[3 * c + 1];
        R[c][4] = P[3 * c + 2];
    }
    // xi mode: columns eta, zeta of G; P_eta -> R(xi eta) b=3, P_zeta -> R(xi zeta) b=5
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        G[3 * c + 1] = T[c][3];
        G[3 * c + 2] = T[c][5];
    }
    iso_flux_cols<Real, 6, 6>(G, Ji, lam, mu, s1, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        R[c][3] = P[3 * c + 1];
        R[c][5] = P[3 * c + 2];
    }
    // eta mode: columns xi, zeta; P_xi -> b=3, P_zeta -> R(eta zeta) b=6
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        G[3 * c + 0] = T[c][3];
        G[3 * c + 2] = T[c][6];
    }
    iso_flux_cols<Real, 5, 5>(G, Ji, lam, mu, s1, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        R[c][3] += P[3 * c + 0];
        R[c][6] = P[3 * c + 2];
    }
    // zeta mode: columns xi, eta; P_xi -> b=5, P_eta -> b=6
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        G[3 * c + 0] = T[c][5];
        G[3 * c + 1] = T[c][6];
    }
    iso_flux_cols<Real, 3, 3>(G, Ji, lam, mu, s1, P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        R[c][5] += P[3 * c + 0];
        R[c][6] += P[3 * c + 1];
    }
    // the three bilinear modes feed R(xi eta zeta) b=7 with one gradient column each, a = T[.][7]
    // times row k of Jinv (b_k): sigma b_k = (lambda + mu)(a . b_k) b_k + mu |b_k|^2 a, so their
    // sum is (lambda + mu) N a + mu tr(N) a with N = sum_k b_k b_k^T (the physical metric)
    Real N[6];  // xx, yy, zz, xy, yz, xz
    N[0] = Ji[0] * Ji[0] + Ji[3] * Ji[3] + Ji[6] * Ji[6];
    N[1] = Ji[1] * Ji[1] + Ji[4] * Ji[4] + Ji[7] * Ji[7];
    N[2] = Ji[2] * Ji[2] + Ji[5] * Ji[5] + Ji[8] * Ji[8];
    N[3] = Ji[0] * Ji[1] + Ji[3] * Ji[4] + Ji[6] * Ji[7];
    N[4] = Ji[1] * Ji[2] + Ji[4] * Ji[5] + Ji[7] * Ji[8];
    N[5] = Ji[0] * Ji[2] + Ji[3] * Ji[5] + Ji[6] * Ji[8];
    const Real trN = N[0] + N[1] + N[2];
    const Real lm = -s2 * (lam + mu), mt = -s2 * mu * trN;
    const Real a0 = T[0][7], a1 = T[1][7], a2 = T[2][7];
    R[0][7] = lm * (N[0] * a0 + N[3] * a1 + N[5] * a2) + mt * a0;
    R[1][7] = lm * (N[3] * a0 + N[1] * a1 + N[4] * a2) + mt * a1;
    R[2][7] = lm * (N[5] * a0 + N[4] * a1 + N[2] * a2) + mt * a2;
#pragma unroll
    for (int c = 0; c < 3; ++c) iwht8<Real>(R[c], re[c]);
}

// Q1 element at the 2x2x2 Gauss points in the modal basis: u = sum_b T_b m_b with T the Walsh-
// Hadamard transform of the corner values over 8 and m_b the products of the reference
// coordinates in b (b = xi + 2 eta + 4 zeta bits), so the reference gradient at a point is
//   du/dxi = T1 + T3 eta + T5 zeta + T7 eta zeta   (and cyclically for eta, zeta),
// and the test sum collects the modal moments R_b = sum_q P(q) : grad m_b(q), re = iwht8(R).
// The pointwise flux is evaluated at every point, with the optional per-point stress `pre` added
// to sigma: six components (xx, yy, zz, xy, yz, xz), component k of point q = qi + 2 qj + 4 qk
// at pre[(8 k + q) * pstride], scaled by prescale (the state-dependent part of a viscoelastic
// stress).  AFFINE: the cell's constant Jacobian (edge vectors at corner 0); otherwise the
// Jacobian at each point from the modal coefficients of the coordinates.  re[c][l] in
// lexicographic corner order.
template <typename Real, bool AFFINE, typename XAcc, typename UAcc>
__host__ __device__ inline void q1_element_qp(const XAcc &X, const UAcc &U, Real lam, Real mu, Real re[3][8],
                                              const double *pre = nullptr, size_t pstride = 0,
                                              Real prescale = Real(0)) {
    constexpr double g = 0.57735026918962576451;  // 1 / sqrt(3)
    Real T[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Real u[8];
#pragma unroll
        for (int l = 0; l < 8; ++l) u[l] = U(c, l);
        wht8<Real>(u, T[c]);
#pragma unroll
        for (int b = 1; b < 8; ++b) T[c][b] *= Real(0.125);
    }
    double TX[3][8], Ja[9], Jia[9], deta = 0.0;
    if constexpr (AFFINE) {
        q1_affine_jacobian(X, Ja);
        deta = inv3(Ja, Jia);
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double x[8];
#pragma unroll
            for (int l = 0; l < 8; ++l) x[l] = X(d, l);
            wht8<double>(x, TX[d]);
#pragma unroll
            for (int b = 1; b < 8; ++b) TX[d][b] *= 0.125;
        }
    }
    Real R[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int b = 0; b < 8; ++b) R[c][b] = Real(0);
#pragma unroll 1
    for (int q = 0; q < 8; ++q) {
        const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g, zeta = (q & 4) ? g : -g;
        double Jid[9], det;
        if constexpr (AFFINE) {
#pragma unroll
            for (int i = 0; i < 9; ++i) Jid[i] = Jia[i];
            det = deta;
        } else {
            double J[9];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                J[3 * d + 0] = TX[d][1] + TX[d][3] * eta + TX[d][5] * zeta + TX[d][7] * (eta * zeta);
                J[3 * d + 1] = TX[d][2] + TX[d][3] * xi + TX[d][6] * zeta + TX[d][7] * (xi * zeta);
                J[3 * d + 2] = TX[d][4] + TX[d][5] * xi + TX[d][6] * eta + TX[d][7] * (xi * eta);
            }
            det = inv3(J, Jid);
        }
        const Real x = (Real)xi, y = (Real)eta, z = (Real)zeta, yz = (Real)(eta * zeta),
                   xz = (Real)(xi * zeta), xy = (Real)(xi * eta);
        Real Ji[9], G[9], P[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) Ji[i] = (Real)Jid[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            G[3 * c + 0] = T[c][1] + T[c][3] * y + T[c][5] * z + T[c][7] * yz;
            G[3 * c + 1] = T[c][2] + T[c][3] * x + T[c][6] * z + T[c][7] * xz;
            G[3 * c + 2] = T[c][4] + T[c][5] * x + T[c][6] * y + T[c][7] * xy;
        }
        // weight 1, and the 1/8 of the test modes
        const Real scale = (Real)(0.125 * det);
        if (pre) {
            Real ps[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) ps[k] = prescale * (Real)pre[(size_t)(8 * k + q) * pstride];
            iso_flux_pre<Real>(G, Ji, lam, mu, scale, ps, P);
        } else {
            iso_flux<Real>(G, Ji, lam, mu, scale, P);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const Real p0 = P[3 * c + 0], p1 = P[3 * c + 1], p2 = P[3 * c + 2];
            R[c][1] += p0;
            R[c][2] += p1;
            R[c][4] += p2;
            R[c][3] += y * p0 + x * p1;
            R[c][5] += z * p0 + x * p2;
            R[c][6] += z * p1 + y * p2;
            R[c][7] += yz * p0 + xz * p1 + xy * p2;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) iwht8<Real>(R[c], re[c]);
}

// P2 affine element, factorized through the P1 structure of the gradient: G(q) = sum_c l_c(q) G_c
// with G_c the reference gradient at vertex c, so P(q) = sum_c l_c(q) P_c and
//   r_i = sum_d M_d : dphi_i(vertex d),  M_d = sum_c W_dc P_c,  W_dc = sum_q w_q l_d(q) l_c(q)
// (W from the 14-point rule: wdiag on the diagonal, woff off it).
template <typename Real, typename UAcc>
__host__ __device__ inline void p2_element_linear(const double X[4][3], const UAcc &U, Real lam,
                                                  Real mu, double wdiag, double woff, Real re[3][10]) {
    double J[9], Jid[9];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
    const double det = inv3(J, Jid);
    Real Ji[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ji[i] = (Real)Jid[i];
    Real P[4][9];
    p2_vertex_grads<Real>(U, P);
    // vertex fluxes, in place
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        Real G[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) G[i] = P[v][i];
        iso_flux<Real>(G, Ji, lam, mu, (Real)det, P[v]);
    }
    // M_d = wdiag' P_d + woff SP with wdiag' = wdiag - woff; sum_d M_d = (wdiag' + 4 woff) SP
    Real SP[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) SP[i] = P[0][i] + P[1][i] + P[2][i] + P[3][i];
    const Real a = (Real)(wdiag - woff), b = (Real)woff;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int i = 0; i < 9; ++i) P[v][i] = a * P[v][i] + b * SP[i];
    const Real sa = a + Real(4) * b;
    Real SM[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) SM[i] = sa * SP[i];
    p2_vertex_contract<Real>(P, SM, re);
}

}  // namespace plx
