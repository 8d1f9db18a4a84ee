// plx_element.hpp -- element math of the residual / Jacobian-action apply, compiled for both
// the gfx950 kernels and the host emulation used by the CPU test suite.
#pragma once

#include <hip/hip_runtime.h>

namespace plx {

constexpr int kP2Quad = 14;

struct P2Tables {
    double w[kP2Quad];
    double dphi[kP2Quad][10][3];
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

// --------------------------------------------------------------------------------------------
// Q1 hexahedron, 2x2x2 Gauss, sum factorization.  Lexicographic corner index l = i + 2j + 4k
// maps to PyLith's vertex order through kLex.
// --------------------------------------------------------------------------------------------
__host__ __device__ constexpr int q1_lex(int l) {
    return l == 2 ? 3 : l == 3 ? 2 : l == 6 ? 7 : l == 7 ? 6 : l;
}

// Gradient of a trilinear field F (lexicographic corners) at the 8 Gauss points.  Each
// reference derivative is constant along its own direction, so it has 4 distinct values:
//   dX[a][b]: d/dxi at (eta=qa, zeta=qb); dY[a][b]: d/deta at (xi=qa, zeta=qb);
//   dZ[a][b]: d/dzeta at (xi=qa, eta=qb).
template <typename Real>
__host__ __device__ inline void q1_interp2(const Real e[2][2], Real out[2][2]) {
    const Real ga = Real(0.78867513459481288225);  // N0(-1/sqrt3) = N1(+1/sqrt3)
    const Real gb = Real(0.21132486540518711775);  // N1(-1/sqrt3) = N0(+1/sqrt3)
    Real t[2][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        t[0][k] = ga * e[0][k] + gb * e[1][k];
        t[1][k] = gb * e[0][k] + ga * e[1][k];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        out[a][0] = ga * t[a][0] + gb * t[a][1];
        out[a][1] = gb * t[a][0] + ga * t[a][1];
    }
}

template <typename Real>
__host__ __device__ inline void q1_ref_grad(const Real F[8], Real dX[2][2], Real dY[2][2],
                                            Real dZ[2][2]) {
    Real e[2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) e[j][k] = Real(0.5) * (F[1 + 2 * j + 4 * k] - F[2 * j + 4 * k]);
    q1_interp2<Real>(e, dX);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int k = 0; k < 2; ++k) e[i][k] = Real(0.5) * (F[i + 2 + 4 * k] - F[i + 4 * k]);
    q1_interp2<Real>(e, dY);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) e[i][j] = Real(0.5) * (F[i + 2 * j + 4] - F[i + 2 * j]);
    q1_interp2<Real>(e, dZ);
}

// Transpose of q1_ref_grad: R[l] += sum_q sum_k S_k(q) dphi_l/dxi_k(q), given the flux sums
// SX[a][b] = sum over the xi-points of P_xi at (eta=a, zeta=b), and likewise SY, SZ.
// (q1_interp2 is symmetric, so it is its own transpose.)
template <typename Real>
__host__ __device__ inline void q1_ref_grad_t(const Real SX[2][2], const Real SY[2][2],
                                              const Real SZ[2][2], Real R[8]) {
    Real h[2][2];
    q1_interp2<Real>(SX, h);
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            R[1 + 2 * j + 4 * k] += Real(0.5) * h[j][k];
            R[2 * j + 4 * k] -= Real(0.5) * h[j][k];
        }
    q1_interp2<Real>(SY, h);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            R[i + 2 + 4 * k] += Real(0.5) * h[i][k];
            R[i + 4 * k] -= Real(0.5) * h[i][k];
        }
    q1_interp2<Real>(SZ, h);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            R[i + 2 * j + 4] += Real(0.5) * h[i][j];
            R[i + 2 * j] -= Real(0.5) * h[i][j];
        }
}

// Element vector of one Q1 cell (lexicographic corners): re[c][l].
template <typename Real>
__host__ __device__ inline void q1_element(const double X[3][8], const Real U[3][8], Real lam,
                                           Real mu, Real re[3][8]) {
    double jX[3][2][2], jY[3][2][2], jZ[3][2][2];
#pragma unroll
    for (int d = 0; d < 3; ++d) q1_ref_grad<double>(X[d], jX[d], jY[d], jZ[d]);
    Real uX[3][2][2], uY[3][2][2], uZ[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c) q1_ref_grad<Real>(U[c], uX[c], uY[c], uZ[c]);
    Real SX[3][2][2], SY[3][2][2], SZ[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) SX[c][a][b] = SY[c][a][b] = SZ[c][a][b] = Real(0);
#pragma unroll
    for (int qk = 0; qk < 2; ++qk)
#pragma unroll
        for (int qj = 0; qj < 2; ++qj)
#pragma unroll
            for (int qi = 0; qi < 2; ++qi) {
                double J[9], Jid[9];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    J[3 * d + 0] = jX[d][qj][qk];
                    J[3 * d + 1] = jY[d][qi][qk];
                    J[3 * d + 2] = jZ[d][qi][qj];
                }
                const double det = inv3(J, Jid);
                Real Ji[9], G[9], P[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) Ji[i] = (Real)Jid[i];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    G[3 * c + 0] = uX[c][qj][qk];
                    G[3 * c + 1] = uY[c][qi][qk];
                    G[3 * c + 2] = uZ[c][qi][qj];
                }
                iso_flux<Real>(G, Ji, lam, mu, (Real)det, P);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    SX[c][qj][qk] += P[3 * c + 0];
                    SY[c][qi][qk] += P[3 * c + 1];
                    SZ[c][qi][qj] += P[3 * c + 2];
                }
            }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int l = 0; l < 8; ++l) re[c][l] = Real(0);
        q1_ref_grad_t<Real>(SX[c], SY[c], SZ[c], re[c]);
    }
}

// Element vector of one P2 tet (affine map, 14-point rule): re[c][b].  Tab::w(q) and
// Tab::dphi(q, b, k) give the reference tables (constant memory on the device).
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
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int b = 0; b < 10; ++b) re[c][b] = Real(0);
#pragma unroll 2
    for (int q = 0; q < kP2Quad; ++q) {
        Real G[9], P[9];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Real acc = Real(0);
#pragma unroll
                for (int b = 0; b < 10; ++b) acc += U[c][b] * (Real)Tab::dphi(q, b, k);
                G[3 * c + k] = acc;
            }
        iso_flux<Real>(G, Ji, lam, mu, (Real)(Tab::w(q) * det), P);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int b = 0; b < 10; ++b)
                re[c][b] += P[3 * c + 0] * (Real)Tab::dphi(q, b, 0) +
                            P[3 * c + 1] * (Real)Tab::dphi(q, b, 1) +
                            P[3 * c + 2] * (Real)Tab::dphi(q, b, 2);
    }
}

}  // namespace plx
