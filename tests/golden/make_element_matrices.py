"""Exact element stiffness matrices by symbolic integration (sympy, rational arithmetic).

Writes tests/golden/element_matrices.npz with, for linear isotropic elasticity,
    K = mu * K_mu + lam * K_lam,   K_ij^{fg} = int C_{f a g b} d_a phi_i d_b phi_j dV,
    C_{f a g b} = lam d_fa d_gb + mu (d_fg d_ab + d_fb d_ag),
for
  * a Q1 box cell [0,2]x[0,3]x[0,5] with PyLith's hex vertex order;
  * a P2 tetrahedron with vertices (0,0,0),(2,0,0),(0,3,0),(0,0,5), nodes = 4 vertices then
    edge midpoints of edges (0,1),(0,2),(0,3),(1,2),(1,3),(2,3).
Rows/columns are ordered (node, component).  These are closed forms of the bilinear form,
computed without any quadrature rule.
"""
import os

import numpy as np
import sympy as sp

x, y, z = sp.symbols("x y z")
X = (x, y, z)


def stiffness(phis, integrate):
    nb = len(phis)
    grads = [[sp.diff(p, v) for v in X] for p in phis]
    Kmu = np.zeros((3 * nb, 3 * nb))
    Klam = np.zeros((3 * nb, 3 * nb))
    cache = {}
    for i in range(nb):
        for j in range(nb):
            for a in range(3):
                for b in range(3):
                    cache[(i, j, a, b)] = integrate(sp.expand(grads[i][a] * grads[j][b]))
    for i in range(nb):
        for j in range(nb):
            for f in range(3):
                for g in range(3):
                    # lam part: d_f phi_i d_g phi_j; mu part: d_fg grad phi_i . grad phi_j
                    # + d_g phi_i d_f phi_j
                    lam = cache[(i, j, f, g)]
                    mu = (sum(cache[(i, j, a, a)] for a in range(3)) if f == g else 0) + cache[(i, j, g, f)]
                    Kmu[3 * i + f, 3 * j + g] = float(mu)
                    Klam[3 * i + f, 3 * j + g] = float(lam)
    return Kmu, Klam


def q1_box(a, b, c):
    ref = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    L = lambda s, t, h: (t / h) if s else (1 - t / h)
    phis = [L(i, x, a) * L(j, y, b) * L(k, z, c) for i, j, k in ref]
    integ = lambda e: sp.integrate(e, (x, 0, a), (y, 0, b), (z, 0, c))
    coords = np.array([[i * a, j * b, k * c] for i, j, k in ref], float)
    return stiffness(phis, integ), coords


def p2_tet(a, b, c):
    l1, l2, l3 = x / a, y / b, z / c
    lam = [1 - l1 - l2 - l3, l1, l2, l3]
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    phis = [lam[v] * (2 * lam[v] - 1) for v in range(4)] + [4 * lam[p] * lam[q] for p, q in edges]
    integ = lambda e: sp.integrate(e, (z, 0, c * (1 - x / a - y / b)), (y, 0, b * (1 - x / a)), (x, 0, a))
    verts = np.array([[0, 0, 0], [a, 0, 0], [0, b, 0], [0, 0, c]], float)
    return stiffness(phis, integ), verts


if __name__ == "__main__":
    (q1_mu, q1_lam), q1_x = q1_box(sp.Integer(2), sp.Integer(3), sp.Integer(5))
    (p2_mu, p2_lam), p2_x = p2_tet(sp.Integer(2), sp.Integer(3), sp.Integer(5))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "element_matrices.npz")
    np.savez(out, q1_Kmu=q1_mu, q1_Klam=q1_lam, q1_coords=q1_x, p2_Kmu=p2_mu, p2_Klam=p2_lam,
             p2_verts=p2_x)
    print(out)
