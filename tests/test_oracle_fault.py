"""Pins of the oracle's cohesive-cell (prescribed slip) residual: the two-block rigid-motion
solution of tests/fullscale/linearelasticity/faults-3d/twoblocks_soln.py, the closed-form Q1
face mass of the Lagrange-multiplier coupling, and linearity.  No GPU."""
import numpy as np

from pylith_amd import meshgen as mg

F1 = 1


def two_block(oracle, n=(4, 3, 3), shear=None):
    m, f = mg.hex_box_with_fault(n, (-2000.0, 0.0, 0.0), (2000.0, 3000.0, 3000.0), n[0] // 2,
                                 brick=2, shear=shear)
    aux, sizes = mg.elastic_aux(m.num_cells)
    return m, f, oracle.Problem(m, aux, sizes)


def test_two_blocks_rigid_motion_is_an_exact_solution(oracle):
    # -x block fixed, +x block moved by the fault slip: stress-free, zero traction (lambda = 0)
    m, f, P = two_block(oracle)
    nrm, t1, t2 = mg.fault_normal_frame(f, m.coords)
    slip = np.zeros((f.num_lag, 3))
    slip[:, 1] = 2.0  # left-lateral
    d = slip[0, 0] * nrm + slip[0, 1] * t1 + slip[0, 2] * t2
    nu = 3 * m.num_nodes
    x = np.zeros(nu + 3 * f.num_lag)
    pos_side = np.zeros(m.num_nodes, bool)
    for c in range(m.num_cells):
        if m.coords[m.cone[c]].mean(0)[0] > 0:
            pos_side[m.closure[c]] = True
    x[:nu].reshape(-1, 3)[pos_side] = d
    C = oracle.Cohesive(m, f, slip, nu)
    r = P.residual(x[:nu], F1, out=np.zeros(nu))
    scale = 2.25e10 * 2.0 * 1000.0  # mu * slip * h: the size of an unbalanced nodal force
    rfull = np.concatenate([r, np.zeros(3 * f.num_lag)])
    rfull = C.residual(x, out=rfull)
    assert np.abs(rfull).max() < 1e-9 * scale
    # a wrong slip sense leaves the Lagrange rows unbalanced
    C2 = oracle.Cohesive(m, f, -slip, nu)
    assert np.abs(C2.residual(x)[nu:]).max() > 1e3


def test_lagrange_coupling_is_the_face_mass(oracle):
    m, f, P = two_block(oracle)
    nu = 3 * m.num_nodes
    lam = np.array([0.3, -1.2, 2.0])
    x = np.zeros(nu + 3 * f.num_lag)
    x[nu:] = np.tile(lam, f.num_lag)
    C = oracle.Cohesive(m, f, np.zeros((f.num_lag, 3)), nu)
    r = C.residual(x)
    area = 3000.0 * 3000.0
    rn = r[:nu].reshape(-1, 3)[f.neg_node]
    rp = r[:nu].reshape(-1, 3)[f.pos_node]
    assert np.allclose(rn.sum(0), -lam * area, rtol=1e-13)
    assert np.allclose(rp.sum(0), lam * area, rtol=1e-13)
    # interior fault node: consistent Q1 face mass row sum = h^2
    h2 = 1000.0 * 1000.0
    jk = f.lag_jk
    inner = np.nonzero((jk[:, 0] > 0) & (jk[:, 0] < 3) & (jk[:, 1] > 0) & (jk[:, 1] < 3))[0]
    assert np.allclose(rp[inner], lam * h2, rtol=1e-13)
    # a point load on one Lagrange node spreads with the Q1 face mass stencil h^2/36 [4 2 1]
    i = inner[0]
    x[nu:] = 0.0
    x[nu + 3 * i] = 1.0
    r = C.residual(x)
    rp = r[:nu].reshape(-1, 3)[f.pos_node][:, 0]
    dj = np.abs(jk[:, 0] - jk[i, 0])
    dk = np.abs(jk[:, 1] - jk[i, 1])
    w = {(0, 0): 16, (1, 0): 4, (0, 1): 4, (1, 1): 1}
    exp = np.array([w.get((a, b), 0) for a, b in zip(dj, dk)]) * h2 / 36.0
    assert np.allclose(rp, exp, rtol=1e-12, atol=1e-9)


def test_cohesive_jacobian_action_is_linear_part(oracle):
    m, f, P = two_block(oracle, shear=0.4)
    nu = 3 * m.num_nodes
    rng = np.random.default_rng(1)
    v = rng.standard_normal(nu + 3 * f.num_lag)
    C0 = oracle.Cohesive(m, f, np.zeros((f.num_lag, 3)), nu)
    assert np.allclose(C0.jacobian_action(v), C0.residual(v), rtol=0, atol=1e-12 * np.abs(v).max() * 1e6)
    slip = rng.standard_normal((f.num_lag, 3))
    C = oracle.Cohesive(m, f, slip, nu)
    # F(v) - F(0) = J v
    assert np.allclose(C.residual(v) - C.residual(np.zeros_like(v)), C.jacobian_action(v), atol=1e-6)
