"""Pins of the CPU oracle (oracle/oracle.c) against closed forms, the reference's own fixtures
and exact symbolic integration.  No GPU.

Each check is chosen so that a plausible restatement mistake (dropped term, wrong sign, wrong
index, transposed operand) fails it; none of them re-types the oracle's formula.
"""
import math
import os

import numpy as np
import pytest

from pylith_amd import meshgen as mg
from pylith_amd import meshio_ascii

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
F1 = 1   # IsotropicLinearElasticity3D::f1v_infinitesimalStrain
JF3 = 2  # IsotropicLinearElasticity3D::Jf3vu_infinitesimalStrain


# ------------------------------------------------------------------------------------------
# Quadrature and tabulation
# ------------------------------------------------------------------------------------------
def test_hex_gauss_exact_to_cubic(oracle):
    p, w = oracle.quadrature(mg.HEX_Q1)
    assert p.shape == (8, 3) and abs(w.sum() - 8.0) < 1e-15
    for a in range(4):
        for b in range(4):
            for c in range(4):
                exact = 1.0
                for e in (a, b, c):
                    exact *= 0.0 if e % 2 else 2.0 / (e + 1)
                got = np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c)
                assert abs(got - exact) < 1e-14


def test_tet_rule_exact_to_degree5(oracle):
    p, w = oracle.quadrature(mg.TET_P2)
    assert p.shape == (14, 3) and abs(w.sum() - 1.0 / 6.0) < 1e-16
    for a in range(6):
        for b in range(6 - a):
            for c in range(6 - a - b):
                exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                got = np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c)
                assert abs(got - exact) < 1e-16, (a, b, c)
    # and not degree 6: the rule is the 14-point degree-5 one, not something larger
    assert abs(np.sum(w * p[:, 0] ** 6) - 720.0 / math.factorial(9)) > 1e-8


@pytest.mark.parametrize("ct,nodes", [
    (mg.HEX_Q1, np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                          [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)),
    (mg.TET_P2, np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [0, .5, 0],
                          [0, 0, .5], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]], float)),
])
def test_tabulation_nodal_and_gradients(oracle, ct, nodes):
    for i, xi in enumerate(nodes):
        phi, _ = oracle.tabulate(ct, xi)
        assert np.allclose(phi, np.eye(len(nodes))[i], atol=1e-15)
    rng = np.random.default_rng(0)
    for _ in range(5):
        xi = rng.uniform(0.05, 0.3, 3) if ct == mg.TET_P2 else rng.uniform(-0.9, 0.9, 3)
        phi, dphi = oracle.tabulate(ct, xi)
        assert abs(phi.sum() - 1) < 1e-14 and np.abs(dphi.sum(0)).max() < 1e-14
        h = 1e-6
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            fd = (oracle.tabulate(ct, xi + e)[0] - oracle.tabulate(ct, xi - e)[0]) / (2 * h)
            assert np.allclose(fd, dphi[:, d], atol=1e-8)


# ------------------------------------------------------------------------------------------
# Pointwise kernels: physical definitions of the moduli
# ------------------------------------------------------------------------------------------
RHO, VS, VP = 2500.0, 3000.0, 5292.0


def _aux():
    mu, K = mg.moduli(RHO, VS, VP)
    return np.array([RHO, mu, K]), [1, 1, 1], float(mu), float(K)


def test_f1_rotation_is_stress_free(oracle):
    a, sz, _, _ = _aux()
    W = np.array([[0, 0.3, -0.2], [-0.3, 0, 0.5], [0.2, -0.5, 0]])
    f1 = oracle.point_f1(F1, W.ravel(), a, sz)[:9]
    assert np.abs(f1).max() == 0.0


def test_f1_moduli_definitions(oracle):
    a, sz, mu, K = _aux()
    e = 1e-4
    # uniform dilatation: sigma = 3 K e I (bulk modulus), f1 = -sigma
    f1 = oracle.point_f1(F1, (e * np.eye(3)).ravel(), a, sz)[:9].reshape(3, 3)
    assert np.allclose(f1, -3 * K * e * np.eye(3), rtol=1e-14, atol=0)
    # engineering shear gamma: sigma_xy = mu gamma (shear modulus = rho Vs^2)
    G = np.zeros((3, 3))
    G[0, 1] = e
    f1 = oracle.point_f1(F1, G.ravel(), a, sz)[:9].reshape(3, 3)
    assert np.isclose(f1[0, 1], -RHO * VS ** 2 * e, rtol=1e-14) and f1[0, 1] == f1[1, 0]
    assert np.count_nonzero(f1) == 2
    # uniaxial strain: sigma_xx = rho Vp^2 e (P-wave modulus), sigma_yy = rho (Vp^2 - 2 Vs^2) e
    G = np.zeros((3, 3))
    G[0, 0] = e
    f1 = oracle.point_f1(F1, G.ravel(), a, sz)[:9].reshape(3, 3)
    assert np.isclose(-f1[0, 0], RHO * VP ** 2 * e, rtol=1e-12)
    assert np.isclose(-f1[1, 1], RHO * (VP ** 2 - 2 * VS ** 2) * e, rtol=1e-12)
    assert np.isclose(-f1[2, 2], RHO * (VP ** 2 - 2 * VS ** 2) * e, rtol=1e-12)


def test_jf3_symmetries_and_consistency_with_f1(oracle):
    a, sz, mu, K = _aux()
    J = oracle.point_jf3(JF3, a, sz).reshape(3, 3, 3, 3)  # [f][g][df][dg]
    assert np.count_nonzero(J) == 21
    assert np.array_equal(J, J.transpose(1, 0, 3, 2))  # major symmetry
    assert np.array_equal(J, J.transpose(2, 1, 0, 3))  # minor: swap f <-> df
    rng = np.random.default_rng(1)
    for _ in range(4):
        G = rng.standard_normal((3, 3)) * 1e-3
        f1 = oracle.point_f1(F1, G.ravel(), a, sz)[:9].reshape(3, 3)
        lin = np.einsum("fgab,gb->fa", J, G)
        assert np.allclose(f1, lin, rtol=1e-13, atol=1e-13 * np.abs(f1).max())


# ------------------------------------------------------------------------------------------
# Element matrices vs exact symbolic integration
# ------------------------------------------------------------------------------------------
def _single_cell_problem(oracle, ct, verts, mu, lam):
    nb = 8 if ct == mg.HEX_Q1 else 10
    nv = 8 if ct == mg.HEX_Q1 else 4
    closure = np.arange(nb, dtype=np.int32)[None, :]
    cone = np.arange(nv, dtype=np.int32)[None, :]
    m = mg.Mesh(ct, closure, cone, np.asarray(verts, float), None, (1, 1, 1))
    aux = np.array([[1.0, mu, lam + 2.0 * mu / 3.0]])
    return oracle.Problem(m, aux, [1, 1, 1])


@pytest.mark.parametrize("ct,key", [(mg.HEX_Q1, "q1"), (mg.TET_P2, "p2")])
def test_element_matrix_matches_symbolic(oracle, ct, key):
    d = np.load(os.path.join(GOLDEN, "element_matrices.npz"))
    verts = d["q1_coords"] if key == "q1" else d["p2_verts"]
    mu, lam = 3.0, 2.0
    P = _single_cell_problem(oracle, ct, verts, mu, lam)
    Ke = P.element_matrix(0, JF3)
    exact = -(mu * d[f"{key}_Kmu"] + lam * d[f"{key}_Klam"])  # Jf3 = -C
    assert np.allclose(Ke, exact, rtol=0, atol=1e-13 * np.abs(exact).max())


# ------------------------------------------------------------------------------------------
# Assembled residual: rigid modes, uniform strain (MMS), energy, Jacobian action
# ------------------------------------------------------------------------------------------
def _perturbed_hex(n=3, seed=3):
    m = mg.hex_box(n, (0, 0, 0), (3000, 3000, 3000), brick=2)
    rng = np.random.default_rng(seed)
    x = m.coords.copy()
    interior = np.all((x > 1) & (x < 2999), axis=1)
    x[interior] += rng.uniform(-150, 150, (interior.sum(), 3))
    m.coords = x
    m.node_coords = x
    return m


def _meshes():
    ref, _ = meshio_ascii.read(os.path.join(GOLDEN, "hex.mesh"))
    return {
        "hex_ref_fixture": ref,
        "hex_perturbed": _perturbed_hex(),
        "tet_p2": mg.tet_p2_box(2, (0, 0, 0), (4000, 4000, 4000), brick=1),
    }


def _problem(oracle, m):
    aux, sizes = mg.elastic_aux(m.num_cells)
    return oracle.Problem(m, aux, sizes)


def _linear_field(m, A, b=np.zeros(3)):
    return (m.node_coords @ A.T + b).ravel()


@pytest.mark.parametrize("name", ["hex_ref_fixture", "hex_perturbed", "tet_p2"])
def test_rigid_body_modes_give_zero_residual(oracle, name):
    m = _meshes()[name]
    P = _problem(oracle, m)
    scale = np.abs(P.residual(_linear_field(m, np.diag([1e-4, 0, 0])), F1)).max()
    W = np.array([[0, 2e-4, -1e-4], [-2e-4, 0, 3e-4], [1e-4, -3e-4, 0]])
    r = P.residual(_linear_field(m, W, np.array([0.3, -0.1, 0.2])), F1)
    assert np.abs(r).max() < 1e-12 * scale


def _interior_nodes(m):
    x = m.node_coords
    lo, hi = x.min(0), x.max(0)
    tol = 1e-9 * np.abs(hi - lo).max()
    return np.nonzero(np.all((x > lo + tol) & (x < hi - tol), axis=1))[0]


# Strains of tests/mmstests/linearelasticity/nofaults-3d/UniformStrain3D.cc:64-122, including
# the reference's quirk: strain_yz() returns PETSC_SUCCESS (0) and disp_y/disp_z multiply
# instead of add, so disp = (0.1x + 0.3y - 0.08z, 0.3x, -0.08x).
MMS_DISP = np.array([[0.1, 0.3, -0.08], [0.3, 0.0, 0.0], [-0.08, 0.0, 0.0]])


@pytest.mark.parametrize("name", ["hex_ref_fixture", "hex_perturbed", "tet_p2"])
def test_uniform_strain_mms_interior_residual_vanishes(oracle, name):
    m = _meshes()[name]
    # UniformStrain3D material: Vs = 3000, Vp = sqrt(3) Vs, density 2500 (UniformStrain3D.cc:30-60)
    aux, sizes = mg.elastic_aux(m.num_cells, 2500.0, 3000.0, math.sqrt(3.0) * 3000.0)
    P = oracle.Problem(m, aux, sizes)
    r = P.residual(_linear_field(m, MMS_DISP), F1).reshape(-1, 3)
    inner = _interior_nodes(m)
    assert inner.size > 0
    bnd = np.setdiff1d(np.arange(m.num_nodes), inner)
    scale = np.abs(r[bnd]).max()
    assert scale > 0
    # TestLinearElasticity.cc:150 uses 1e-9 on the (nondimensional) residual norm
    assert np.abs(r[inner]).max() < 1e-12 * scale


@pytest.mark.parametrize("name", ["hex_ref_fixture", "hex_perturbed", "tet_p2"])
def test_uniform_strain_energy(oracle, name):
    m = _meshes()[name]
    aux, sizes = mg.elastic_aux(m.num_cells)
    P = oracle.Problem(m, aux, sizes)
    E = np.array([[2e-4, 1e-4, 0], [0, -1e-4, 3e-4], [5e-5, 0, 1e-4]])
    u = _linear_field(m, E)
    r = P.residual(u, F1)
    eps = 0.5 * (E + E.T)
    mu, K = aux[0, 1], aux[0, 2]
    lam = K - 2 * mu / 3
    x = m.coords
    vol = np.prod(x.max(0) - x.min(0))
    energy = vol * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
    assert np.isclose(-u @ r, energy, rtol=1e-12)


@pytest.mark.parametrize("name", ["hex_ref_fixture", "hex_perturbed", "tet_p2"])
def test_jacobian_action_equals_residual_and_dense_assembly(oracle, name):
    m = _meshes()[name]
    P = _problem(oracle, m)
    rng = np.random.default_rng(7)
    v = rng.uniform(-1e-3, 1e-3, 3 * m.num_nodes)
    Jv = P.jacobian_action(v, JF3)
    r = P.residual(v, F1)  # linear: F(v) - F(0) = J v
    assert np.allclose(Jv, r, rtol=0, atol=1e-12 * np.abs(r).max())
    # brute force: dense assembly of element matrices
    nd = 3 * m.num_nodes
    A = np.zeros((nd, nd))
    for c in range(m.num_cells):
        Ke = P.element_matrix(c, JF3)
        dofs = (3 * m.closure[c][:, None] + np.arange(3)).ravel()
        A[np.ix_(dofs, dofs)] += Ke
    assert np.allclose(A, A.T, atol=1e-12 * np.abs(A).max())
    assert np.allclose(A @ v, r, rtol=0, atol=1e-12 * np.abs(r).max())
    ev = np.linalg.eigvalsh(A)
    assert np.sum(np.abs(ev) < 1e-9 * np.abs(ev).max()) == 6  # rigid modes only
    assert ev.max() < 1e-9 * np.abs(ev).max()                  # Jf3 = -C: negative semidefinite
    # point-block diagonal equals the assembled diagonal blocks
    B = P.pbdiag(JF3, m.num_nodes)
    for n in range(m.num_nodes):
        assert np.allclose(B[n], A[3 * n:3 * n + 3, 3 * n:3 * n + 3], atol=1e-12 * np.abs(A).max())


def test_colored_residual_matches_serial(oracle):
    m = mg.hex_box(6, (0, 0, 0), (6000, 6000, 6000), brick=2)
    P = _problem(oracle, m)
    u = mg.random_field(m.num_nodes, 5)
    a = P.residual(u, F1)
    b = P.residual_colored(u, F1, m.colors, nthreads=4)
    assert np.allclose(a, b, rtol=0, atol=1e-12 * np.abs(a).max())
    t = mg.tet_p2_box(2, (0, 0, 0), (1, 1, 1), brick=1)
    P = _problem(oracle, t)
    u = mg.random_field(t.num_nodes, 6)
    a = P.residual(u, F1)
    b = P.residual_colored(u, F1, t.colors, nthreads=4)
    assert np.allclose(a, b, rtol=0, atol=1e-12 * np.abs(a).max())


def test_local_to_global_drops_constrained(oracle):
    loc = np.arange(12, dtype=float)
    g = oracle.local_to_global(loc, [0, 5, 11])
    assert np.array_equal(g, np.delete(loc, [0, 5, 11]))


def test_tangential_directions(oracle):
    # vertical fault, normal +x: strike along +y, up-dip +z (FaultCohesive.py ref_dir defaults)
    t1, t2 = oracle.tangential_directions([0, 0, 1], [0, 1, 0], [1, 0, 0])
    assert np.allclose(t1, [0, 1, 0]) and np.allclose(t2, [0, 0, 1])
    rng = np.random.default_rng(2)
    for _ in range(10):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        t1, t2 = oracle.tangential_directions([0, 0, 1], [0, 1, 0], n)
        M = np.stack([n, t1, t2])
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-14)
    # normal along the first reference direction switches to the second
    t1, t2 = oracle.tangential_directions([0, 0, 1], [0, 1, 0], [0, 0, 1])
    assert np.allclose(np.abs(t1), [1, 0, 0])


@pytest.mark.parametrize("kind", ["q1", "p2"])
def test_time_derivative_enters_the_test_f0(oracle, kind):
    """The test-only f0 adds a0 * s_t (a0 the first aux subfield): a uniform unit u_t in one
    component changes that component's residual by sum_i int a0 phi_i = sum_cells a0_c vol_c
    (partition of unity), and nothing else; Jf0 carries s_tshift * a0 on its diagonal."""
    if kind == "q1":
        m = mg.hex_box((3, 2, 2), (0, 0, 0), (3000.0, 2000.0, 2000.0), brick=2, shear=0.2)
    else:
        m = mg.tet_p2_box(2, (0, 0, 0), (2000.0, 2000.0, 2000.0), brick=1)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=3, vs_sigma=0.2, vp_vs=1.8)
    aux[:, 0] *= np.random.default_rng(5).uniform(0.5, 1.5, m.num_cells)  # density varies by cell
    consts = [2.5e-3, 3.0e7, 0.3]
    P = oracle.Problem(m, aux, sizes, constants=consts)
    # congruent cells (a sheared brick grid / the six tets of each cube): equal volumes
    vol = np.full(m.num_cells, (3000.0 if kind == "q1" else 2000.0) * 2000.0 * 2000.0 / m.num_cells)
    n = 3 * m.num_nodes
    u = np.random.default_rng(4).standard_normal(n) * 1e-3
    T_F0, T_F1, T_JF0 = 90, 91, 92
    r0 = P.residual(u, T_F1, f0=T_F0, t=0.4)
    for comp in range(3):
        ut = np.zeros(n)
        ut[comp::3] = 1.0
        d = P.residual(u, T_F1, f0=T_F0, t=0.4, locX_t=ut) - r0
        assert abs(d[comp::3].sum() - (aux[:, 0] * vol).sum()) <= 1e-10 * (aux[:, 0] * vol).sum()
        others = [k for k in range(3) if k != comp]
        assert np.abs(d[others[0]::3]).max() == 0.0 and np.abs(d[others[1]::3]).max() == 0.0
    # the s_tshift term of Jf0: J(s_tshift) - J(0) applied to a uniform field, summed, = s_tshift int a0
    e = np.zeros(n)
    e[0::3] = 1.0
    dj = P.jacobian_action(e, 0, jf0=T_JF0, t=0.4, s_tshift=2.0) - P.jacobian_action(e, 0, jf0=T_JF0, t=0.4)
    assert abs(dj[0::3].sum() - 2.0 * (aux[:, 0] * vol).sum()) <= 1e-10 * (aux[:, 0] * vol).sum()
