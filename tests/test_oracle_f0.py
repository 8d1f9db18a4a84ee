"""Body force and gravity, the f0 of the quasi-static displacement equation
(Elasticity::g0v_bodyforce / g0v_grav / g0v_gravbodyforce, fekernels/Elasticity.hh:249-385), and
the reference-state stress (IsotropicLinearElasticity3D::f1v_infinitesimalStrain_refState,
IsotropicLinearElasticity.hh:848-880), in the oracle.

Pins:
* the reference's MMS cases Gravity3D and BodyForce3D (tests/mmstests/linearelasticity/
  nofaults-3d/{Gravity3D,BodyForce3D}.cc, TetP2 on its tet.msh): their exact displacements are
  quadratic, inside the P2 space, so every interior residual row vanishes once f0 balances the
  divergence of the stress; without f0, or with f0 of the wrong sign or without the density
  factor, those rows do not vanish;
* closed-form moments of the basis on affine P2 tets (vertex functions integrate to -V/20,
  edge functions to V/5) and the partition of unity on non-affine Q1 hexes (the rows of a
  constant f0 sum to f0 times the volume);
* the reference state: the pointwise stress against sigma_ref + lambda tr(e - e_ref) I +
  2 mu (e - e_ref) written with lambda (the kernel uses K and the deviator), a consistent
  reference state reducing to the plain kernel, and the reference's GravityRefState3D MMS case
  (HexQ1 and TetP2: a lithostatic reference stress balances gravity at u = 0).
"""
import math
import os

import numpy as np
import pytest

from pylith_amd import meshgen as mg
from pylith_amd import meshio_gmsh

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
F1 = 1
F0_BODY, F0_GRAV, F0_GRAVBODY = 7, 8, 9
G_ACC = 9.80665  # utils/constants.hh:17


def f0_aux(base, body=None, grav=None):
    """[density, (body_force), (gravitational_acceleration), shear_modulus, bulk_modulus] per cell
    (materials/Elasticity.cc:234-241 subfield order)."""
    n = base.shape[0]
    cols, sizes = [base[:, :1]], [1]
    if body is not None:
        cols.append(np.broadcast_to(np.asarray(body, float), (n, 3)))
        sizes.append(3)
    if grav is not None:
        cols.append(np.broadcast_to(np.asarray(grav, float), (n, 3)))
        sizes.append(3)
    cols.append(base[:, 1:3])
    sizes += [1, 1]
    return np.ascontiguousarray(np.concatenate(cols, 1)), np.array(sizes, np.int32)


def mms_material(m):
    # Gravity3D.cc / BodyForce3D.cc: density 2500, Vs 3000, Vp = sqrt(3) Vs
    base, _ = mg.elastic_aux(m.num_cells, density=2500.0, vs=3000.0, vp=math.sqrt(3.0) * 3000.0)
    rho, mu, K = base[0]
    return base, rho, K + 4.0 / 3.0 * mu  # lambda + 2 mu


def gravity_disp(x, rho, l2m, zmin=-4.0e3, zmax=4.0e3):
    """Gravity3D.cc disp_z: rho g / (lambda + 2 mu) (0.5 (z^2 - zmin^2) - zmax (z - zmin))."""
    z = x[:, 2]
    uz = rho * G_ACC / l2m * (0.5 * (z * z - zmin * zmin) - zmax * (z - zmin))
    return np.stack([0 * z, 0 * z, uz], 1).ravel()


def bodyforce_disp(x, l2m, b=5.0e3, xmax=4.0e3):
    """BodyForce3D.cc disp_x: -0.5 b / (lambda + 2 mu) (x - xmax)^2."""
    xp = x[:, 0] - xmax
    return np.stack([-0.5 * b / l2m * xp * xp, 0 * xp, 0 * xp], 1).ravel()


def mms_case(kind):
    m, groups = meshio_gmsh.read(os.path.join(GOLDEN, "mms_tet.msh"))
    base, rho, l2m = mms_material(m)
    x = m.node_coords
    grav = (0.0, 0.0, -G_ACC)
    body = (5.0e3, 0.0, 0.0)
    if kind == F0_GRAV:
        aux = f0_aux(base, grav=grav)
        u = gravity_disp(x, rho, l2m)
    elif kind == F0_BODY:
        aux = f0_aux(base, body=body)
        u = bodyforce_disp(x, l2m)
    else:
        aux = f0_aux(base, body=body, grav=grav)
        u = gravity_disp(x, rho, l2m) + bodyforce_disp(x, l2m)
    interior = np.setdiff1d(np.arange(m.num_nodes), groups["boundary"])
    return m, aux, u, interior


@pytest.mark.parametrize("kind", [F0_GRAV, F0_BODY, F0_GRAVBODY])
def test_mms_interior_rows_vanish(oracle, kind):
    m, (aux, sizes), u, interior = mms_case(kind)
    P = oracle.Problem(m, aux, sizes)
    r = P.residual(u, F1, f0=kind).reshape(-1, 3)
    load = P.residual(np.zeros_like(u), 0, f0=kind).reshape(-1, 3)
    scale = np.abs(load[interior]).max()
    assert scale > 0.0
    assert np.abs(r[interior]).max() <= 1e-10 * scale
    # the check can fail: without f0 the interior rows carry minus the load
    r0 = P.residual(u, F1).reshape(-1, 3)
    assert np.abs(r0[interior]).max() > 0.5 * scale


def test_gravity_sign_and_density_matter(oracle):
    """A density of 1 or a reversed gravity vector leaves the interior rows unbalanced."""
    m, (aux, sizes), u, interior = mms_case(F0_GRAV)
    P = oracle.Problem(m, aux, sizes)
    scale = np.abs(P.residual(np.zeros_like(u), 0, f0=F0_GRAV).reshape(-1, 3)[interior]).max()
    for change in ("density", "sign"):
        a = aux.copy()
        if change == "density":
            a[:, 0] = 1.0
        else:
            a[:, 1:4] *= -1.0
        r = oracle.Problem(m, a, sizes).residual(u, F1, f0=F0_GRAV).reshape(-1, 3)
        assert np.abs(r[interior]).max() > 0.5 * scale, change


def test_p2_moments_affine_tet(oracle):
    """One affine P2 tet, constant f0: vertex rows -V/20 f0, edge rows V/5 f0."""
    rng = np.random.default_rng(5)
    X = np.array([[0, 0, 0], [1.3, 0.1, -0.2], [0.2, 0.9, 0.1], [0.1, -0.3, 1.1]]) * 1e3
    X += rng.uniform(-50, 50, X.shape)
    edges = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))
    node_coords = np.concatenate([X, [(X[a] + X[b]) / 2 for a, b in edges]])
    closure = np.arange(10, dtype=np.int32)[None]
    m = mg.Mesh(mg.TET_P2, closure, closure[:, :4].copy(), X, node_coords, None, np.zeros(1, np.int32))
    vol = np.linalg.det(np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6
    assert vol > 0
    base = np.array([[2700.0, 3.0e10, 5.0e10]])
    body, grav = np.array([1.5, -2.0, 0.25]), np.array([0.3, 0.1, -9.8])
    for kind, aux in ((F0_BODY, f0_aux(base, body=body)), (F0_GRAV, f0_aux(base, grav=grav)),
                      (F0_GRAVBODY, f0_aux(base, body=body, grav=grav))):
        f = {F0_BODY: body, F0_GRAV: 2700.0 * grav, F0_GRAVBODY: body + 2700.0 * grav}[kind]
        r = oracle.Problem(m, *aux).residual(np.zeros(30), 0, f0=kind).reshape(10, 3)
        np.testing.assert_allclose(r[:4], np.tile(-vol / 20 * f, (4, 1)), rtol=1e-12, atol=1e-12 * abs(vol) * np.abs(f).max())
        np.testing.assert_allclose(r[4:], np.tile(vol / 5 * f, (6, 1)), rtol=1e-12)


def test_q1_partition_of_unity_nonaffine(oracle):
    """Non-affine hexes, per-cell f0: the rows of each component sum to sum_c V_c f0_c."""
    from test_meshio_gmsh import _volume
    m = mg.hex_box((5, 4, 3), (0, 0, 0), (500.0, 400.0, 300.0))
    rng = np.random.default_rng(8)
    x = m.coords.copy()
    inner = np.all((x > 1) & (x < np.array([499, 399, 299])), axis=1)
    x[inner] += rng.uniform(-20, 20, (inner.sum(), 3))
    m.coords = m.node_coords = x
    base, _ = mg.elastic_aux(m.num_cells)
    base[:, 0] = rng.uniform(2000, 3000, m.num_cells)
    body = rng.normal(size=(m.num_cells, 3))
    grav = rng.normal(size=(m.num_cells, 3))
    aux, sizes = f0_aux(base, body=body, grav=grav)
    r = oracle.Problem(m, aux, sizes).residual(np.zeros(3 * m.num_nodes), 0, f0=F0_GRAVBODY).reshape(-1, 3)
    vols = np.array([_volume(mg.Mesh(m.cell_type, m.closure[c:c + 1], m.cone[c:c + 1], m.coords,
                                     m.node_coords, None, m.labels[c:c + 1])) for c in range(m.num_cells)])
    f = body + base[:, :1] * grav
    np.testing.assert_allclose(r.sum(0), (vols[:, None] * f).sum(0), rtol=1e-12)


# --- reference stress / strain (IsotropicLinearElasticity3D::f1v_infinitesimalStrain_refState) ---
F1_REF = 10


def qp_coords(m, oracle):
    """Physical quadrature points [num_cells, nq, 3] of the oracle's rule (affine P2 geometry from
    the vertices, trilinear Q1 geometry)."""
    p, _ = oracle.quadrature(m.cell_type)
    if m.cell_type == mg.TET_P2:
        lam = np.stack([1.0 - p.sum(1), p[:, 0], p[:, 1], p[:, 2]], 1)  # [nq, 4]
        return np.einsum("qv,cvd->cqd", lam, m.coords[m.cone])
    psi = np.stack([oracle.tabulate(m.cell_type, xi)[0] for xi in p])  # [nq, 8]
    return np.einsum("qv,cvd->cqd", psi, m.coords[m.closure])


def refstate_aux(base, nq, ref_stress, ref_strain, grav=None):
    """Per quadrature point [density, (gravity), reference_stress(6), reference_strain(6), mu, K]
    (Elasticity.cc:234-241 with IsotropicLinearElasticity.cc:88-93)."""
    n = ref_stress.shape[0]
    b = np.repeat(base, nq, 0)
    cols, sizes = [b[:, :1]], [1]
    if grav is not None:
        cols.append(np.broadcast_to(np.asarray(grav, float), (n, 3)))
        sizes.append(3)
    cols += [ref_stress, ref_strain, b[:, 1:3]]
    sizes += [6, 6, 1, 1]
    return np.ascontiguousarray(np.concatenate(cols, 1)), np.array(sizes, np.int32)


def gravity_refstate_case(oracle, name, zmax=4.0e3):
    """GravityRefState3D.cc: u = 0, gravity (0, 0, -g), reference stress rho g (z - zmax) on the
    diagonal (basis order 1: exact at the quadrature points), reference strain 0."""
    m, groups = meshio_gmsh.read(os.path.join(GOLDEN, name))
    base, rho, _ = mms_material(m)
    xq = qp_coords(m, oracle).reshape(-1, 3)
    nq = oracle.quadrature(m.cell_type)[1].size
    rs = np.zeros((xq.shape[0], 6))
    rs[:, :3] = (rho * G_ACC * (xq[:, 2] - zmax))[:, None]
    aux, sizes = refstate_aux(base, nq, rs, np.zeros_like(rs), grav=(0.0, 0.0, -G_ACC))
    interior = np.setdiff1d(np.arange(m.num_nodes), groups["boundary"])
    return m, aux, sizes, interior


def test_refstate_point_kernel_closed_form(oracle):
    """sigma = sigma_ref + lambda tr(e - e_ref) I + 2 mu (e - e_ref); f1 = -sigma."""
    rng = np.random.default_rng(12)
    mu, K = 3.0e10, 5.0e10
    lam = K - 2.0 / 3.0 * mu
    for _ in range(5):
        g = rng.normal(size=(3, 3)) * 1e-3
        rs, re = rng.normal(size=6) * 1e6, rng.normal(size=6) * 1e-4
        aux = np.concatenate([[2500.0], rs, re, [mu, K]])
        f1 = oracle.point_f1(F1_REF, g.ravel(), aux, [1, 6, 6, 1, 1])[:9].reshape(3, 3)
        e = 0.5 * (g + g.T)
        E = np.array([[re[0], re[3], re[5]], [re[3], re[1], re[4]], [re[5], re[4], re[2]]])
        S = np.array([[rs[0], rs[3], rs[5]], [rs[3], rs[1], rs[4]], [rs[5], rs[4], rs[2]]])
        d = e - E
        sigma = S + lam * np.trace(d) * np.eye(3) + 2.0 * mu * d
        np.testing.assert_allclose(f1, -sigma, rtol=1e-13, atol=1e-13 * np.abs(sigma).max())
        # consistent reference state (sigma_ref = C : e_ref): the plain kernel's stress
        Sc = lam * np.trace(E) * np.eye(3) + 2.0 * mu * E
        rsc = np.array([Sc[0, 0], Sc[1, 1], Sc[2, 2], Sc[0, 1], Sc[1, 2], Sc[0, 2]])
        f1c = oracle.point_f1(F1_REF, g.ravel(), np.concatenate([[2500.0], rsc, re, [mu, K]]), [1, 6, 6, 1, 1])
        f1p = oracle.point_f1(F1, g.ravel(), [2500.0, mu, K], [1, 1, 1])
        np.testing.assert_allclose(f1c[:9], f1p[:9], rtol=1e-9, atol=1e-9 * np.abs(f1p[:9]).max())


@pytest.mark.parametrize("name", ["mms_hex.msh", "mms_tet.msh"])
def test_gravity_refstate_mms(oracle, name):
    """GravityRefState3D (HexQ1 and TetP2 cases): the lithostatic reference stress balances
    gravity at u = 0, so every interior row vanishes; without gravity they do not."""
    m, aux, sizes, interior = gravity_refstate_case(oracle, name)
    P = oracle.Problem(m, aux, sizes, per_qp=True)
    u = np.zeros(3 * m.num_nodes)
    r = P.residual(u, F1_REF, f0=F0_GRAV).reshape(-1, 3)
    rg = P.residual(u, 0, f0=F0_GRAV).reshape(-1, 3)
    scale = np.abs(rg[interior]).max()
    assert scale > 0.0
    assert np.abs(r[interior]).max() <= 1e-10 * scale
    r_nog = P.residual(u, F1_REF).reshape(-1, 3)
    assert np.abs(r_nog[interior]).max() > 0.5 * scale
