"""Body force and gravity on the device (f0 = Elasticity::g0v_bodyforce / g0v_grav /
g0v_gravbodyforce, fekernels/Elasticity.hh:249-385) and the reference-state stress
(IsotropicLinearElasticity3D::f1v_infinitesimalStrain_refState, IsotropicLinearElasticity.hh:848)
against the oracle, through the C ABI.

Bar: FP64 within 1e-12 relative, FP32 within 2e-5; the Jacobian action carries neither term; the
reference's Gravity3D / BodyForce3D (TetP2 on its tet.msh) and GravityRefState3D (HexQ1 and
TetP2) MMS cases leave the interior rows of the device residual at zero.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import meshgen as mg  # noqa: E402
from test_oracle_f0 import F0_BODY, F0_GRAV, F0_GRAVBODY, f0_aux, mms_case  # noqa: E402

pytestmark = pytest.mark.gpu
F1, JF3 = 1, 2
TOL64, TOL32 = 1e-12, 2e-5


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def mesh_case(name):
    if name == "q1_nonaffine":
        m = mg.hex_box((9, 7, 5), (0, 0, 0), (900.0, 700.0, 500.0), ordered=False)
        rng = np.random.default_rng(11)
        x = m.coords.copy()
        inner = np.all((x > 1) & (x < np.array([899, 699, 499])), axis=1)
        x[inner] += rng.uniform(-25, 25, (inner.sum(), 3))
        m.coords = m.node_coords = x
        return m
    return mg.tet_p2_box(6, (0, 0, 0), (600.0, 600.0, 600.0))


def random_f0_aux(m, kind, seed, per_qp=False, nq=1):
    """Per-cell moduli; density, body force and gravity per cell, or per point when per_qp."""
    rng = np.random.default_rng(seed)
    base, _ = mg.elastic_aux(m.num_cells, seed=seed, vs_sigma=0.2, vp_vs=1.732)
    n = m.num_cells * (nq if per_qp else 1)
    b = np.repeat(base, nq, 0) if per_qp else base
    b = b.copy()
    b[:, 0] = rng.uniform(2000.0, 3000.0, n)
    body = rng.normal(scale=1e4, size=(n, 3)) if kind in (F0_BODY, F0_GRAVBODY) else None
    grav = rng.normal(scale=10.0, size=(n, 3)) if kind in (F0_GRAV, F0_GRAVBODY) else None
    return f0_aux(b, body=body, grav=grav)


@pytest.mark.parametrize("kind", [F0_BODY, F0_GRAV, F0_GRAVBODY])
@pytest.mark.parametrize("mesh_name", ["q1_nonaffine", "p2"])
def test_residual_with_f0_matches_oracle(oracle, kind, mesh_name):
    from pylith_amd.dm import DM
    m = mesh_case(mesh_name)
    aux, sizes = random_f0_aux(m, kind, seed=kind)
    u = mg.random_field(m.num_nodes, 3) * 1e-3
    P = oracle.Problem(m, aux, sizes)
    ref = P.residual(u, F1, f0=kind)
    ref_j = P.residual(u, F1)  # J v = F(v) - F(0): no f0 term
    print(f"[f0] {mesh_name} kind={kind} cells={m.num_cells}", flush=True)
    for order in ("residual_first", "aux_first"):
        dm = DM(m)
        if order == "residual_first":
            dm.add_residual(0, 0, kind, F1)
            dm.add_jacobian(0, 0, 0, 0, 0, JF3)
            dm.set_auxiliary(0, 0, aux, sizes)
        else:
            dm.add_jacobian(0, 0, 0, 0, 0, JF3)
            dm.set_auxiliary(0, 0, aux, sizes)
            dm.add_residual(0, 0, kind, F1)
        for dt, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
            x = dev(u, dt)
            r = torch.zeros_like(x)
            dm.compute_residual(0, 0, x, r, mode=1)
            torch.cuda.synchronize()
            assert relerr(r.double().cpu().numpy(), ref) <= tol, (order, dt)
            r2 = torch.full_like(x, 1.0)
            dm.compute_residual(0, 0, x, r2, mode=0)
            torch.cuda.synchronize()
            assert relerr(r2.double().cpu().numpy() - 1.0, ref) <= 10 * tol, (order, dt)
            jv = torch.zeros_like(x)
            dm.compute_jacobian_action(0, 0, x, jv, mode=1)
            torch.cuda.synchronize()
            assert relerr(jv.double().cpu().numpy(), ref_j) <= tol, (order, dt)
        dm.close()


def test_body_force_per_quadrature_point(oracle):
    """Aux given per quadrature point (Q1, 8 points): the body force varies inside a cell."""
    from pylith_amd.dm import DM
    m = mesh_case("q1_nonaffine")
    aux, sizes = random_f0_aux(m, F0_GRAVBODY, seed=4, per_qp=True, nq=8)
    u = mg.random_field(m.num_nodes, 5) * 1e-3
    ref = oracle.Problem(m, aux, sizes, per_qp=True).residual(u, F1, f0=F0_GRAVBODY)
    dm = DM(m)
    dm.add_residual(0, 0, F0_GRAVBODY, F1)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(0, 0, aux, sizes, per_qp=True)
    x = dev(u)
    r = torch.zeros_like(x)
    dm.compute_residual(0, 0, x, r, mode=1)
    torch.cuda.synchronize()
    assert relerr(r.cpu().numpy(), ref) <= TOL64
    dm.close()


@pytest.mark.parametrize("kind", [F0_GRAV, F0_BODY, F0_GRAVBODY])
def test_mms_interior_rows_vanish_on_device(oracle, kind):
    from pylith_amd.dm import DM
    m, (aux, sizes), u, interior = mms_case(kind)
    P = oracle.Problem(m, aux, sizes)
    scale = np.abs(P.residual(np.zeros_like(u), 0, f0=kind).reshape(-1, 3)[interior]).max()
    dm = DM(m)
    dm.add_residual(24, 0, kind, F1)
    dm.add_jacobian(24, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(24, 0, aux, sizes)
    x = dev(u)
    r = torch.zeros_like(x)
    dm.compute_residual(24, 0, x, r, mode=1)
    torch.cuda.synchronize()
    rr = r.cpu().numpy().reshape(-1, 3)
    assert np.abs(rr[interior]).max() <= 1e-10 * scale
    assert relerr(r.cpu().numpy(), P.residual(u, F1, f0=kind)) <= TOL64
    dm.close()


def test_f0_layout_errors():
    from pylith_amd.dm import DM, PlxError
    m = mesh_case("p2")
    base, sizes = mg.elastic_aux(m.num_cells)
    dm = DM(m)
    dm.add_residual(0, 0, F0_GRAV, F1)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    with pytest.raises(PlxError):
        dm.set_auxiliary(0, 0, base, sizes)  # no gravity subfield
    with pytest.raises(PlxError):
        dm.add_residual(0, 0, 42, F1)  # no device twin
    dm.close()


F1_REF = 10


@pytest.mark.parametrize("mesh_name", ["q1_nonaffine", "p2"])
@pytest.mark.parametrize("with_gravity", [False, True])
def test_reference_state_matches_oracle(oracle, mesh_name, with_gravity):
    """f1 with reference stress and strain per quadrature point (and gravity): device = oracle;
    the Jacobian action is the plain elastic one."""
    from pylith_amd.dm import DM
    from test_oracle_f0 import refstate_aux
    m = mesh_case(mesh_name)
    nq = 8 if m.cell_type == mg.HEX_Q1 else 14
    rng = np.random.default_rng(21)
    base, _ = mg.elastic_aux(m.num_cells, seed=3, vs_sigma=0.2, vp_vs=1.732)
    n = m.num_cells * nq
    rs = rng.normal(scale=1e6, size=(n, 6))
    re = rng.normal(scale=1e-4, size=(n, 6))
    grav = (0.1, -0.2, -9.8) if with_gravity else None
    aux, sizes = refstate_aux(base, nq, rs, re, grav=grav)
    f0 = F0_GRAV if with_gravity else 0
    u = mg.random_field(m.num_nodes, 6) * 1e-3
    P = oracle.Problem(m, aux, sizes, per_qp=True)
    ref = P.residual(u, F1_REF, f0=f0)
    ref_j = oracle.Problem(m, *mg.elastic_aux(m.num_cells, seed=3, vs_sigma=0.2, vp_vs=1.732)).residual(u, F1)
    print(f"[refstate] {mesh_name} gravity={with_gravity}", flush=True)
    dm = DM(m)
    dm.add_residual(0, 0, f0, F1_REF)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(0, 0, aux, sizes, per_qp=True)
    for dt, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
        x = dev(u, dt)
        r = torch.zeros_like(x)
        dm.compute_residual(0, 0, x, r, mode=1)
        torch.cuda.synchronize()
        assert relerr(r.double().cpu().numpy(), ref) <= tol, dt
        jv = torch.zeros_like(x)
        dm.compute_jacobian_action(0, 0, x, jv, mode=1)
        torch.cuda.synchronize()
        assert relerr(jv.double().cpu().numpy(), ref_j) <= tol, dt
    dm.close()


@pytest.mark.parametrize("name", ["mms_hex.msh", "mms_tet.msh"])
def test_gravity_reference_state_mms_on_device(oracle, name):
    from pylith_amd.dm import DM
    from test_oracle_f0 import gravity_refstate_case
    m, aux, sizes, interior = gravity_refstate_case(oracle, name)
    P = oracle.Problem(m, aux, sizes, per_qp=True)
    u = np.zeros(3 * m.num_nodes)
    scale = np.abs(P.residual(u, 0, f0=F0_GRAV).reshape(-1, 3)[interior]).max()
    dm = DM(m)
    dm.add_residual(24, 0, F0_GRAV, F1_REF)
    dm.add_jacobian(24, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(24, 0, aux, sizes, per_qp=True)
    x = dev(u)
    r = torch.zeros_like(x)
    dm.compute_residual(24, 0, x, r, mode=1)
    torch.cuda.synchronize()
    rr = r.cpu().numpy()
    assert np.abs(rr.reshape(-1, 3)[interior]).max() <= 1e-10 * scale
    assert relerr(rr, P.residual(u, F1_REF, f0=F0_GRAV)) <= TOL64
    dm.close()


@pytest.mark.parametrize("kind", [F0_GRAV, F0_BODY, F0_GRAVBODY])
def test_solve_recovers_mms_solution(kind):
    """Dirichlet values of the exact MMS displacement on the boundary, right-hand side F(0) (the
    load): the FP64 CG solve reproduces the quadratic exact solution, which P2 represents."""
    from pylith_amd import _lib as L
    from pylith_amd.dm import DM, Solver
    m, (aux, sizes), u_exact, interior = mms_case(kind)
    boundary = np.setdiff1d(np.arange(m.num_nodes), interior)
    dofs = (3 * boundary[:, None] + np.arange(3)).ravel().astype(np.int32)
    dm = DM(m)
    dm.add_residual(24, 0, kind, F1)
    dm.add_jacobian(24, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(24, 0, aux, sizes)
    dm.set_constraints(dofs)
    solver = Solver(dm, label_value=24)
    zero = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(zero)
    dm.compute_residual(24, 0, zero, b, mode=1)
    x = torch.zeros_like(zero)
    dm.insert_boundary_values(dev(u_exact[dofs]), x)
    st = solver.solve(x, b=b, method=L.KSP_CG, rtol=1e-13, maxit=20000)
    torch.cuda.synchronize()
    assert st["converged"], st
    err = np.abs(x.cpu().numpy() - u_exact).max() / np.abs(u_exact).max()
    assert err <= 1e-8, (err, st)
    solver.close()
    dm.close()
