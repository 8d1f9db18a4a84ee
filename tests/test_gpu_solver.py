"""Solves with the matrix-free operator: FP64 PCG, FP64 FGMRES with FP32 PCG preconditioning,
and FP32 iterative refinement, against dense direct solves of the oracle's assembled operator
(small meshes) and against the oracle residual (config-1 mesh)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import _lib as L  # noqa: E402
from pylith_amd import meshgen as mg  # noqa: E402

pytestmark = pytest.mark.gpu
F1, JF3 = 1, 2


def build(m, aux, sizes, dofs):
    from pylith_amd.dm import DM, Solver
    dm = DM(m)
    dm.add_residual(0, 0, 0, F1)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(0, 0, aux, sizes)
    dm.set_constraints(dofs)
    return dm, Solver(dm)


def dense_solution(oracle, m, aux, sizes, dofs, vals):
    P = oracle.Problem(m, aux, sizes)
    nd = 3 * m.num_nodes
    A = np.zeros((nd, nd))
    for c in range(m.num_cells):
        d = (3 * m.closure[c][:, None] + np.arange(3)).ravel()
        A[np.ix_(d, d)] -= P.element_matrix(c, JF3)  # A = -J
    u = np.zeros(nd)
    u[dofs] = vals
    free = np.setdiff1d(np.arange(nd), dofs)
    rhs = -A[np.ix_(free, dofs)] @ vals
    u[free] = np.linalg.solve(A[np.ix_(free, free)], rhs)
    return u


@pytest.mark.parametrize("kind", ["hex", "tet"])
@pytest.mark.parametrize("method", [L.KSP_CG, L.KSP_FGMRES_IR, L.KSP_IR])
def test_solve_matches_dense_direct(oracle, kind, method):
    if kind == "hex":
        m = mg.hex_box(6, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=2)
    else:
        m = mg.tet_p2_box(3, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=1)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=4, vs_sigma=0.2, vp_vs=1.732)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    ref = dense_solution(oracle, m, aux, sizes, dofs, vals)
    dm, solver = build(m, aux, sizes, dofs)
    print(f"[solver] kind={kind} method={method} dofs={3 * m.num_nodes}", flush=True)
    x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
    st = solver.solve(x, method=method, rtol=1e-12, maxit=5000)
    torch.cuda.synchronize()
    err = np.abs(x.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert st["converged"], st
    assert err <= 1e-9, (err, st)


@pytest.mark.parametrize("kind", ["hex", "tet"])
@pytest.mark.parametrize("method", [L.KSP_CG, L.KSP_FGMRES_IR, L.KSP_IR])
def test_multigrid_solve_matches_dense_direct(oracle, kind, method):
    """The V-cycle preconditioner: Q1 from the element-matrix assembly, P2 from the assembled
    Jacobian (plx_compute_jacobian) renumbered in Morton order; 1,331 P2 nodes keep a
    multi-level hierarchy (the coarsest level holds at most 768 nodes)."""
    if kind == "hex":
        m = mg.hex_box(6, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=2)
    else:
        m = mg.tet_p2_box(5, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=2)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=4, vs_sigma=0.2, vp_vs=1.732)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    ref = dense_solution(oracle, m, aux, sizes, dofs, vals)
    dm, solver = build(m, aux, sizes, dofs)
    solver.set_preconditioner("amg")
    print(f"[solver-amg] kind={kind} method={method} dofs={3 * m.num_nodes}", flush=True)
    x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
    st = solver.solve(x, method=method, rtol=1e-12, maxit=5000)
    torch.cuda.synchronize()
    err = np.abs(x.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert st["converged"], st
    assert err <= 1e-9, (err, st)


def test_multigrid_cuts_iterations_on_config1():
    """32^3 Q1 (the config-1 mesh, lognormal moduli): FP64 CG with the V-cycle reaches the
    Jacobi-preconditioned solution in a fraction of the iterations; set-up rebuilds it."""
    m = mg.hex_box(32, (0, 0, 0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(m.num_cells, seed=5, vs_sigma=0.2, vp_vs=1.732)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    dm, solver = build(m, aux, sizes, dofs)
    sols, its = {}, {}
    for pc in ("pbjacobi", "amg", "amg-rebuilt"):
        if pc == "amg":
            solver.set_preconditioner("amg")
        if pc == "amg-rebuilt":
            solver.setup()
        x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
        dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
        st = solver.solve(x, method=L.KSP_CG, rtol=1e-10, maxit=20000)
        print(f"[solver-amg] config1 pc={pc} {st}", flush=True)
        assert st["converged"], st
        sols[pc], its[pc] = x.cpu().numpy(), st["outer_iterations"]
    assert its["amg"] * 8 <= its["pbjacobi"], its
    assert its["amg-rebuilt"] == its["amg"]
    ref = sols["pbjacobi"]
    for pc in ("amg", "amg-rebuilt"):
        assert np.abs(sols[pc] - ref).max() <= 1e-8 * np.abs(ref).max()


def test_multigrid_cuts_iterations_on_p2():
    """16^3 x 6 P2 tets (lognormal moduli): FP64 CG with the V-cycle of the assembled P2
    operator against point-block Jacobi; both reach the same solution."""
    m = mg.tet_p2_box(16, (0, 0, 0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(m.num_cells, seed=5, vs_sigma=0.2, vp_vs=1.732)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    dm, solver = build(m, aux, sizes, dofs)
    sols, its = {}, {}
    for pc in ("pbjacobi", "amg"):
        if pc == "amg":
            solver.set_preconditioner("amg")
        x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
        dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
        st = solver.solve(x, method=L.KSP_CG, rtol=1e-10, maxit=20000)
        print(f"[solver-amg] p2 16^3 pc={pc} {st}", flush=True)
        assert st["converged"], st
        sols[pc], its[pc] = x.cpu().numpy(), st["outer_iterations"]
    assert its["amg"] * 4 <= its["pbjacobi"], its
    ref = sols["pbjacobi"]
    assert np.abs(sols["amg"] - ref).max() <= 1e-8 * np.abs(ref).max()


def test_config1_fp32_refinement_matches_fp64(oracle):
    m = mg.hex_box(32, (0, 0, 0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(m.num_cells)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    dm, solver = build(m, aux, sizes, dofs)
    sols = {}
    for method in (L.KSP_CG, L.KSP_FGMRES_IR):
        x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
        dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
        st = solver.solve(x, method=method, rtol=1e-11, maxit=20000)
        print(f"[solver] config1 method={method} {st}", flush=True)
        assert st["converged"], st
        sols[method] = x.cpu().numpy()
    a, b = sols[L.KSP_CG], sols[L.KSP_FGMRES_IR]
    assert np.abs(a - b).max() / np.abs(a).max() <= 1e-8
    # the oracle's FP64 residual of the refined solution vanishes on the unconstrained dofs
    P = oracle.Problem(m, aux, sizes)
    r = P.residual(b, F1)
    u0 = np.zeros_like(b)
    u0[dofs] = vals
    r0 = P.residual(u0, F1)
    free = np.setdiff1d(np.arange(b.size), dofs)
    assert np.linalg.norm(r[free]) <= 1e-10 * np.linalg.norm(r0[free])


@pytest.mark.parametrize("pc", ["pbjacobi", "amg"])
def test_captured_iterations_follow_changes_of_the_dm(pc):
    """Krylov chunks run as captured graphs: after new materials (reallocated device buffers) and
    set-up, the same solver gives what a fresh handle gives; event logging (eager path) agrees."""
    m = mg.hex_box(10, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=4)
    aux1, sizes = mg.elastic_aux(m.num_cells, seed=7, vs_sigma=0.2, vp_vs=1.732)
    aux2, _ = mg.elastic_aux(m.num_cells, seed=8, vs_sigma=0.3, vp_vs=1.8)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)

    def solve(dm, solver, method):
        x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
        dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
        st = solver.solve(x, method=method, rtol=1e-11, maxit=5000)
        assert st["converged"], st
        return x.cpu().numpy()

    dm, solver = build(m, aux1, sizes, dofs)
    solver.set_preconditioner(pc)
    for method in (L.KSP_CG, L.KSP_FGMRES_IR):
        solve(dm, solver, method)  # captures the chunks
    dm.set_auxiliary(0, 0, aux2, sizes)
    solver.setup()
    dm2, solver2 = build(m, aux2, sizes, dofs)
    solver2.set_preconditioner(pc)
    for method in (L.KSP_CG, L.KSP_FGMRES_IR):
        got, ref = solve(dm, solver, method), solve(dm2, solver2, method)
        print(f"[solver-graph] pc={pc} method={method}", flush=True)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
        dm.set_event_logging(True)
        eager = solve(dm, solver, method)
        dm.set_event_logging(False)
        assert np.abs(eager - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("kind", ["hex", "tet"])
@pytest.mark.parametrize("pc", ["pbjacobi", "amg"])
@pytest.mark.parametrize("method", [L.KSP_CG, L.KSP_FGMRES_IR])
def test_two_materials_solve_matches_dense_direct(oracle, method, pc, kind):
    """A solver over every key (two materials, labels 1 and 2): the operator is the sum of the
    keys' operators; against a dense direct solve with each cell's own moduli."""
    from pylith_amd.dm import DM, Solver
    if kind == "hex":
        m = mg.hex_box(6, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=2)
    else:
        m = mg.tet_p2_box(4, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=2)
    m.labels = (m.coords[m.cone[:, 0], 2] >= 3000.0).astype(np.int32) + 1
    aux1, sizes = mg.elastic_aux(m.num_cells, seed=4, vs_sigma=0.2, vp_vs=1.732)
    aux2, _ = mg.elastic_aux(m.num_cells, vs=1500.0, vp=2800.0)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    P1, P2 = oracle.Problem(m, aux1, sizes), oracle.Problem(m, aux2, sizes)
    nd = 3 * m.num_nodes
    A = np.zeros((nd, nd))
    for c in range(m.num_cells):
        d = (3 * m.closure[c][:, None] + np.arange(3)).ravel()
        A[np.ix_(d, d)] -= (P1 if m.labels[c] == 1 else P2).element_matrix(c, JF3)
    ref = np.zeros(nd)
    ref[dofs] = vals
    free = np.setdiff1d(np.arange(nd), dofs)
    ref[free] = np.linalg.solve(A[np.ix_(free, free)], -A[np.ix_(free, dofs)] @ vals)
    dm = DM(m)
    for lab, aux in ((1, aux1), (2, aux2)):
        dm.add_residual(lab, 0, 0, F1)
        dm.add_jacobian(lab, 0, 0, 0, 0, JF3)
        dm.set_auxiliary(lab, 0, aux, sizes)
    dm.set_constraints(dofs)
    solver = Solver(dm, label_value=L.ALL_KEYS)
    solver.set_preconditioner(pc)
    print(f"[solver-2mat] kind={kind} method={method} pc={pc}", flush=True)
    x = torch.zeros(nd, dtype=torch.float64, device="cuda")
    dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
    st = solver.solve(x, method=method, rtol=1e-12, maxit=5000)
    assert st["converged"], st
    err = np.abs(x.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err <= 1e-9, err
