"""The three system workloads at their full sizes, checked against FP64 references.

* C3: 48^3 Q1 generalized Maxwell (3 elements per point, viscosities log-uniform in
  [1e17, 1e19] Pa s) + a vertical fault with a tapered 2 m patch, 10 implicit steps of 0.1 yr.
  At every step the oracle's full residual (domain + cohesive rows) at the device solution is
  <= 1e-8 in each block against that block's own scale (force rows: the nodal forces of the
  domain at the solution; Lagrange-multiplier rows: the slip forcing), and the device's state
  update equals the oracle's to 1e-12.
* C4: 64^3 Q1, 80 km cube, fault dipping 60 degrees, all 512 vertex impulses as one batched
  solve: every impulse's FP64 residual (the matrix-free device apply, which matches the oracle
  to 1e-12) is <= the solve tolerance, and every impulse's error against a 1e-14 FP64 solve is
  <= 1e-8.  A 20-impulse block on a 12 x 8 x 8 mesh matches a direct solve of the oracle's
  assembled saddle-point system.
* C2: P2 64^3 x 6 (6.44 M dofs, lognormal moduli, shear Dirichlet values) solved by FGMRES with
  the FP32 multigrid PCG inside FP64 refinement to 1e-10: the oracle's FP64 residual of that
  solution on the free dofs is <= 1e-9 of the initial one, and it equals an FP64 CG solve to
  1e-12 (multigrid-preconditioned) within 1e-8.
* C5: the serving mix of 16^3 / 32^3 / 64^3 / 96^3 meshes: every request's FP64 residual is
  <= the serving tolerance (1e-9), its error against a 1e-13 FP64 solve is <= 1e-8, and
  requests solved alone equal the same requests solved in the mix to 1e-8.

Reference: problems/GreensFns.cc:318-331 (one solve per impulse), faults/FaultCohesiveImpulses.cc
:263-300 (impulse basis), materials/IsotropicLinearGenMaxwell.hh, faults/FaultCohesiveKin.hh.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import _lib as L  # noqa: E402
from pylith_amd import meshgen as mg  # noqa: E402
from pylith_amd import workloads  # noqa: E402

pytestmark = pytest.mark.gpu
F1_GM, JF3_GM, VISC, TOTAL = 3, 4, 5, 6


def gm_update(P, u):
    new = P.update_state(u, VISC, 5)
    tot = P.update_state(u, TOTAL, 6)
    new[:, 27:33] = tot[:, 27:33]
    return new


def test_c3_genmaxwell_fault_48cube_ten_steps(oracle):
    from pylith_amd.dm import DM, Solver
    n = 48
    m, f = mg.hex_box_with_fault(n, (0.0, 0.0, -6000.0), (6000.0, 6000.0, 0.0), n // 2, brick=8)
    aux, sizes = mg.genmax_aux(m.num_cells, 8, seed=20261202, log_eta_range=(1e17, 1e19),
                               ratios=(0.3, 0.3, 0.3))
    dt = 0.1 * mg.YEAR
    dm = DM(m)
    dm.add_residual(0, 0, 0, F1_GM)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3_GM)
    dm.set_auxiliary(0, 0, aux, sizes, per_qp=True)
    dm.set_constants(0, 0, [dt])
    dm.add_fault(f)
    slip = mg.tapered_patch_slip(f, (3000.0, -3000.0), 1000.0, 2.0, component=1, taper=1000.0)
    dm.set_fault_slip(slip)
    fixed = np.nonzero(np.isclose(m.node_coords[:, 0], 0.0) | np.isclose(m.node_coords[:, 0], 6000.0))[0]
    dofs = np.sort((3 * fixed[:, None] + np.arange(3)).ravel()).astype(np.int32)
    dm.set_constraints(dofs)
    solver = Solver(dm)
    solver.set_preconditioner("amg")
    nloc = dm.sizes()[0]
    nu = 3 * m.num_nodes
    coh = oracle.Cohesive(m, f, slip, nu)
    x = torch.zeros(nloc, dtype=torch.float64, device="cuda")
    zero = torch.zeros_like(x)
    b = torch.zeros_like(x)
    print(f"[c3] cells={m.num_cells} dofs={nu} lagrange={nloc - nu}", flush=True)
    for step in range(10):
        P = oracle.Problem(m, aux, sizes, per_qp=True, constants=[dt])

        dm.compute_residual(0, 0, zero, b, mode=L.INSERT_VALUES)  # the state term F(0)
        st = solver.solve(x, b=b, method=L.KSP_FGMRES_IR, rtol=1e-10, maxit=200)
        assert st["converged"], (step, st)
        xs = x.cpu().numpy()
        dom = P.residual(xs[:nu], F1_GM)
        full = coh.residual(xs)
        full[:nu] += dom
        full[dofs] = 0.0
        # force rows against the domain's nodal forces at the solution; constraint rows
        # (u+ - u- - d) against the slip forcing (their value at u = 0)
        rel_u = np.abs(full[:nu]).max() / np.abs(np.delete(dom, dofs)).max()
        rel_l = np.abs(full[nu:]).max() / np.abs(coh.residual(np.zeros(nloc))[nu:]).max()
        rel = max(rel_u, rel_l)
        dm.update_state_vars(0, 0, x)
        got = dm.get_auxiliary(0, 0, aux)
        aux = gm_update(P, xs[:nu])
        dstate = max(np.abs(got[:, 9:27] - aux[:, 9:27]).max() / np.abs(aux[:, 9:27]).max(),
                     np.abs(got[:, 27:33] - aux[:, 27:33]).max() / np.abs(aux[:, 27:33]).max())
        print(f"[c3] step {step}: oracle full residual {rel_u:.2e} (forces) {rel_l:.2e} (constraints), state update "
              f"{dstate:.2e}, solve {st}", flush=True)
        assert rel <= 1e-8, (step, rel)
        assert dstate <= 1e-12, (step, dstate)
        assert np.array_equal(got[:, :9], aux[:, :9])
    solver.close()
    dm.close()


def merged_residual(dm, f, dofs, u):
    """The constraint-eliminated residual of a displacement u (local vector, Lagrange entries
    zero) with the matrix-free FP64 apply: J u with each positive split node's row added to its
    negative twin, positive rows and constrained dofs dropped."""
    nu = 3 * f_num_nodes(dm)
    jv = torch.zeros_like(u)
    dm.compute_jacobian_action(0, 0, u, jv, mode=L.INSERT_VALUES)
    r = jv[:nu].clone().view(-1, 3)
    neg = torch.from_numpy(f.neg_node.astype(np.int64)).cuda()
    pos = torch.from_numpy(f.pos_node.astype(np.int64)).cuda()
    r[neg] += r[pos]
    r[pos] = 0.0
    r = r.view(-1)
    r[torch.from_numpy(dofs.astype(np.int64)).cuda()] = 0.0
    return r


def f_num_nodes(dm):
    return dm.mesh.num_nodes


def jump_vector(m, f, slip, nloc):
    """u_jump: the slip (fault frame) as a displacement of the positive split nodes."""
    nrm, t1, t2 = mg.fault_normal_frame(f, m.coords)
    v = np.zeros(nloc)
    v[:3 * m.num_nodes].reshape(-1, 3)[f.pos_node] = slip @ np.stack([nrm, t1, t2])
    return torch.from_numpy(v).cuda()


def test_c4_all_512_impulses_64cube():
    from pylith_amd.dm import Solver
    from pylith_amd.greensfns import GreensFunctions
    prob = workloads.faulted_box(64, length=80000.0, dip_deg=60.0)
    m, f, dm = prob.mesh, prob.fault, prob.dm
    slip, _ = mg.fault_vertex_impulses(f, (16, 48), (48, 64), component=1)
    assert slip.shape[0] == 512
    nloc, nu = prob.local_size, 3 * m.num_nodes
    gf = GreensFunctions(dm)
    gf.set_preconditioner("amg")
    gf.set_options(block=64)
    rtol = 1e-8
    U = torch.zeros(512, nloc, dtype=torch.float64, device="cuda")
    st = gf.solve(torch.from_numpy(slip).cuda(), U, rtol=rtol, maxit=30)
    print(f"[c4] block solve {st}", flush=True)
    assert st["converged"], st
    # every impulse: FP64 residual through the matrix-free apply
    res = []
    for j in range(512):
        r = merged_residual(dm, f, prob.dofs, U[j])
        b = merged_residual(dm, f, prob.dofs, jump_vector(m, f, slip[j], nloc))
        res.append((r.norm() / b.norm()).item())
    res = np.array(res)
    print(f"[c4] FP64 relative residual: max {res.max():.2e} median {np.median(res):.2e}", flush=True)
    assert res.max() <= 1.01 * rtol, res.max()
    # every impulse: error against a tight FP64 solve (FGMRES on the matrix-free FP64 operator)
    solver = Solver(dm)
    solver.set_preconditioner("amg")
    solver.set_inner(rtol=1e-3, maxit=400, restart=30)
    err = np.zeros(512)
    for j in range(512):
        dm.set_fault_slip(slip[j])
        x = torch.zeros(nloc, dtype=torch.float64, device="cuda")
        s = solver.solve(x, method=L.KSP_FGMRES_IR, rtol=1e-14, maxit=400)
        assert s["converged"], (j, s)
        d = (U[j, :nu] - x[:nu]).abs().max() / x[:nu].abs().max()
        err[j] = d.item()
        if j % 128 == 0:
            print(f"[c4] impulse {j}: error vs FP64 {err[j]:.2e}", flush=True)
    print(f"[c4] error vs 1e-14 FP64 solves: max {err.max():.2e} median {np.median(err):.2e}", flush=True)
    assert err.max() <= 1e-8, err.max()
    solver.close()
    gf.close()
    dm.close()


def test_c4_block_matches_oracle_direct_solve_12x8x8(oracle):
    """20 impulses of a sheared (dipping) 12 x 8 x 8 faulted mesh in one block against a direct
    solve of the full saddle-point system assembled from the oracle (domain + cohesive rows)."""
    import scipy.linalg
    from pylith_amd.dm import DM
    from pylith_amd.greensfns import GreensFunctions
    n, shear = (12, 8, 8), 0.5
    hi = 1000.0 * np.array(n, float)
    m, f = mg.hex_box_with_fault(n, (0.0, 0.0, -hi[2]), (hi[0], hi[1], 0.0), n[0] // 2, brick=4, shear=shear)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=4, vs_sigma=0.1, vp_vs=1.75)
    dm = DM(m)
    dm.add_residual(0, 0, 0, 1)
    dm.add_jacobian(0, 0, 0, 0, 0, 2)
    dm.set_auxiliary(0, 0, aux, sizes)
    dm.add_fault(f)
    x0 = m.node_coords[:, 0] - shear * m.node_coords[:, 2]
    fixed = np.nonzero(np.isclose(x0, 0.0) | np.isclose(x0, hi[0]))[0]
    dofs = np.sort((3 * fixed[:, None] + np.arange(3)).ravel()).astype(np.int32)
    dm.set_constraints(dofs)
    slip, _ = mg.fault_vertex_impulses(f, (1, 8), (1, 6), component=1)
    slip = slip[:20]
    nu = 3 * m.num_nodes
    nloc = dm.sizes()[0]
    # oracle Jacobian of F(x) = domain + cohesive rows, column by column
    P = oracle.Problem(m, aux, sizes)
    C0 = oracle.Cohesive(m, f, np.zeros((f.num_lag, 3)), nu)
    J = np.zeros((nloc, nloc))
    e = np.zeros(nloc)
    for i in range(nloc):
        e[i] = 1.0
        col = C0.jacobian_action(e)
        col[:nu] += P.residual(e[:nu], 1)
        J[:, i] = col
        e[i] = 0.0
    free = np.setdiff1d(np.arange(nloc), dofs)
    lu = scipy.linalg.lu_factor(J[np.ix_(free, free)])
    gf = GreensFunctions(dm)
    gf.set_preconditioner("amg")
    gf.set_options(block=32)
    U = torch.zeros(20, nloc, dtype=torch.float64, device="cuda")
    st = gf.solve(torch.from_numpy(slip).cuda(), U, rtol=1e-11, maxit=30)
    assert st["converged"] and st["num_blocks"] == 1, st
    worst = 0.0
    for j in range(20):
        F0 = oracle.Cohesive(m, f, slip[j], nu).residual(np.zeros(nloc))
        x = np.zeros(nloc)
        x[free] = scipy.linalg.lu_solve(lu, -F0[free])
        got = U[j].cpu().numpy()[:nu]
        worst = max(worst, np.abs(got - x[:nu]).max() / np.abs(x[:nu]).max())
    print(f"[c4 direct] 20 impulses, max relative error vs direct solve {worst:.2e}", flush=True)
    assert worst <= 1e-8, worst
    gf.close()
    dm.close()


def test_c5_serving_mix_16_to_96():
    from pylith_amd import serving
    from pylith_amd.dm import Solver
    sizes, probs = (16, 32, 64, 96), (0.4, 0.3, 0.2, 0.1)
    # a relative residual of 1e-8 left errors up to 1.6e-8 against FP64 on this mix: serving
    # solves to 1e-9 so that the error stays within 1e-8
    rtol = 1e-9
    eng = serving.ServingEngine(sizes, block=16, rtol=rtol, keep_results=True)
    eng.warm()
    reqs = serving.make_requests(40, sizes, probs, rate=None, seed=17)
    assert {r.size for r in reqs} == set(sizes)
    mixed = eng.run(reqs)
    print(f"[c5] mix of {len(mixed)} requests: {serving.latency_stats(mixed)}", flush=True)
    worst_res, worst_err = 0.0, 0.0
    solvers = {}
    for r in mixed:
        assert getattr(r, "error", None) is None, (r.rid, r.error)
        st = eng.states[r.size]
        prob = st.prob
        nloc = prob.local_size
        slip = workloads.patch_slip(prob.fault, st.patch, r.slip_values)
        with torch.cuda.stream(st.stream):  # the stream the mesh state's DM launches on
            u = torch.from_numpy(r.u).cuda()
            res = merged_residual(prob.dm, prob.fault, prob.dofs, u)
            b = merged_residual(prob.dm, prob.fault, prob.dofs, jump_vector(prob.mesh, prob.fault, slip, nloc))
            worst_res = max(worst_res, (res.norm() / b.norm()).item())
            if r.size not in solvers:
                solvers[r.size] = Solver(prob.dm)
                solvers[r.size].set_preconditioner("amg")
                solvers[r.size].set_inner(rtol=1e-3, maxit=400, restart=30)
            prob.dm.set_fault_slip(slip)
            x = torch.zeros(nloc, dtype=torch.float64, device="cuda")
            s = solvers[r.size].solve(x, method=L.KSP_FGMRES_IR, rtol=1e-13, maxit=400)
            assert s["converged"], (r.rid, s)
            nu = 3 * prob.mesh.num_nodes
            worst_err = max(worst_err, ((u[:nu] - x[:nu]).abs().max() / x[:nu].abs().max()).item())
    print(f"[c5] per request: FP64 residual max {worst_res:.2e}, error vs FP64 max {worst_err:.2e}",
          flush=True)
    assert worst_res <= 1.01 * rtol, worst_res
    assert worst_err <= 1e-8, worst_err
    # the same requests alone: one of each size
    for sz in sizes:
        r = next(q for q in mixed if q.size == sz)
        (a,) = eng.run([serving.Request(r.rid, r.size, r.slip_values)])
        d = np.abs(a.u - r.u).max() / np.abs(a.u).max()
        print(f"[c5] size {sz}: alone vs mix {d:.2e}", flush=True)
        assert d <= 1e-8, (sz, d)
    for s in solvers.values():
        s.close()


def test_c2_fp32_refinement_against_tight_fp64(oracle):
    from pylith_amd.dm import DM, Solver
    m = mg.tet_p2_box(64, (0.0, 0.0, 0.0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(m.num_cells, seed=20261202, vs_sigma=0.2, vp_vs=1.732)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    dm = DM(m)
    dm.add_residual(0, 0, 0, 1)
    dm.add_jacobian(0, 0, 0, 0, 0, 2)
    dm.set_auxiliary(0, 0, aux, sizes)
    dm.set_constraints(dofs)
    solver = Solver(dm)
    solver.set_preconditioner("amg")
    solver.set_inner(1e-3, 20000, 30)
    print(f"[c2] dofs={3 * m.num_nodes} hierarchy built", flush=True)
    sols = {}
    for name, method, rtol in (("fp64_cg", L.KSP_CG, 1e-12), ("fp32_ir", L.KSP_FGMRES_IR, 1e-10)):
        x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
        dm.insert_boundary_values(torch.from_numpy(vals).cuda(), x)
        st = solver.solve(x, method=method, rtol=rtol, maxit=20000)
        print(f"[c2] {name} {st}", flush=True)
        assert st["converged"], st
        sols[name] = x.cpu().numpy()
    a, b = sols["fp32_ir"], sols["fp64_cg"]
    assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()
    # the oracle's FP64 residual of the refined solution vanishes on the unconstrained dofs
    P = oracle.Problem(m, aux, sizes)
    r = P.residual(a, 1)
    u0 = np.zeros_like(a)
    u0[dofs] = vals
    r0 = P.residual(u0, 1)
    free = np.setdiff1d(np.arange(a.size), dofs)
    rel = np.linalg.norm(r[free]) / np.linalg.norm(r0[free])
    print(f"[c2] oracle residual of the FP32-refined solution {rel:.3e}", flush=True)
    assert rel <= 1e-9
    solver.close()
    dm.close()
