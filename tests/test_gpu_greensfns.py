"""Batched Green's functions on the device: the assembled merged operator against the matrix-free
apply, block solves against one-at-a-time FP64 fault solves, and block-width invariance."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import _lib as L  # noqa: E402
from pylith_amd import meshgen as mg  # noqa: E402

pytestmark = pytest.mark.gpu
F1, JF3 = 1, 2


def faulted_problem(n=(8, 6, 6), shear=0.5):
    from pylith_amd.dm import DM
    hi = 1000.0 * np.array(n, float)
    m, f = mg.hex_box_with_fault(n, (0.0, 0.0, -hi[2]), (hi[0], hi[1], 0.0), n[0] // 2, brick=4, shear=shear)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=4, vs_sigma=0.1, vp_vs=1.75)
    dm = DM(m)
    dm.add_residual(0, 0, 0, F1)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(0, 0, aux, sizes)
    dm.add_fault(f)
    x0 = m.node_coords[:, 0] - shear * m.node_coords[:, 2]
    fixed = np.nonzero(np.isclose(x0, 0.0) | np.isclose(x0, hi[0]))[0]
    dofs = np.sort((3 * fixed[:, None] + np.arange(3)).ravel()).astype(np.int32)
    dm.set_constraints(dofs)
    return m, f, dm, dofs


def test_assembled_operator_matches_matrix_free():
    from pylith_amd.greensfns import GreensFunctions
    m, f, dm, dofs = faulted_problem()
    gf = GreensFunctions(dm)
    nd = 3 * m.num_nodes
    k = 8
    free = np.ones(nd, bool)
    free[dofs] = False
    free.reshape(-1, 3)[f.pos_node] = False
    rng = np.random.default_rng(2)
    X = rng.standard_normal((nd, k)) * free[:, None]
    Y = torch.zeros(nd, k, dtype=torch.float64, device="cuda")
    gf.apply(torch.from_numpy(X).cuda(), Y, k)
    Yb = Y.cpu().numpy()
    nloc = dm.sizes()[0]
    for c in range(k):
        xf = np.zeros(nloc)
        xf[:nd] = X[:, c]
        xf[:nd].reshape(-1, 3)[f.pos_node] = xf[:nd].reshape(-1, 3)[f.neg_node]
        jv = torch.zeros(nloc, dtype=torch.float64, device="cuda")
        dm.compute_jacobian_action(0, 0, torch.from_numpy(xf).cuda(), jv, mode=L.INSERT_VALUES)
        y = -jv.cpu().numpy()[:nd]
        y.reshape(-1, 3)[f.neg_node] += y.reshape(-1, 3)[f.pos_node]
        err = np.abs(y[free] - Yb[free, c]).max() / np.abs(y[free]).max()
        assert err <= 1e-12, (c, err)
    # constrained dofs and positive split nodes are identity rows
    assert np.array_equal(Yb[~free], X[~free])
    # FP32 copy of the operator
    Y32 = torch.zeros(nd, k, dtype=torch.float32, device="cuda")
    gf.apply(torch.from_numpy(X).float().cuda(), Y32, k)
    assert np.abs(Y32.double().cpu().numpy() - Yb).max() <= 1e-5 * np.abs(Yb).max()


def test_p2_assembled_operator_matches_matrix_free():
    """P2 tets, two materials: the handle's operator (the assembled Jacobian of both keys,
    negated, rows in Morton order) against the matrix-free Jacobian action; constrained dofs
    are identity rows."""
    from pylith_amd.dm import DM
    from pylith_amd.greensfns import GreensFunctions
    m = mg.tet_p2_box(4, (0, 0, 0), (4000.0, 4000.0, 4000.0), brick=2)
    m.labels = (m.coords[m.cone[:, 0], 2] >= 2000.0).astype(np.int32) + 1
    dm = DM(m)
    for key in (1, 2):
        aux, sizes = mg.elastic_aux(m.num_cells, seed=7 + key, vs_sigma=0.2, vp_vs=1.75)
        dm.add_residual(key, 0, 0, F1)
        dm.add_jacobian(key, 0, 0, 0, 0, JF3)
        dm.set_auxiliary(key, 0, aux, sizes)
    dofs, _ = mg.shear_bcs(m.node_coords, 0.0, 4000.0)
    dm.set_constraints(dofs)
    gf = GreensFunctions(dm, label_value=L.ALL_KEYS)
    nd = 3 * m.num_nodes
    k = 4
    free = np.ones(nd, bool)
    free[dofs] = False
    rng = np.random.default_rng(3)
    X = rng.standard_normal((nd, k)) * free[:, None]
    Y = torch.zeros(nd, k, dtype=torch.float64, device="cuda")
    gf.apply(torch.from_numpy(X).cuda(), Y, k)
    Yb = Y.cpu().numpy()
    for c in range(k):
        jv = torch.zeros(nd, dtype=torch.float64, device="cuda")
        for key in (1, 2):
            dm.compute_jacobian_action(key, 0, torch.from_numpy(X[:, c].copy()).cuda(), jv, mode=L.ADD_VALUES)
        y = -jv.cpu().numpy()
        err = np.abs(y[free] - Yb[free, c]).max() / np.abs(y[free]).max()
        assert err <= 1e-12, (c, err)
    assert np.array_equal(Yb[~free], X[~free])


def test_block_solves_match_single_fp64_solves():
    from pylith_amd.dm import Solver
    from pylith_amd.greensfns import GreensFunctions
    m, f, dm, dofs = faulted_problem()
    slip, _ = mg.fault_vertex_impulses(f, (1, 6), (1, 4), component=1)
    slip = slip[:12]
    nimp = slip.shape[0]
    nloc = dm.sizes()[0]
    nu = 3 * m.num_nodes
    gf = GreensFunctions(dm)
    gf.set_options(block=8)
    U = torch.zeros(nimp, nloc, dtype=torch.float64, device="cuda")
    st = gf.solve(torch.from_numpy(slip).cuda(), U, rtol=1e-11)
    print(f"[gf] {st}", flush=True)
    assert st["converged"] and st["num_blocks"] == 2, st
    solver = Solver(dm)
    for j in (0, 5, 11):
        dm.set_fault_slip(slip[j])
        x = torch.zeros(nloc, dtype=torch.float64, device="cuda")
        s1 = solver.solve(x, method=L.KSP_CG, rtol=1e-12, maxit=5000)
        assert s1["converged"], s1
        ref = x.cpu().numpy()[:nu]
        got = U[j].cpu().numpy()[:nu]
        assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max(), j
    # another block width gives the same solutions
    gf.set_options(block=16)
    U2 = torch.zeros_like(U)
    st2 = gf.solve(torch.from_numpy(slip).cuda(), U2, rtol=1e-11)
    assert st2["num_blocks"] == 1
    d = (U2 - U).abs().max().item()
    assert d <= 1e-9 * U.abs().max().item()


@pytest.mark.parametrize("block", [8, 64])
def test_multigrid_preconditioner_matches_jacobi(block):
    """The multigrid-preconditioned block PCG reaches the same displacements as the Jacobi one in
    far fewer inner iterations (block 64 takes the wave-per-row fine-level product)."""
    from pylith_amd.greensfns import GreensFunctions
    m, f, dm, dofs = faulted_problem((12, 8, 8))
    slip, _ = mg.fault_vertex_impulses(f, (1, 8), (1, 6), component=1)
    slip = slip[:20]
    nloc = dm.sizes()[0]
    gf = GreensFunctions(dm)
    gf.set_options(block=block)
    U = torch.zeros(slip.shape[0], nloc, dtype=torch.float64, device="cuda")
    sj = gf.solve(torch.from_numpy(slip).cuda(), U, rtol=1e-10)
    gf.set_preconditioner("amg")
    for rep in range(2):  # the second solve replays the captured chunks
        Ua = torch.zeros_like(U)
        sa = gf.solve(torch.from_numpy(slip).cuda(), Ua, rtol=1e-10)
        print(f"[gf-amg] block={block} rep={rep} jacobi={sj} amg={sa}", flush=True)
        assert sa["converged"] and sa["max_relative_residual"] <= 1e-10, sa
        assert sa["inner_iterations"] * 4 <= sj["inner_iterations"], (sa, sj)
        d = (Ua - U).abs().max().item()
        assert d <= 1e-8 * U.abs().max().item(), d
    # back to Jacobi: same answer again
    gf.set_preconditioner("jacobi")
    Uj = torch.zeros_like(U)
    sj2 = gf.solve(torch.from_numpy(slip).cuda(), Uj, rtol=1e-10)
    assert sj2["inner_iterations"] == sj["inner_iterations"]
    assert (Uj - U).abs().max().item() <= 1e-9 * U.abs().max().item()
    with pytest.raises(Exception):
        gf.set_preconditioner(7)


def test_autotune_picks_a_candidate():
    from pylith_amd.greensfns import GreensFunctions
    m, f, dm, dofs = faulted_problem((6, 4, 4))
    gf = GreensFunctions(dm)
    best, times = gf.autotune()
    assert best in (8, 16, 32, 64)
    assert all(t > 0 for t in times.values())
    assert times[best] == min(times.values())
