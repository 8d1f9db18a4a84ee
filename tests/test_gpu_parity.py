"""Parity of the HIP residual / Jacobian-action path (through the C ABI) with the CPU oracle.

Bar: FP64 within 1e-12 relative (||r_gpu - r_cpu||_inf / ||r_cpu||_inf); FP32 within 2e-5.
"""
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import meshgen as mg  # noqa: E402
from pylith_amd import meshio_ascii  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
F1, JF3 = 1, 2
TOL64, TOL32 = 1e-12, 2e-5


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def make_dm(m, aux, sizes, label=0):
    from pylith_amd.dm import DM
    dm = DM(m)
    dm.add_residual(label, 0, 0, F1)
    dm.add_jacobian(label, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(label, 0, aux, sizes)
    return dm


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def check_residual(oracle, m, u, label=0, aux=None, dtypes=(torch.float64, torch.float32)):
    if aux is None:
        aux, sizes = mg.elastic_aux(m.num_cells)
    else:
        aux, sizes = aux
    P = oracle.Problem(m, aux, sizes)
    ref = P.residual(u, F1, cells=np.nonzero(m.labels == label)[0] if label else None)
    dm = make_dm(m, aux, sizes, label)
    print(f"[parity] cells={m.num_cells} nodes={m.num_nodes} ct={m.cell_type}", flush=True)
    from pylith_amd import _lib as L
    for algo, acc, dt in [(a, c, d) for a in (L.ALGO_QUADRATURE_POINTS, L.ALGO_FACTORED)
                          for c in (L.ACCUMULATE_COLORED, L.ACCUMULATE_LDS_ATOMIC) for d in dtypes]:
        dm.set_kernel_options(algo, acc, 0)
        x = dev(u, dt)
        r = torch.zeros_like(x)
        dm.compute_residual(label, 0, x, r, mode=1)
        torch.cuda.synchronize()
        e = relerr(r.double().cpu().numpy(), ref)
        tol = TOL64 if dt == torch.float64 else TOL32
        assert e <= tol, (dt, algo, acc, e)
        # ADD mode accumulates on top of existing values
        r2 = torch.full_like(x, 1.0)
        dm.compute_residual(label, 0, x, r2, mode=0)
        torch.cuda.synchronize()
        assert relerr(r2.double().cpu().numpy() - 1.0, ref) <= 10 * tol
        # Jacobian action (linear elasticity: J v = F(v))
        jv = torch.zeros_like(x)
        dm.compute_jacobian_action(label, 0, x, jv, mode=1)
        torch.cuda.synchronize()
        assert relerr(jv.double().cpu().numpy(), ref) <= tol
    return dm, ref


def test_q1_config1_32cube(oracle):
    m = mg.hex_box(32, (0, 0, 0), (6000.0, 6000.0, 6000.0))
    u = mg.random_field(m.num_nodes, 20261202)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    u[dofs] = vals
    check_residual(oracle, m, u)


def test_q1_nonaffine_and_unordered(oracle):
    m = mg.hex_box((9, 7, 5), (0, 0, 0), (900.0, 700.0, 500.0), ordered=False)
    rng = np.random.default_rng(11)
    x = m.coords.copy()
    inner = np.all((x > 1) & (x < np.array([899, 699, 499])), axis=1)
    x[inner] += rng.uniform(-25, 25, (inner.sum(), 3))
    m.coords = m.node_coords = x
    u = mg.random_field(m.num_nodes, 4)
    check_residual(oracle, m, u)


def test_q1_reference_fixture_mesh_and_label(oracle):
    m, groups = meshio_ascii.read(os.path.join(GOLDEN, "hex.mesh"))
    u = mg.random_field(m.num_nodes, 9)
    aux, sizes = mg.elastic_aux(m.num_cells, 2500.0, 3000.0, math.sqrt(3.0) * 3000.0)
    check_residual(oracle, m, u, label=24, aux=(aux, sizes))


@pytest.mark.parametrize("name", ["mms_hex.msh", "mms_tet.msh"])
def test_reference_mms_meshes(oracle, name):
    """The reference's MMS meshes (Gmsh 4.1 binary; tet.msh raised to P2): device residual of the
    UniformStrain3D displacement against the oracle, interior rows zero (UniformStrain3D.cc)."""
    from pylith_amd import meshio_gmsh
    from test_meshio_gmsh import mms_aux, uniform_strain_disp
    m, groups = meshio_gmsh.read(os.path.join(GOLDEN, name))
    print(f"[parity] {name} cells={m.num_cells} nodes={m.num_nodes}", flush=True)
    aux = mms_aux(m)
    u = uniform_strain_disp(m.node_coords)
    check_residual(oracle, m, u, label=24, aux=aux)
    check_residual(oracle, m, mg.random_field(m.num_nodes, 21), label=24, aux=aux)
    from pylith_amd.dm import DM
    dm = DM(m)
    dm.add_residual(24, 0, 0, F1)
    dm.set_auxiliary(24, 0, *aux)
    x = dev(u)
    r = torch.zeros_like(x)
    dm.compute_residual(24, 0, x, r, mode=1)
    torch.cuda.synchronize()
    rr = r.cpu().numpy().reshape(-1, 3)
    interior = np.setdiff1d(np.arange(m.num_nodes), groups["boundary"])
    assert np.abs(rr[interior]).max() <= 1e-12 * np.abs(rr).max()


def test_q1_two_materials(oracle):
    m = mg.hex_box(8, (0, 0, 0), (800.0, 800.0, 800.0), brick=4)
    m.labels = (m.coords[m.cone[:, 0], 2] >= 400).astype(np.int32) + 1
    u = mg.random_field(m.num_nodes, 12)
    aux1, sizes = mg.elastic_aux(m.num_cells, vs=3000.0, vp=5292.0)
    aux2, _ = mg.elastic_aux(m.num_cells, vs=2000.0, vp=3600.0)
    P1, P2 = oracle.Problem(m, aux1, sizes), oracle.Problem(m, aux2, sizes)
    ref = P1.residual(u, F1, cells=np.nonzero(m.labels == 1)[0])
    ref = P2.residual(u, F1, cells=np.nonzero(m.labels == 2)[0], out=ref)
    from pylith_amd.dm import DM
    dm = DM(m)
    for lab, aux in ((1, aux1), (2, aux2)):
        dm.add_residual(lab, 0, 0, F1)
        dm.set_auxiliary(lab, 0, aux, sizes)
    x = dev(u)
    r = torch.zeros_like(x)
    dm.compute_residual(1, 0, x, r, mode=0)
    dm.compute_residual(2, 0, x, r, mode=0)
    torch.cuda.synchronize()
    assert relerr(r.cpu().numpy(), ref) <= TOL64


def test_single_cell_and_ragged_tiles(oracle):
    for n in ((1, 1, 1), (13, 3, 2)):
        m = mg.hex_box(n, (0, 0, 0), (1.0 * n[0], 1.0 * n[1], 1.0 * n[2]))
        check_residual(oracle, m, mg.random_field(m.num_nodes, 1))
    t = mg.tet_p2_box((1, 1, 1), (0, 0, 0), (1, 1, 1), brick=1)
    check_residual(oracle, t, mg.random_field(t.num_nodes, 2))


def test_p2_tets_small(oracle):
    m = mg.tet_p2_box(8, (0, 0, 0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(m.num_cells, seed=3, vs_sigma=0.2, vp_vs=1.732)
    u = mg.sinusoid_field(m.node_coords, 17)
    check_residual(oracle, m, u, aux=(aux, sizes))


def test_p2_tets_config2_64cube(oracle):
    m = mg.tet_p2_box(64, (0, 0, 0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(m.num_cells, seed=20261202, vs_sigma=0.2, vp_vs=1.732)
    u = mg.sinusoid_field(m.node_coords, 20261202)
    check_residual(oracle, m, u, aux=(aux, sizes))


def test_deterministic_bitwise(oracle):
    m = mg.tet_p2_box(8, (0, 0, 0), (1.0, 1.0, 1.0))
    aux, sizes = mg.elastic_aux(m.num_cells)
    dm = make_dm(m, aux, sizes)
    from pylith_amd import _lib as L
    dm.set_kernel_options(L.ALGO_AUTO, L.ACCUMULATE_COLORED)
    x = dev(mg.random_field(m.num_nodes, 3))
    outs = []
    for _ in range(3):
        r = torch.zeros_like(x)
        dm.compute_residual(0, 0, x, r, mode=1)
        outs.append(r.cpu().numpy())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_pbdiag_matches_oracle(oracle):
    for m in (mg.hex_box(4, (0, 0, 0), (4.0, 4.0, 4.0), brick=2),
              mg.tet_p2_box(2, (0, 0, 0), (1.0, 1.0, 1.0), brick=1)):
        aux, sizes = mg.elastic_aux(m.num_cells)
        ref = oracle.Problem(m, aux, sizes).pbdiag(JF3, m.num_nodes)
        dm = make_dm(m, aux, sizes)
        b = torch.zeros(m.num_nodes * 9, dtype=torch.float64, device="cuda")
        dm.compute_jacobian_pbdiag(0, 0, b)
        torch.cuda.synchronize()
        assert relerr(b.cpu().numpy().reshape(-1, 3, 3), ref) <= TOL64


def test_scatters_and_reductions(oracle):
    m = mg.hex_box(4, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=2)
    aux, sizes = mg.elastic_aux(m.num_cells)
    dm = make_dm(m, aux, sizes)
    dofs, vals = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    dm.set_constraints(dofs)
    nloc, nglob = dm.sizes()
    assert nloc == 3 * m.num_nodes and nglob == nloc - dofs.size
    loc = np.random.default_rng(0).standard_normal(nloc)
    g = torch.zeros(nglob, dtype=torch.float64, device="cuda")
    dm.local_to_global(dev(loc), g, mode=0)
    ref = oracle.local_to_global(loc, dofs)
    assert np.array_equal(g.cpu().numpy(), ref)
    back = torch.zeros(nloc, dtype=torch.float64, device="cuda")
    dm.global_to_local(g, back)
    dm.insert_boundary_values(dev(vals), back)
    exp = loc.copy()
    exp[dofs] = vals
    assert np.array_equal(back.cpu().numpy(), exp)
    a = np.random.default_rng(1).standard_normal(100003)
    b = np.random.default_rng(2).standard_normal(100003)
    assert np.isclose(dm.dot(dev(a), dev(b)), a @ b, rtol=1e-13)
    assert np.isclose(dm.norm2(dev(a, torch.float32)), np.linalg.norm(a.astype(np.float32)), rtol=1e-6)


def test_bad_geometry_rejected():
    from pylith_amd.dm import DM, PlxError
    m = mg.hex_box(1, (0, 0, 0), (1.0, 1.0, 1.0))
    m.closure = m.closure[:, [1, 0, 3, 2, 5, 4, 7, 6]].copy()  # mirrored: negative determinant
    m.cone = m.closure.copy()
    with pytest.raises(PlxError):
        DM(m)


@pytest.mark.parametrize("ct", ["q1", "p2"])
def test_graph_captured_applies_match_eager(oracle, ct):
    """Residual and Jacobian action captured into one HIP graph on the handle's stream replay to
    the eager bits (colour-ordered accumulation), replay after replay, with eager work between."""
    from pylith_amd import _lib as L
    if ct == "q1":
        m = mg.hex_box((10, 9, 7), (0, 0, 0), (1000.0, 900.0, 700.0))
    else:
        m = mg.tet_p2_box(5, (0, 0, 0), (500.0, 500.0, 500.0))
    aux, sizes = mg.elastic_aux(m.num_cells, seed=2, vs_sigma=0.2, vp_vs=1.732)
    dm = make_dm(m, aux, sizes)
    dm.set_kernel_options(L.ALGO_AUTO, L.ACCUMULATE_COLORED, 0)
    x = dev(mg.random_field(m.num_nodes, 17))
    r_e, j_e = torch.zeros_like(x), torch.zeros_like(x)
    dm.compute_residual(0, 0, x, r_e, mode=1)
    dm.compute_jacobian_action(0, 0, x, j_e, mode=1)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    dm.set_stream(s)
    r, jv = torch.zeros_like(x), torch.zeros_like(x)
    with torch.cuda.stream(s):
        dm.compute_residual(0, 0, x, r, mode=1)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        dm.compute_residual(0, 0, x, r, mode=1)
        dm.compute_jacobian_action(0, 0, x, jv, mode=1)
    for k in range(4):
        r.fill_(float("nan"))
        jv.fill_(float("nan"))
        y = (x * 2.0).sum()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(r, r_e), k
        assert torch.equal(jv, j_e), k
        assert torch.isfinite(y)
    dm.set_stream(torch.cuda.current_stream())
    dm.close()


def test_q1_192cube_linearity_and_rigid_motion():
    """Full-size Q1 mesh (192^3, 21.6 M dofs; the oracle takes minutes there): size-independent
    properties of the device residual, FP64.  Linearity F(u + 2 v) = F(u) + 2 F(v), and a
    rigid-body motion (translation + rotation, zero strain) leaves every row at rounding level."""
    m = mg.hex_box(192, (0.0, 0.0, 0.0), (6000.0, 6000.0, 6000.0))
    aux, sizes = mg.elastic_aux(m.num_cells, seed=5, vs_sigma=0.2, vp_vs=1.732)
    dm = make_dm(m, aux, sizes)
    print(f"[parity] 192^3 nodes={m.num_nodes}", flush=True)

    def F(field):
        x = dev(field)
        r = torch.zeros_like(x)
        dm.compute_residual(0, 0, x, r, mode=1)
        torch.cuda.synchronize()
        return r

    u = mg.random_field(m.num_nodes, 31)
    v = mg.random_field(m.num_nodes, 32)
    fu, fv, fuv = F(u), F(v), F(u + 2.0 * v)
    scale = float(fuv.abs().max())
    assert scale > 0.0
    assert float((fuv - fu - 2.0 * fv).abs().max()) <= 1e-12 * scale
    x = m.node_coords - 3000.0
    w = np.array([1e-3, -2e-3, 5e-4]) + x @ np.array([[0.0, 2e-4, -1e-4], [-2e-4, 0.0, 3e-4], [1e-4, -3e-4, 0.0]]).T
    fw = F(np.ascontiguousarray(w).ravel())
    # a field of the same size without the rigid structure, for the scale of a real residual
    fr = F(mg.random_field(m.num_nodes, 33) * float(np.abs(w).max()) / 1e-3)
    assert float(fw.abs().max()) <= 1e-10 * float(fr.abs().max())
    dm.close()


@pytest.mark.parametrize("kind", ["q1_lexicographic", "q1_random", "p2_random"])
def test_badly_numbered_meshes_are_tiled_in_morton_order(oracle, kind):
    """A mesh numbered row by row or at random tiles with huge halos in input order; the planner
    then takes the cells in Morton order of their centroids (few halo slots per cell) and the
    device residual still matches the oracle."""
    import ctypes
    from pylith_amd import _lib as L
    if kind == "q1_lexicographic":
        brick = mg.hex_box(24, (0, 0, 0), (6000.0, 6000.0, 6000.0))
        m = mg.hex_box(24, (0, 0, 0), (6000.0, 6000.0, 6000.0), ordered=False)
    elif kind == "q1_random":
        brick = mg.hex_box(24, (0, 0, 0), (6000.0, 6000.0, 6000.0))
        m = mg.renumbered(brick, 3)
    else:
        brick = mg.tet_p2_box(12, (0, 0, 0), (6000.0, 6000.0, 6000.0))
        m = mg.renumbered(brick, 4)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=6, vs_sigma=0.2, vp_vs=1.8)

    def plan(mesh):
        d = make_dm(mesh, aux, sizes)
        tiles, halo, reordered = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int()
        d._check(d.lib.plx_debug_plan_info(d.h, 0, 0, ctypes.byref(tiles), ctypes.byref(halo), ctypes.byref(reordered)))
        return d, tiles.value, halo.value, reordered.value

    db, _, halo_brick, _ = plan(brick)
    db.close()
    dm, tiles, halo, reordered = plan(m)
    print(f"[parity] {kind}: tiles={tiles} halo slots per cell={halo / m.num_cells:.2f} (brick order "
          f"{halo_brick / m.num_cells:.2f}) reordered={reordered}", flush=True)
    # compact tiles again: within 1.5x of the halo of the brick-numbered mesh
    assert reordered == 1 and halo <= 1.5 * halo_brick
    u = mg.random_field(m.num_nodes, 9)
    ref = oracle.Problem(m, aux, sizes).residual(u, F1)
    r = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    dm.compute_residual(0, 0, dev(u), r, mode=L.INSERT_VALUES)
    torch.cuda.synchronize()
    assert relerr(r.cpu().numpy(), ref) <= TOL64
    dm.close()


def test_dmplex_vertex_order_and_tensor_point_order(oracle):
    """PyLith builds its DM from cells inverted by DMPlexInvertCell (meshio/MeshBuilder.cc:151:
    hexahedron vertices 1 and 3 swapped) and may hand per-point aux in a tensor rule's order with
    xi slowest.  The reference fixture mesh in that vertex order is refused as inverted cells
    without the order and accepted with it; per-point generalized Maxwell state that varies inside
    each cell, given in the xi-slowest order, gives the oracle's residual, and the updated state
    comes back in the caller's order."""
    import dataclasses
    from pylith_amd.dm import DM, PlxError
    m, groups = meshio_ascii.read(os.path.join(GOLDEN, "hex.mesh"))
    inv = np.ascontiguousarray(m.closure[:, [0, 3, 2, 1, 4, 5, 6, 7]])  # DMPlexInvertCell
    m_dm = dataclasses.replace(m, closure=inv, cone=inv.copy())
    with pytest.raises(PlxError):
        DM(m_dm)  # reflected cells: non-positive Jacobians
    vperm = [0, 3, 2, 1, 4, 5, 6, 7]
    qperm = np.array([0, 4, 2, 6, 1, 5, 3, 7])  # our q = i + 2j + 4k is the caller's 4i + 2j + k
    rng = np.random.default_rng(31)
    aux, sizes = mg.genmax_aux(m.num_cells, 8, seed=4, log_eta_range=(1e17, 1e19), ratios=(0.3, 0.2, 0.1))
    aux[:, 9:33] = rng.standard_normal((aux.shape[0], 24)) * 1e-4  # state varies point by point
    caller = aux.reshape(m.num_cells, 8, -1)[:, np.argsort(qperm)].reshape(aux.shape)
    assert np.array_equal(caller.reshape(m.num_cells, 8, -1)[:, qperm], aux.reshape(m.num_cells, 8, -1))
    dt = 0.1 * mg.YEAR
    dm = DM(m_dm, vertex_perm=vperm)
    dm.set_point_order(qperm)
    dm.add_residual(24, 0, 0, 3)
    dm.add_jacobian(24, 0, 0, 0, 0, 4)
    dm.set_auxiliary(24, 0, np.ascontiguousarray(caller), sizes, per_qp=True)
    dm.set_constants(24, 0, [dt])
    u = mg.random_field(m.num_nodes, 12)
    cells = np.nonzero(m.labels == 24)[0]
    P = oracle.Problem(m, aux, sizes, per_qp=True, constants=[dt])
    ref = P.residual(u, 3, cells=cells)
    r = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    dm.compute_residual(24, 0, dev(u), r, mode=1)
    torch.cuda.synchronize()
    assert relerr(r.cpu().numpy(), ref) <= TOL64
    # the state update returns per-point records in the caller's order
    dm.update_state_vars(24, 0, dev(u))
    got = dm.get_auxiliary(24, 0, caller)
    new = P.update_state(u, 5, 5)
    new[:, 27:33] = P.update_state(u, 6, 6)[:, 27:33]
    new_caller = new.reshape(m.num_cells, 8, -1)[:, np.argsort(qperm)].reshape(new.shape)
    sel = np.repeat(np.isin(np.arange(m.num_cells), cells), 8)
    assert relerr(got[sel][:, 9:33], new_caller[sel][:, 9:33]) <= TOL64
    dm.close()
