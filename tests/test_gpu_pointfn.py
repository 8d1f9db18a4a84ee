"""Pointwise kernels in the PetscPointFn / PetscPointJacFn form on the device, and the assembled
Jacobian (DMPlexComputeJacobianByKey, IntegratorDomain.cc:489-490), against the oracle (FP64,
1e-12 relative).

* The library's twins of the reference kernels written in that form (PLX_ALGO_POINTFN): the
  elastic, reference-state, body-force and gravity residuals and the elastic Jacobian action on
  non-affine Q1 and on P2 meshes, through the tile path and through the colour passes (bitwise
  reproducible).
* Test-only kernels in a library of their own (tests/cpp/pointfn_test_kernels.hip), compiled
  against the public header only and registered through plx_dm_add_*_pointfn: residual, action and
  assembled Jacobian, reading the point x, time, constants, solution values and aux offsets.
* The assembled Jacobian of built-in keys (two materials adding into one matrix, P2, generalized
  Maxwell) equals the oracle's dense Jacobian block for block, and its product with a vector equals
  the matrix-free action.
"""
import ctypes
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import _lib as L  # noqa: E402
from pylith_amd import meshgen as mg  # noqa: E402

pytestmark = pytest.mark.gpu
F1, JF3, F1_GM, JF3_GM = 1, 2, 3, 4
F0_BODY, F0_GRAV, F0_GRAVBODY, F1_REF = 7, 8, 9, 10
T_F0, T_F1, T_JF0, T_JF3 = 90, 91, 92, 93
TOL = 1e-12
TEST_LIB = os.path.join(os.path.dirname(__file__), "cpp", "libpointfn_test.so")


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def q1_perturbed(n=(5, 4, 3), seed=2):
    m = mg.hex_box(n, (0, 0, 0), (5000.0, 4000.0, 3000.0), brick=2, shear=0.15)
    rng = np.random.default_rng(seed)
    x = m.coords + rng.uniform(-80, 80, m.coords.shape)
    m.coords = m.node_coords = x
    return m


def p2_mesh(n=3):
    return mg.tet_p2_box(n, (0, 0, 0), (3000.0, 3000.0, 3000.0), brick=1)


def dense_jacobian(P, n, jf3, jf0=0, t=0.0):
    J = np.zeros((n, n))
    e = np.zeros(n)
    for i in range(n):
        e[i] = 1.0
        J[:, i] = P.jacobian_action(e, jf3, jf0=jf0, t=t)
        e[i] = 0.0
    return J


def bsr_to_dense(row_ptr, col, values, num_nodes):
    n = 3 * num_nodes
    D = np.zeros((n, n))
    for i in range(num_nodes):
        for e in range(row_ptr[i], row_ptr[i + 1]):
            j = col[e]
            D[3 * i:3 * i + 3, 3 * j:3 * j + 3] += values[e]
    return D


@pytest.mark.parametrize("kind", ["q1", "p2"])
def test_builtin_kernels_in_pointfn_form(oracle, kind):
    from pylith_amd.dm import DM
    m = q1_perturbed() if kind == "q1" else p2_mesh()
    nn = m.num_nodes
    rng = np.random.default_rng(11)
    u = rng.standard_normal(3 * nn) * 1e-3
    print(f"[pointfn builtin] {kind} cells={m.num_cells}", flush=True)
    # elastic residual + Jacobian action
    aux, sizes = mg.elastic_aux(m.num_cells, seed=5, vs_sigma=0.2, vp_vs=1.8)
    P = oracle.Problem(m, aux, sizes)
    dm = DM(m)
    dm.add_residual(0, 0, 0, F1)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(0, 0, aux, sizes)
    ref, jref = P.residual(u, F1), P.jacobian_action(u, JF3)
    # colour passes (bitwise reproducible) and the tile path (LDS accumulation, halo slots)
    for acc in (L.ACCUMULATE_COLORED, L.ACCUMULATE_AUTO):
        dm.set_kernel_options(L.ALGO_POINTFN, acc, 0)
        r = torch.zeros(3 * nn, dtype=torch.float64, device="cuda")
        dm.compute_residual(0, 0, dev(u), r, mode=L.INSERT_VALUES)
        r2 = torch.full_like(r, 3.0)
        dm.compute_residual(0, 0, dev(u), r2, mode=L.INSERT_VALUES)
        jv = torch.full_like(r, 1.0)
        dm.compute_jacobian_action(0, 0, dev(u), jv, mode=L.ADD_VALUES)
        torch.cuda.synchronize()
        assert relerr(r.cpu().numpy(), ref) <= TOL, acc
        if acc == L.ACCUMULATE_COLORED:
            assert torch.equal(r, r2)  # colour-ordered: the same bits every call
        else:
            assert relerr(r2.cpu().numpy(), ref) <= TOL
        assert relerr(jv.cpu().numpy() - 1.0, jref) <= TOL, acc
    dm.close()
    # reference state with gravity and body force: aux [density, body_force, gravity, ref stress,
    # ref strain, mu, K]
    base, _ = mg.elastic_aux(m.num_cells, seed=6, vs_sigma=0.1, vp_vs=1.8)
    nc = m.num_cells
    aux = np.concatenate([base[:, :1], rng.standard_normal((nc, 3)) * 1e3, np.tile([0.0, 0.0, -9.8], (nc, 1)),
                          rng.standard_normal((nc, 6)) * 1e6, rng.standard_normal((nc, 6)) * 1e-4, base[:, 1:]], 1)
    sizes = np.array([1, 3, 3, 6, 6, 1, 1], np.int32)
    P = oracle.Problem(m, aux, sizes)
    dm = DM(m)
    dm.add_residual(0, 0, F0_GRAVBODY, F1_REF)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(0, 0, aux, sizes)
    dm.set_kernel_options(L.ALGO_POINTFN, 0, 0)
    r = torch.zeros(3 * nn, dtype=torch.float64, device="cuda")
    dm.compute_residual(0, 0, dev(u), r, mode=L.INSERT_VALUES)
    torch.cuda.synchronize()
    assert relerr(r.cpu().numpy(), P.residual(u, F1_REF, f0=F0_GRAVBODY)) <= TOL
    # the tile kernels agree with the PetscPointFn form
    dm.set_kernel_options(L.ALGO_AUTO, 0, 0)
    rt = torch.zeros_like(r)
    dm.compute_residual(0, 0, dev(u), rt, mode=L.INSERT_VALUES)
    torch.cuda.synchronize()
    assert relerr(rt.cpu().numpy(), r.cpu().numpy()) <= TOL
    dm.close()


def test_launchers_of_the_test_library_exist():
    if not os.path.exists(TEST_LIB):
        pytest.fail("tests/cpp/libpointfn_test.so is not built (run __graft_entry__.build())")


def load_launchers():
    lib = ctypes.CDLL(TEST_LIB)
    out = (ctypes.c_void_p * 3)()
    assert lib.plx_test_pointfn_launchers(out) == 0
    return [out[0], out[1], out[2]], lib


@pytest.mark.parametrize("kind", ["q1", "p2"])
def test_caller_kernels_through_launchers(oracle, kind):
    """Kernels without a built-in twin, written in PETSc's form in another library."""
    from pylith_amd.dm import DM
    launchers, keep = load_launchers()
    m = q1_perturbed((4, 3, 3)) if kind == "q1" else p2_mesh(2)
    nn, n = m.num_nodes, 3 * m.num_nodes
    aux, sizes = mg.elastic_aux(m.num_cells, seed=8, vs_sigma=0.2, vp_vs=1.8)
    consts = [2.5e-3, 3.0e7, 0.3]
    t = 0.7
    dm = DM(m)
    dm.add_jacobian_pointfn(0, 0, launchers[1], launchers[2])
    dm.add_residual_pointfn(0, 0, launchers[0])
    dm.set_auxiliary(0, 0, aux, sizes)
    dm.set_constants(0, 0, consts)
    P = oracle.Problem(m, aux, sizes, constants=consts)
    rng = np.random.default_rng(12)
    u = rng.standard_normal(n) * 1e-3
    print(f"[pointfn caller] {kind} cells={m.num_cells}", flush=True)
    r = torch.zeros(n, dtype=torch.float64, device="cuda")
    dm.compute_residual(0, 0, dev(u), r, mode=L.INSERT_VALUES, t=t)
    jv = torch.zeros_like(r)
    dm.compute_jacobian_action(0, 0, dev(u), jv, mode=L.INSERT_VALUES, t=t)
    torch.cuda.synchronize()
    ref = P.residual(u, T_F1, f0=T_F0, t=t)
    assert relerr(r.cpu().numpy(), ref) <= TOL
    jref = P.jacobian_action(u, T_JF3, jf0=T_JF0, t=t)
    assert relerr(jv.cpu().numpy(), jref) <= TOL
    # with a time derivative (the test f0 reads s_t): tile path and colour passes
    ut = rng.standard_normal(n) * 1e-2
    ref_t = P.residual(u, T_F1, f0=T_F0, t=t, locX_t=ut)
    for acc in (L.ACCUMULATE_AUTO, L.ACCUMULATE_COLORED):
        dm.set_kernel_options(L.ALGO_AUTO, acc, 0)
        rt = torch.zeros_like(r)
        dm.compute_residual(0, 0, dev(u), rt, mode=L.INSERT_VALUES, t=t, locX_t=dev(ut))
        torch.cuda.synchronize()
        assert relerr(rt.cpu().numpy(), ref_t) <= TOL, acc
    dm.set_kernel_options(L.ALGO_AUTO, L.ACCUMULATE_AUTO, 0)
    # residual is affine in u: F(u) - F(0) = J u
    r0 = torch.zeros_like(r)
    dm.compute_residual(0, 0, torch.zeros_like(r), r0, mode=L.INSERT_VALUES, t=t)
    torch.cuda.synchronize()
    assert relerr((r - r0).cpu().numpy(), jref) <= 1e-10
    # assembled Jacobian against the oracle's dense Jacobian
    row_ptr, col = dm.jacobian_pattern()
    vals = torch.zeros(col.size, 3, 3, dtype=torch.float64, device="cuda")
    dm.compute_jacobian(0, 0, vals, mode=L.INSERT_VALUES, t=t, locX=dev(u))
    torch.cuda.synchronize()
    D = bsr_to_dense(row_ptr, col, vals.cpu().numpy(), nn)
    Jd = dense_jacobian(P, n, T_JF3, jf0=T_JF0, t=t)
    assert np.abs(D - Jd).max() <= TOL * np.abs(Jd).max()
    del keep


def test_assembled_jacobian_builtin_keys(oracle):
    from pylith_amd.dm import DM
    # two materials adding into one matrix; non-affine Q1
    m = q1_perturbed((5, 4, 4), seed=9)
    labels = (np.arange(m.num_cells) % 3 == 0).astype(np.int32)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=10, vs_sigma=0.3, vp_vs=1.75)
    dm = DM(m, labels=labels)
    for lab in (0, 1):
        dm.add_residual(lab, 0, 0, F1)
        dm.add_jacobian(lab, 0, 0, 0, 0, JF3)
        dm.set_auxiliary(lab, 0, aux, sizes)
    row_ptr, col = dm.jacobian_pattern()
    # every pair of nodes of a cell is in the pattern, nothing else
    pairs = {(int(a), int(b)) for cl in m.closure for a in cl for b in cl}
    got = {(i, int(col[e])) for i in range(m.num_nodes) for e in range(row_ptr[i], row_ptr[i + 1])}
    assert got == pairs
    vals = torch.full((col.size, 3, 3), 7.0, dtype=torch.float64, device="cuda")
    dm.compute_jacobian(0, 0, vals, mode=L.INSERT_VALUES)
    dm.compute_jacobian(1, 0, vals, mode=L.ADD_VALUES)
    torch.cuda.synchronize()
    n = 3 * m.num_nodes
    P = oracle.Problem(m, aux, sizes)
    Jd = dense_jacobian(P, n, JF3)
    D = bsr_to_dense(row_ptr, col, vals.cpu().numpy(), m.num_nodes)
    assert np.abs(D - Jd).max() <= TOL * np.abs(Jd).max()
    # the product with a vector equals the matrix-free action of both keys
    v = np.random.default_rng(3).standard_normal(n)
    jv = torch.zeros(n, dtype=torch.float64, device="cuda")
    dm.compute_jacobian_action(0, 0, dev(v), jv, mode=L.INSERT_VALUES)
    dm.compute_jacobian_action(1, 0, dev(v), jv, mode=L.ADD_VALUES)
    torch.cuda.synchronize()
    assert relerr(D @ v, jv.cpu().numpy()) <= TOL
    dm.close()
    # P2 and generalized Maxwell (effective shear modulus of the time step)
    m = p2_mesh(2)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=4, vs_sigma=0.2, vp_vs=1.8)
    dm = DM(m)
    dm.add_residual(0, 0, 0, F1)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(0, 0, aux, sizes)
    row_ptr, col = dm.jacobian_pattern()
    vals = torch.zeros(col.size, 3, 3, dtype=torch.float64, device="cuda")
    dm.compute_jacobian(0, 0, vals, mode=L.INSERT_VALUES)
    torch.cuda.synchronize()
    Jd = dense_jacobian(oracle.Problem(m, aux, sizes), 3 * m.num_nodes, JF3)
    D = bsr_to_dense(row_ptr, col, vals.cpu().numpy(), m.num_nodes)
    assert np.abs(D - Jd).max() <= TOL * np.abs(Jd).max()
    dm.close()
    m = mg.hex_box((3, 3, 2), (0, 0, 0), (3000.0, 3000.0, 2000.0), brick=2, shear=0.2)
    aux, sizes = mg.genmax_aux(m.num_cells, 8, seed=5, log_eta_range=(1e17, 1e19), ratios=(0.3, 0.2, 0.1))
    dt = 0.1 * mg.YEAR
    dm = DM(m)
    dm.add_residual(0, 0, 0, F1_GM)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3_GM)
    dm.set_auxiliary(0, 0, aux, sizes, per_qp=True)
    dm.set_constants(0, 0, [dt])
    row_ptr, col = dm.jacobian_pattern()
    vals = torch.zeros(col.size, 3, 3, dtype=torch.float64, device="cuda")
    dm.compute_jacobian(0, 0, vals, mode=L.INSERT_VALUES)
    torch.cuda.synchronize()
    Pg = oracle.Problem(m, aux, sizes, per_qp=True, constants=[dt])
    Jd = dense_jacobian(Pg, 3 * m.num_nodes, JF3_GM)
    D = bsr_to_dense(row_ptr, col, vals.cpu().numpy(), m.num_nodes)
    assert np.abs(D - Jd).max() <= TOL * np.abs(Jd).max()
    dm.close()
