"""Host set-up of the smoothed-aggregation multigrid (pylith_amd/csrc/plx_amg.cpp, the GAMG role of
materials/Elasticity.cc:307-318) on the merged operator of a faulted Q1 box assembled from the
oracle's element matrices, through the g++-built shim tests/cpp/libamg_probe.so.  No GPU.

Pins (each fails for a plausible set-up mistake):
* R is P transposed and every coarse operator is the Galerkin product R A P (symmetric);
* the stored inverse diagonal blocks invert the diagonal blocks; lambda_max bounds D^-1 A;
* the prolongator reproduces all six rigid-body modes on rows away from the Dirichlet nodes
  (a wrong sign or index in the rotation modes, or a broken per-aggregate QR, fails this);
* the V-cycle is symmetric positive definite and cuts PCG iterations several-fold;
* on a P2 tet box with Dirichlet faces that fix one component of a node (a diagonal block with
  an identity row beside stiffness rows ~1e14), the stored inverse blocks still invert the
  diagonal blocks, lambda_max stays O(1), and the V-cycle over strong couplings (theta 0.1)
  cuts PCG iterations more than tenfold against point-block Jacobi;
* the P2 form with a P2 -> P1 first level: vertex rows of P are identity blocks to their vertex
  and edge rows half blocks to both ends (zero at fixed dofs), R = P^T, the P1 operator is the
  Galerkin product P^T A P, and the V-cycle is SPD and cuts PCG iterations more than tenfold.
"""
import ctypes
import os

import numpy as np
import pytest

scipy_sparse = pytest.importorskip("scipy.sparse")
sp = scipy_sparse

from pylith_amd import meshgen as mg  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "libamg_probe.so")
JF3 = 2


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(LIB):
        from pylith_amd import build
        build.build_amg_probe()
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.amg_probe_build.argtypes = [ctypes.c_int, P, P, P, P, P, ctypes.c_int, ctypes.c_int]
    lib.amg_probe_level.argtypes = [ctypes.c_int, P, P]
    lib.amg_probe_array.argtypes = [ctypes.c_int, ctypes.c_int, P]
    lib.amg_probe_coarse.argtypes = [P]
    lib.amg_probe_vcycle.argtypes = [P, P]
    lib.amg_probe_error.restype = ctypes.c_char_p
    lib.amg_probe_set_theta.argtypes = [ctypes.c_double]
    lib.amg_probe_build_p2.argtypes = [ctypes.c_int, P, P, P, P, ctypes.c_int, P, P, P, P, ctypes.c_int,
                                       ctypes.c_int]
    lib.amg_probe_level_bc.argtypes = [ctypes.c_int]
    lib.amg_probe_build_p2f.argtypes = lib.amg_probe_build_p2.argtypes
    return lib


@pytest.fixture(scope="module")
def problem(oracle):
    """Merged operator of a sheared 10^3 box with a dipping fault (positive split nodes and the
    x-face nodes are identity rows), as the Green's-function assembly builds it."""
    n, L = 10, 6000.0
    shear = 0.5
    m, f = mg.hex_box_with_fault(n, (0, 0, -L), (L, L, 0), n // 2, brick=4, shear=shear)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=3, vs_sigma=0.2, vp_vs=1.732)
    nn = m.num_nodes
    merged = np.arange(nn)
    merged[f.pos_node] = f.neg_node
    cl = merged[m.closure]
    # the cells are congruent parallelepipeds: K_cell = lambda K_lambda + mu K_mu
    Pmu = oracle.Problem(m, np.tile([2500.0, 1.0, 2.0 / 3.0], (m.num_cells, 1)), sizes)
    Plam = oracle.Problem(m, np.tile([2500.0, 0.0, 1.0], (m.num_cells, 1)), sizes)
    Kmu, Klam = -Pmu.element_matrix(0, JF3), -Plam.element_matrix(0, JF3)
    mu = aux[:, 1]
    lam = aux[:, 2] - 2.0 / 3.0 * mu
    d = (3 * cl[:, :, None] + np.arange(3)).reshape(m.num_cells, 24)
    rows = np.repeat(d, 24, axis=1).ravel()
    cols = np.tile(d, (1, 24)).ravel()
    vals = (lam[:, None] * Klam.ravel()[None, :] + mu[:, None] * Kmu.ravel()[None, :]).ravel()
    A = sp.csr_matrix((vals, (rows, cols)), shape=(3 * nn, 3 * nn))
    x0 = m.node_coords[:, 0] - shear * m.node_coords[:, 2]
    fixed_nodes = np.nonzero(np.isclose(x0, 0.0) | np.isclose(x0, L))[0]
    fixed = np.zeros(3 * nn, bool)
    fixed[(3 * fixed_nodes[:, None] + np.arange(3)).ravel()] = True
    fixed[(3 * f.pos_node[:, None] + np.arange(3)).ravel()] = True
    keep = sp.diags((~fixed).astype(float))
    A = (keep @ A @ keep + sp.diags(fixed.astype(float))).tocsr()
    # 3x3 block pattern of the node graph (the identity rows keep their diagonal block only)
    Ab = sp.bsr_matrix(A, blocksize=(3, 3))
    Ab.sort_indices()
    return dict(m=m, f=f, A=A, Ab=Ab, fixed=fixed, nn=nn, fixed_nodes=fixed_nodes)


@pytest.fixture(scope="module")
def hierarchy(probe, problem):
    Ab, m = problem["Ab"], problem["m"]
    ptr = Ab.indptr.astype(np.int32)
    col = Ab.indices.astype(np.int32)
    val = np.ascontiguousarray(Ab.data, np.float64)
    coords = np.ascontiguousarray(m.node_coords, np.float64)
    fx = problem["fixed"].astype(np.uint8)
    nl = probe.amg_probe_build(problem["nn"], ptr.ctypes.data, col.ctypes.data, val.ctypes.data,
                               coords.ctypes.data, fx.ctypes.data, 300, 8)
    assert nl >= 1, probe.amg_probe_error()

    def arr(l, which, dtype):
        cnt = probe.amg_probe_array(l, which, None)
        out = np.zeros(cnt, dtype)
        probe.amg_probe_array(l, which, out.ctypes.data)
        return out

    levels = []
    for l in range(nl):
        info = np.zeros(3, np.int32)
        lmax = ctypes.c_double()
        probe.amg_probe_level(l, info.ctypes.data, ctypes.byref(lmax))
        n, b, nc = (int(v) for v in info)
        lv = dict(n=n, b=b, nc=nc, lmax=lmax.value)
        lv["P"] = sp.bsr_matrix((arr(l, 5, np.float64).reshape(-1, b, 6), arr(l, 4, np.int32),
                                 arr(l, 3, np.int32)), shape=(n * b, nc * 6)).tocsr()
        lv["R"] = sp.bsr_matrix((arr(l, 8, np.float64).reshape(-1, 6, b), arr(l, 7, np.int32),
                                 arr(l, 6, np.int32)), shape=(nc * 6, n * b)).tocsr()
        lv["Dinv"] = arr(l, 9, np.float64).reshape(n, b, b)
        if l == 0:
            lv["A"] = problem["A"]
        else:
            lv["A"] = sp.bsr_matrix((arr(l, 2, np.float64).reshape(-1, 6, 6), arr(l, 1, np.int32),
                                     arr(l, 0, np.int32)), shape=(n * 6, n * 6)).tocsr()
        levels.append(lv)
    cn = probe.amg_probe_coarse_n()
    cinv = np.zeros(cn * cn)
    probe.amg_probe_coarse(cinv.ctypes.data)
    return dict(levels=levels, coarse_inv=cinv.reshape(cn, cn))


def test_galerkin_coarse_operators(hierarchy):
    lv = hierarchy["levels"]
    for l, L in enumerate(lv):
        assert abs(L["R"] - L["P"].T).max() == 0.0
        Ac = (L["R"] @ L["A"] @ L["P"]).toarray()
        nxt = lv[l + 1]["A"].toarray() if l + 1 < len(lv) else np.linalg.inv(hierarchy["coarse_inv"])
        # the set-up writes 1 on the diagonal of dummy (rank-deficient) coarse dofs
        dummy = np.abs(Ac).sum(1) == 0.0
        Ac[dummy, dummy] = 1.0
        scale = np.abs(Ac).max()
        assert np.abs(nxt - Ac).max() <= 1e-9 * scale, l
        assert np.abs(nxt - nxt.T).max() <= 1e-9 * scale, l


def test_inverse_diagonal_blocks_and_spectrum_bound(hierarchy):
    for L in hierarchy["levels"]:
        b = L["b"]
        A = sp.bsr_matrix(L["A"], blocksize=(b, b))
        D = np.zeros((L["n"], b, b))
        for i in range(L["n"]):
            for k in range(A.indptr[i], A.indptr[i + 1]):
                if A.indices[k] == i:
                    D[i] = A.data[k]
        prod = np.einsum("nij,njk->nik", L["Dinv"], D)
        ok = np.abs(np.linalg.det(D)) > 1e-12 * np.abs(D).max()
        assert np.allclose(prod[ok], np.eye(b), atol=1e-10)
        # lambda_max(D^-1 A) <= lmax <= 1.2 x it (the set-up takes 1.1 x a power-iteration estimate)
        Dinv = sp.block_diag(list(L["Dinv"])).tocsr()
        ev = np.linalg.eigvals((Dinv @ L["A"]).toarray())
        top = ev.real.max()
        assert top <= L["lmax"] <= 1.2 * top


def test_prolongator_reproduces_rigid_body_modes(problem, hierarchy):
    m, fixed = problem["m"], problem["fixed"]
    P = hierarchy["levels"][0]["P"]
    x = m.node_coords - m.node_coords.mean(0)
    nn = problem["nn"]
    modes = []
    for c in range(3):  # translations
        t = np.zeros((nn, 3))
        t[:, c] = 1.0
        modes.append(t.ravel())
    for w in np.eye(3):  # rotations w x x
        modes.append(np.cross(w, x).ravel())
    # rows whose stencil holds no fixed dof (there A m = 0 for a rigid mode m, so the smoothed
    # prolongator keeps the tentative one's exact reproduction); positive split nodes are merged
    A = problem["A"]
    merged = np.arange(nn)
    merged[problem["f"].pos_node] = problem["f"].neg_node
    cl = merged[m.closure]
    is_fixed_node = np.zeros(nn, bool)
    is_fixed_node[problem["fixed_nodes"]] = True
    near = np.zeros(nn, bool)
    near[np.unique(cl[is_fixed_node[cl].any(1)])] = True
    touched = np.repeat(near, 3)
    rows = np.nonzero(~fixed & ~touched)[0]
    assert np.abs(A[rows] @ (modes_check := np.tile([1.0, 0.0, 0.0], nn) * ~fixed)).max() <= 1e-6 * abs(A).max()
    assert rows.size > 0.3 * A.shape[0]
    Pd = P.toarray()
    for k, mode in enumerate(modes):
        mvec = mode * (~fixed)
        c, *_ = np.linalg.lstsq(Pd[rows], mvec[rows], rcond=None)
        err = np.abs(Pd[rows] @ c - mvec[rows]).max() / np.abs(mvec[rows]).max()
        assert err <= 1e-10, (k, err)


def test_vcycle_is_spd_and_cuts_iterations(probe, problem, hierarchy):
    A, fixed = problem["A"], problem["fixed"]
    n = A.shape[0]

    def M(r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        probe.amg_probe_vcycle(r.ctypes.data, z.ctypes.data)
        return z

    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(n) * ~fixed, rng.standard_normal(n) * ~fixed
    assert abs(u @ M(v) - v @ M(u)) <= 1e-11 * abs(u @ M(u))
    for _ in range(5):
        w = rng.standard_normal(n) * ~fixed
        assert w @ M(w) > 0.0

    def pcg(prec, b, rtol=1e-8, maxit=2000):
        x = np.zeros_like(b)
        r = b.copy()
        z = prec(r)
        p = z.copy()
        rz = r @ z
        b0 = np.linalg.norm(b)
        for it in range(maxit):
            q = A @ p
            a = rz / (p @ q)
            x += a * p
            r -= a * q
            if np.linalg.norm(r) <= rtol * b0:
                return x, it + 1
            z = prec(r)
            rzn = r @ z
            p = z + rzn / rz * p
            rz = rzn
        return x, maxit

    b = rng.standard_normal(n) * ~fixed
    Ab = sp.bsr_matrix(A, blocksize=(3, 3))
    D = np.zeros((n // 3, 3, 3))
    for i in range(n // 3):
        for k in range(Ab.indptr[i], Ab.indptr[i + 1]):
            if Ab.indices[k] == i:
                D[i] = Ab.data[k]
    Dinv = np.linalg.inv(D)
    jac = lambda r: np.einsum("nij,nj->ni", Dinv, r.reshape(-1, 3)).ravel()  # noqa: E731
    xa, ita = pcg(M, b)
    xj, itj = pcg(jac, b)
    assert np.linalg.norm(A @ xa - b) <= 1e-8 * np.linalg.norm(b)
    assert ita * 3 <= itj, (ita, itj)
    assert ita <= 25, ita


def _pcg(A, prec, b, rtol=1e-8, maxit=3000):
    x = np.zeros_like(b)
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    b0 = np.linalg.norm(b)
    for it in range(maxit):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= rtol * b0:
            return x, it + 1
        z = prec(r)
        rzn = r @ z
        p = z + rzn / rz * p
        rz = rzn
    return x, maxit


def test_p2_operator_with_partly_fixed_nodes(probe, oracle):
    """Runs last: it rebuilds the probe's hierarchy."""
    m = mg.tet_p2_box(8, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=2)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=4, vs_sigma=0.2, vp_vs=1.732)
    dofs, _ = mg.shear_bcs(m.node_coords, 0.0, 6000.0)  # +x face: only u_y fixed
    P = oracle.Problem(m, aux, sizes)
    nn = m.num_nodes
    rows, cols, vals = [], [], []
    for c in range(m.num_cells):
        d = (3 * m.closure[c][:, None] + np.arange(3)).ravel()
        rows.append(np.repeat(d, d.size))
        cols.append(np.tile(d, d.size))
        vals.append(-P.element_matrix(c, JF3).ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * nn, 3 * nn))
    fixed = np.zeros(3 * nn, bool)
    fixed[dofs] = True
    keep = sp.diags((~fixed).astype(float))
    A = (keep @ A @ keep + sp.diags(fixed.astype(float))).tocsr()
    Ab = sp.bsr_matrix(A, blocksize=(3, 3))
    Ab.sort_indices()
    ptr, col = Ab.indptr.astype(np.int32), Ab.indices.astype(np.int32)
    val = np.ascontiguousarray(Ab.data, np.float64)
    coords = np.ascontiguousarray(m.node_coords, np.float64)
    fx = fixed.astype(np.uint8)
    probe.amg_probe_set_theta(0.1)
    try:
        nl = probe.amg_probe_build(nn, ptr.ctypes.data, col.ctypes.data, val.ctypes.data, coords.ctypes.data,
                                   fx.ctypes.data, 768, 8)
    finally:
        probe.amg_probe_set_theta(0.0)
    assert nl >= 1, probe.amg_probe_error()
    D = np.zeros((nn, 3, 3))
    for i in range(nn):
        for k in range(Ab.indptr[i], Ab.indptr[i + 1]):
            if Ab.indices[k] == i:
                D[i] = Ab.data[k]
    partly = fixed.reshape(-1, 3).any(axis=1) & ~fixed.reshape(-1, 3).all(axis=1)
    assert partly.sum() > 0 and np.abs(D[partly]).max() > 1e13  # the blocks of the regression
    cnt = probe.amg_probe_array(0, 9, None)
    Dinv = np.zeros(cnt)
    probe.amg_probe_array(0, 9, Dinv.ctypes.data)
    Dinv = Dinv.reshape(nn, 3, 3)
    eye = np.einsum("nij,njk->nik", Dinv, D)
    assert np.abs(eye - np.eye(3)[None]).max() <= 1e-10
    info = np.zeros(3, np.int32)
    lmax = ctypes.c_double()
    probe.amg_probe_level(0, info.ctypes.data, ctypes.byref(lmax))
    assert 1.0 < lmax.value < 5.0, lmax.value

    def M(r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        probe.amg_probe_vcycle(r.ctypes.data, z.ctypes.data)
        return z

    rng = np.random.default_rng(1)
    b = rng.standard_normal(3 * nn) * ~fixed
    xa, ita = _pcg(A, M, b)
    xj, itj = _pcg(A, lambda r: np.einsum("nij,nj->ni", Dinv, r.reshape(-1, 3)).ravel(), b)
    assert np.linalg.norm(A @ xa - b) <= 1e-8 * np.linalg.norm(b)
    assert ita * 10 <= itj, (ita, itj)



def _p2_operator(oracle, n):
    m = mg.tet_p2_box(n, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=2)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=4, vs_sigma=0.2, vp_vs=1.732)
    dofs, _ = mg.shear_bcs(m.node_coords, 0.0, 6000.0)
    P = oracle.Problem(m, aux, sizes)
    nn = m.num_nodes
    rows, cols, vals = [], [], []
    for c in range(m.num_cells):
        d = (3 * m.closure[c][:, None] + np.arange(3)).ravel()
        rows.append(np.repeat(d, d.size))
        cols.append(np.tile(d, d.size))
        vals.append(-P.element_matrix(c, JF3).ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * nn, 3 * nn))
    fixed = np.zeros(3 * nn, bool)
    fixed[dofs] = True
    keep = sp.diags((~fixed).astype(float))
    A = (keep @ A @ keep + sp.diags(fixed.astype(float))).tocsr()
    return m, A, fixed


def test_p2_to_p1_first_level(probe, oracle):
    """Runs last: it rebuilds the probe's hierarchy."""
    m, A, fixed = _p2_operator(oracle, 6)
    nn, nv = m.num_nodes, m.num_vertices
    par = [None] * nn
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    for c in range(m.num_cells):
        cl, cn = m.closure[c], m.cone[c]
        for b in range(4):
            par[cl[b]] = ([int(cn[b])], [1.0])
        for e, (a, b) in enumerate(edges):
            par[cl[4 + e]] = (sorted([int(cn[a]), int(cn[b])]), [0.5, 0.5])
    par_ptr = np.zeros(nn + 1, np.int32)
    for i in range(nn):
        par_ptr[i + 1] = par_ptr[i] + len(par[i][0])
    pidx = np.concatenate([np.array(p[0], np.int32) for p in par])
    pw = np.concatenate([np.array(p[1], float) for p in par])
    Ab = sp.bsr_matrix(A, blocksize=(3, 3))
    Ab.sort_indices()
    ptr, col = Ab.indptr.astype(np.int32), Ab.indices.astype(np.int32)
    val = np.ascontiguousarray(Ab.data, np.float64)
    fx = fixed.astype(np.uint8)
    vxyz = np.ascontiguousarray(m.coords, np.float64)
    probe.amg_probe_set_theta(0.05)
    try:
        nl = probe.amg_probe_build_p2(nn, ptr.ctypes.data, col.ctypes.data, val.ctypes.data, fx.ctypes.data, nv,
                                      par_ptr.ctypes.data, pidx.ctypes.data, pw.ctypes.data, vxyz.ctypes.data, 300, 8)
    finally:
        probe.amg_probe_set_theta(0.0)
    assert nl >= 2, probe.amg_probe_error()
    assert probe.amg_probe_level_bc(0) == 3 and probe.amg_probe_level_bc(1) == 6

    def arr(l, which, dtype):
        cnt = probe.amg_probe_array(l, which, None)
        out = np.zeros(cnt, dtype)
        probe.amg_probe_array(l, which, out.ctypes.data)
        return out

    Pm = sp.bsr_matrix((arr(0, 5, np.float64).reshape(-1, 3, 3), arr(0, 4, np.int32), arr(0, 3, np.int32)),
                       shape=(3 * nn, 3 * nv)).tocsr()
    Rm = sp.bsr_matrix((arr(0, 8, np.float64).reshape(-1, 3, 3), arr(0, 7, np.int32), arr(0, 6, np.int32)),
                       shape=(3 * nv, 3 * nn)).tocsr()
    assert abs(Rm - Pm.T).max() == 0.0
    # the P1 interpolation, zero at fixed dofs (fine or coarse)
    fixed1 = np.zeros(3 * nv, bool)
    for i in range(nn):
        if len(par[i][0]) == 1:
            fixed1[3 * par[i][0][0]:3 * par[i][0][0] + 3] |= fixed[3 * i:3 * i + 3]
    Pd = np.zeros((3 * nn, 3 * nv))
    for i in range(nn):
        for I, w in zip(*par[i]):
            for c in range(3):
                if not fixed[3 * i + c] and not fixed1[3 * I + c]:
                    Pd[3 * i + c, 3 * I + c] = w
    assert np.abs(Pm.toarray() - Pd).max() == 0.0
    # the P1 operator: Galerkin product, identity rows / zero columns at fixed vertex dofs
    A1 = sp.bsr_matrix((arr(1, 2, np.float64).reshape(-1, 3, 3), arr(1, 1, np.int32), arr(1, 0, np.int32)),
                       shape=(3 * nv, 3 * nv)).toarray()
    G = Pd.T @ A.toarray() @ Pd + np.diag(fixed1.astype(float))
    assert np.abs(A1 - G).max() <= 1e-12 * np.abs(G).max()

    def M(r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        probe.amg_probe_vcycle(r.ctypes.data, z.ctypes.data)
        return z

    rng = np.random.default_rng(2)
    n = A.shape[0]
    u, v = rng.standard_normal(n) * ~fixed, rng.standard_normal(n) * ~fixed
    assert abs(u @ M(v) - v @ M(u)) <= 1e-11 * abs(u @ M(u))
    assert all((w @ M(w)) > 0.0 for w in (rng.standard_normal(n) * ~fixed for _ in range(3)))
    b = rng.standard_normal(n) * ~fixed
    D = np.zeros((nn, 3, 3))
    for i in range(nn):
        for k in range(Ab.indptr[i], Ab.indptr[i + 1]):
            if Ab.indices[k] == i:
                D[i] = Ab.data[k]
    Dinv = np.linalg.inv(D)
    xa, ita = _pcg(A, M, b)
    xj, itj = _pcg(A, lambda r: np.einsum("nij,nj->ni", Dinv, r.reshape(-1, 3)).ravel(), b)
    assert np.linalg.norm(A @ xa - b) <= 1e-8 * np.linalg.norm(b)
    assert ita * 10 <= itj, (ita, itj)


def test_p2_hierarchy_from_fp32_values(probe, oracle):
    """The device path's set-up reads the operator's FP32 copy: the hierarchy built from FP32 values
    equals, bit for bit, the one built from the same values widened to FP64 (every level's
    operator, transfers, inverse diagonal blocks and spectrum bound, and the coarsest inverse)."""
    m, A, fixed = _p2_operator(oracle, 4)
    nn, nv = m.num_nodes, m.num_vertices
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    par = [None] * nn
    for c in range(m.num_cells):
        cl, cn = m.closure[c], m.cone[c]
        for b in range(4):
            par[cl[b]] = ([int(cn[b])], [1.0])
        for e, (a, b) in enumerate(edges):
            par[cl[4 + e]] = (sorted([int(cn[a]), int(cn[b])]), [0.5, 0.5])
    par_ptr = np.zeros(nn + 1, np.int32)
    for i in range(nn):
        par_ptr[i + 1] = par_ptr[i] + len(par[i][0])
    pidx = np.concatenate([np.array(p[0], np.int32) for p in par])
    pw = np.concatenate([np.array(p[1], float) for p in par])
    Ab = sp.bsr_matrix(A, blocksize=(3, 3))
    Ab.sort_indices()
    ptr, col = Ab.indptr.astype(np.int32), Ab.indices.astype(np.int32)
    val32 = np.ascontiguousarray(Ab.data, np.float32)
    val64 = val32.astype(np.float64)
    fx = fixed.astype(np.uint8)
    vxyz = np.ascontiguousarray(m.coords, np.float64)

    def snapshot(nl):
        out = []
        for l in range(nl):
            info, lmax = np.zeros(3, np.int32), ctypes.c_double()
            probe.amg_probe_level(l, info.ctypes.data, ctypes.byref(lmax))
            out.append(("info", l, info.tolist(), lmax.value))
            for which in range(10):
                cnt = probe.amg_probe_array(l, which, None)
                dt = np.float64 if which in (2, 5, 8, 9) else np.int32
                a = np.zeros(max(cnt, 0), dt)
                if cnt > 0:
                    probe.amg_probe_array(l, which, a.ctypes.data)
                out.append((which, l, a))
        cnt = probe.amg_probe_coarse(None)
        cinv = np.zeros(cnt, np.float64)
        probe.amg_probe_coarse(cinv.ctypes.data)
        out.append(("coarse", cinv))
        return out

    probe.amg_probe_set_theta(0.05)
    try:
        args = (par_ptr.ctypes.data, pidx.ctypes.data, pw.ctypes.data, vxyz.ctypes.data, 300, 8)
        nl64 = probe.amg_probe_build_p2(nn, ptr.ctypes.data, col.ctypes.data, val64.ctypes.data, fx.ctypes.data,
                                        nv, *args)
        assert nl64 >= 2, probe.amg_probe_error()
        ref = snapshot(nl64)
        nl32 = probe.amg_probe_build_p2f(nn, ptr.ctypes.data, col.ctypes.data, val32.ctypes.data, fx.ctypes.data,
                                         nv, *args)
        assert nl32 == nl64, probe.amg_probe_error()
        got = snapshot(nl32)
    finally:
        probe.amg_probe_set_theta(0.0)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        if g[0] == "info":
            assert g == r
        elif g[0] == "coarse":
            assert np.array_equal(g[1], r[1])
        else:
            assert g[0] == r[0] and g[1] == r[1] and np.array_equal(g[2], r[2]), (g[0], g[1])

