"""The Gmsh reader on the reference's MMS fixture meshes (tests/mmstests/linearelasticity/
nofaults-3d/data/{hex,tet}.msh, committed here as tests/golden/mms_{hex,tet}.msh) and the MMS
known answer on them through the oracle: the UniformStrain3D displacement (UniformStrain3D.cc:
63-104, with its strain_yz = 0 quirk that makes u_y and u_z depend on x only) is linear, so the
residual of every interior node vanishes (TestLinearElasticity.cc:150 checks it at 1e-9).  No GPU.
"""
import math
import os

import numpy as np
import pytest

from pylith_amd import meshgen as mg
from pylith_amd import meshio_gmsh

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
F1 = 1


def uniform_strain_disp(x):
    """UniformStrain3D.cc:88-104 as written (strain_yz = PETSC_SUCCESS = 0, products in disp_y/z)."""
    exx, eyy, ezz, exy, eyz, exz = 0.1, 0.25, -0.05, 0.3, 0.0, -0.08
    ux = exx * x[:, 0] + exy * x[:, 1] + exz * x[:, 2]
    uy = exy * x[:, 0] + eyy * x[:, 1] * eyz * x[:, 2]
    uz = exz * x[:, 0] + eyz * x[:, 1] * ezz * x[:, 2]
    return np.stack([ux, uy, uz], 1).ravel()


def mms_aux(m):
    # UniformStrain3D.cc:30-60: density 2500, Vs 3000, Vp = sqrt(3) Vs
    return mg.elastic_aux(m.num_cells, density=2500.0, vs=3000.0, vp=math.sqrt(3.0) * 3000.0)


def _volume(m):
    if m.cell_type == mg.TET_P2:
        x = m.coords[m.cone]
        return np.einsum("ij,ij->i", x[:, 1] - x[:, 0], np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 0])).sum() / 6
    # trilinear hexes: 2x2x2 Gauss integrates det J exactly
    x = m.coords[m.closure]  # [nc, 8, 3], PyLith order 0-3 bottom cyclic, 4-7 top
    ref = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    g = 1.0 / math.sqrt(3.0)
    vol = 0.0
    for q in np.array(np.meshgrid([-g, g], [-g, g], [-g, g])).reshape(3, -1).T:
        dN = np.empty((8, 3))
        for a in range(8):
            f = 1 + ref[a] * q
            dN[a] = ref[a] * np.array([f[1] * f[2], f[0] * f[2], f[0] * f[1]]) / 8
        J = np.einsum("cai,aj->cij", x, dN)
        vol += np.linalg.det(J).sum()
    return vol


@pytest.mark.parametrize("name", ["mms_hex.msh", "mms_tet.msh"])
def test_reader_on_reference_meshes(name):
    m, groups = meshio_gmsh.read(os.path.join(GOLDEN, name))
    lo, hi = m.coords.min(0), m.coords.max(0)
    assert np.allclose(lo, [-4000, -4000, -8000]) and np.allclose(hi, [4000, 4000, 0])
    assert np.all(m.labels == 24)  # physical volume "material-id:24"
    assert abs(_volume(m) - 8000.0 ** 3) <= 1e-9 * 8000.0 ** 3
    # the "boundary" physical surface is exactly the nodes on the box faces
    x = m.node_coords
    on = np.any(np.isclose(x, lo, atol=1e-6) | np.isclose(x, hi, atol=1e-6), axis=1)
    assert np.array_equal(np.sort(groups["boundary"]), np.nonzero(on)[0])
    if m.cell_type == mg.HEX_Q1:
        assert (m.num_cells, m.num_nodes) == (125, 216)
    else:
        assert m.closure.shape[1] == 10 and m.num_vertices == 111
        # edge nodes sit at their edge midpoints
        for k, (a, b) in enumerate(((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))):
            mid = 0.5 * (x[m.closure[:, a]] + x[m.closure[:, b]])
            assert np.allclose(x[m.closure[:, 4 + k]], mid)
        # every edge node is shared consistently: 6 per tet, Euler count for a ball
        nedges = m.num_nodes - m.num_vertices
        faces = np.unique(np.sort(m.cone[:, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]].reshape(-1, 3), 1), axis=0)
        assert m.num_vertices - nedges + len(faces) - m.num_cells == 1


@pytest.mark.parametrize("name", ["mms_hex.msh", "mms_tet.msh"])
def test_mms_uniform_strain_interior_residual_vanishes(oracle, name):
    m, groups = meshio_gmsh.read(os.path.join(GOLDEN, name))
    aux, sizes = mms_aux(m)
    P = oracle.Problem(m, aux, sizes)
    r = P.residual(uniform_strain_disp(m.node_coords), F1).reshape(-1, 3)
    interior = np.setdiff1d(np.arange(m.num_nodes), groups["boundary"])
    scale = np.abs(r).max()
    assert scale > 0.0  # boundary rows carry the tractions
    assert np.abs(r[interior]).max() <= 1e-12 * scale
    # a non-affine field (quadratic in x) does not pass: the check can fail
    x = m.node_coords
    bad = uniform_strain_disp(x) + np.stack([1e-5 * x[:, 0] ** 2, 0 * x[:, 0], 0 * x[:, 0]], 1).ravel()
    rb = P.residual(bad, F1).reshape(-1, 3)
    assert np.abs(rb[interior]).max() > 1e-6 * scale
